"""CPU oracle for the dense sumcheck path — TEST INFRASTRUCTURE ONLY.

This package restates, in plain Python integers (and one C file for large
sizes), the reference's dense sumcheck algorithm and its SHA-256 field
transcript, so that the HIP product can be checked bit for bit. Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import, call, link or execute anything in here, and only as
the checker — never as a code path of the product.

Modules and what pins them (each function cites the reference file:line it
restates; "pinned by" names the independent check in ``tests/``):

* ``fields``      — BabyBear (Montgomery storage), BabyBear^4 = F_p[X]/(X^4-11),
                    Goldilocks. Pinned by closed forms, field axioms,
                    Fermat/irreducibility checks and the Montgomery codec.
* ``sha256``      — pure-Python SHA-256 compression for the streaming midstate.
                    Pinned by ``hashlib`` digests on every length 0..300.
* ``transcript``  — ``Sha256FieldTranscript`` framing. Pinned by ``hashlib``
                    and the reference test's independent byte restatement
                    (zorch/testing/byte_transcript_test.py:40-89).
* ``sumcheck``    — StandardRound / EqPolyRound provers, SumcheckRound
                    verifier. Pinned by hypercube brute force, the
                    s(0)+s(1)=claim identity, the final-claim =
                    prod eval_mle(T_k, point) identity, the eq-poly ==
                    product-over-[A, B, eq] anchor and the reference KATs.
* ``sumcheck_ref.c`` — the same algorithm in C (any n), cross-checked against
                    the Python oracle on the committed fixtures; used for
                    parity at large n and as the timed CPU baseline.

Parity unpinned (no import of the reference stack is possible here; see the
design notes): the raw-vs-reduced storage FRX keeps for a squeezed word >= p
inside ``EvaluationClaim.point`` (values are pinned, raw words are stored),
and the BabyBear^4 modulus / limb order (X^4 - 11, limbs c0..c3).
"""
