"""Small-value eq-poly sumcheck (drop-in for reference
zorch/sumcheck/eq/small_value.py and accumulators.py).

Rounds 1..l0 read precomputed accumulators A_i(v, u): one device pass over
the base-field factors forms, for every beta in U_d^l0 (the first l0
variables extended to {inf, 0, 1, .., d-1}),
T[beta] = sum_x eq(w[l0:], x) prod_k p_k(beta, x); each A_i is a contraction of
T with eq(w[i:l0], .). A round is then R_i . A_i, the round's Lagrange tensor
grows R, and the challenge is one transcript word. The factors are refolded
once over the l0 bound variables and the rest of the proof (the transition
round first) is eq-poly rounds from the bound eq mass, all on the compressed
U-hat_d domain, so the messages equal ``EqPolyRound``'s (``zs_prove_small_value``).

The accumulator routing reads v and y most-significant digit first, the
order of the R tensor and of the e_out rows they meet. The reference's
``_idx4`` writes them least-significant first; the two agree up to l0 = 2
(at most one digit each), and from l0 = 3 on only the order used here gives
the eq-poly messages the engine is specified to send.
"""

from __future__ import annotations

import ctypes

import numpy as np

from zorch_amd import _lib
from zorch_amd.challenge import ChallengePolicy
from zorch_amd.fields import FieldArray, as_field
from zorch_amd.round import RunningClaim
from zorch_amd.runtime import default_context, log2_strict
from zorch_amd.sumcheck._device import make_job
from zorch_amd.sumcheck.prover import FoldingClaim


def prove_eq_poly_small_value(p_initial, w: FieldArray, l_0: int, claim: FieldArray, transcript,
                              *, challenges: ChallengePolicy):
    """Prove the eq-weighted product sumcheck with l_0 small-value rounds
    (reference small_value.py:148-192). Returns (carry, transcript, messages):
    carry.state = (final folded factors (d, 1), eq mass), carry.claim the
    reduced claim; every message on U-hat_d. One device call."""
    d = p_initial.shape[0]
    l = log2_strict(p_initial.shape[1])
    if not 1 <= l_0 <= l - l // 2:
        raise ValueError(f"l_0 must be in [1, {l - l // 2}] (the small-value rounds fit before "
                         f"the out-half); got l_0={l_0} for l={l}")
    chal = as_field(challenges.dtype)
    job = make_job(p_initial, claim, chal, _lib.ZS_MODE_EQ, d, w)
    st = transcript.state.copy()
    _lib.check(_lib.lib().zs_prove_small_value(default_context().handle, ctypes.byref(job.args),
                                               l_0, ctypes.byref(st), ctypes.byref(job.proof)))
    mlen = d  # U-hat_d samples
    msgs = FieldArray(np.ascontiguousarray(
        job.msgs.reshape(-1, chal.words)[: l * mlen].reshape(l, mlen, chal.words)), chal)
    final = FieldArray(job.final[:d], chal).reshape(d, 1)
    mass = FieldArray(job.final[d], chal)
    carry = FoldingClaim((final, mass),
                         RunningClaim(FieldArray(job.value, chal), FieldArray(job.point, chal), l))
    return carry, type(transcript)(st, transcript.dtype), [msgs[i] for i in range(l)]


__all__ = ["prove_eq_poly_small_value"]
