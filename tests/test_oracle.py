"""Pins of the CPU oracle (no GPU).

Every oracle function is checked against something other than a retyping of
its own formula: hashlib digests, the reference test's independent byte
restatement of the transcript framing, field-theoretic facts, hypercube brute
force, the sumcheck identities, the eq-poly == product-over-[A, B, eq] anchor
(reference zorch/sumcheck/eq/testing/eq_poly_test.py:53-72) and the
reference's own small-integer KATs restated on BabyBear.
"""

from __future__ import annotations

import hashlib
import json
import random
from pathlib import Path

import numpy as np
import pytest

from oracle import blake3 as OB3
from oracle import fields as OF
from oracle import sha256 as OSHA
from oracle import sumcheck as OS
from oracle import transcript as OT
from tests import blake3_indep as B3I
from tests.golden.recipes import RECIPES, build, fixture_state
from tests.oracle_glue import run_c, run_py

BB, BB4, GL = OF.BABYBEAR, OF.BABYBEAR4, OF.GOLDILOCKS
GOLDEN = Path(__file__).resolve().parent / "golden"


# ------------------------------------------------------------------ SHA-256
def test_sha256_constants_are_fips():
    assert OSHA.K[0] == 0x428A2F98 and OSHA.K[63] == 0xC67178F2
    assert len(OSHA.K) == 64


@pytest.mark.parametrize("length", list(range(0, 200)) + [255, 256, 257, 1000])
def test_sha256_midstate_matches_hashlib(length):
    data = bytes((i * 37 + length) & 0xFF for i in range(length))
    h, pend, tot = OSHA.midstate(data)
    assert tot == length and len(pend) == length % 64
    assert OSHA.finalize(h, pend, tot) == hashlib.sha256(data).digest()


# ------------------------------------------------------------------- BLAKE3
# Published vectors, as the reference's own test quotes them
# (zorch/testkit/testing/byte_hash_test.py:24-29: input_len 3, input[i] = i % 251,
# 64 bytes of extendable output), and the empty input (BLAKE3 test vectors,
# input_len 0).
_B3_LEN3_XOF64 = bytes.fromhex(
    "e1be4d7a8ab5560aa4199eea339849ba8e293d55ca0a81006726d184519e647f"
    "5b49b82f805a538c68915c1ae8035c900fd1d4b13902920fd05e1450822f36de")
_B3_EMPTY = bytes.fromhex("af1349b9f5f9a1a6a0404dea36dcc9499bcb25c9adc112b7cc9a93cae41f3262")


def test_blake3_published_vectors():
    assert OB3.blake3(b"") == _B3_EMPTY
    for width in (1, 16, 32, 64):
        assert OB3.blake3(bytes(range(3)), width) == _B3_LEN3_XOF64[:width]
        assert OB3.blake3_tree(bytes(range(3)), width) == _B3_LEN3_XOF64[:width]


@pytest.mark.parametrize("length", [0, 1, 63, 64, 65, 127, 128, 129, 1023, 1024, 1025, 2047,
                                    2048, 2049, 3072, 3073, 4096, 4097, 5121, 8193, 16385])
def test_blake3_stream_equals_tree(length):
    # Two restatements of the specification: incremental (chunk state + merged
    # subtree stack) and one-shot recursion over the chunk tree; the stream is
    # also fed in random pieces. Extendable output past one 64-byte block.
    data = bytes(i % 251 for i in range(length))
    want = OB3.blake3_tree(data, 150)
    assert OB3.blake3(data, 150) == want
    rng = random.Random(length)
    s, pos = OB3.Blake3Stream(), 0
    while pos < length:
        k = rng.randrange(0, 700)
        s.absorb(data[pos:pos + k])
        pos += k
    assert s.finalize(150) == want
    assert OB3.blake3(data, 32) == want[:32]  # the hash is a prefix of the XOF


def test_independent_blake3_pinned_to_published_vectors():
    assert B3I.blake3_xof(b"") == _B3_EMPTY
    assert B3I.blake3_xof(bytes(range(3)), 64) == _B3_LEN3_XOF64


@pytest.mark.parametrize("length", [0, 1, 3, 63, 64, 65, 128, 1023, 1024, 1025, 2048, 2049,
                                    3072, 3073, 4095, 5120, 6145, 7169, 8192, 8193,
                                    11264, 16385, 31744, 33793])
def test_blake3_oracle_equals_independent_restatement(length):
    # Multi-block chunks, parent nodes, trees of 1..33 chunks (odd carries at
    # several levels) and extendable output over four output blocks, against
    # a restatement that shares no code with the oracle.
    data = bytes((i * 7 + 3) % 251 for i in range(length))
    assert OB3.blake3(data, 250) == B3I.blake3_xof(data, 250)


def test_independent_restatement_catches_a_planted_parent_slip(monkeypatch):
    # A parent node compressed with the chunk counter instead of 0 (a plausible
    # slip) changes every input of more than one chunk, so the comparison
    # above would see it.
    data = bytes(i % 251 for i in range(2049))
    good = OB3.blake3(data, 32)
    real = OB3._parent
    monkeypatch.setattr(OB3, "_parent",
                        lambda l, r: OB3.Output(OB3.IV, real(l, r).block, 1, OB3.BLOCK_LEN,
                                                OB3.PARENT))
    assert OB3.blake3(data, 32) != good
    assert OB3.blake3(data[:1024], 32) == B3I.blake3_xof(data[:1024], 32)


def test_blake3_transcript_framing():
    # Blake3FieldTranscript's wire (zorch/blake3_field_transcript.py:293-330,
    # 387-395): the same framed stream as the SHA-256 row, squeezed by XOF read.
    t = OT.OracleBlake3Transcript.new(b"dom", 4)
    t, out = t.observe_and_sample(b"\x01\x00\x00\x00\x02\x00\x00\x00", 3)
    stream = (bytes([1]) + _le8(3) + b"dom" + bytes([3, 2]) + _le8(2)
              + b"\x01\x00\x00\x00\x02\x00\x00\x00" + bytes([4, 2]) + _le8(3))
    assert out == OB3.blake3_tree(stream, 12)
    assert t.stream == stream + out
    t2, sc = t.sample_scalar()
    assert sc == OB3.blake3_tree(t.stream + bytes([4, 1]), 4)


# ----------------------------------------------------- transcript framing
def _le8(n):
    return int(n).to_bytes(8, "little")


def _ref_squeeze(buf: bytes, n: int) -> bytes:
    out, ctr = b"", 0
    while len(out) < n:
        out += hashlib.sha256(buf + _le8(ctr)).digest()
        ctr += 1
    return out[:n]


def test_transcript_matches_independent_byte_restatement():
    # The reference's independent hashlib framing (byte_transcript_test.py:40-89).
    buf = bytes([1]) + _le8(3) + b"dom"
    t = OT.OracleTranscript.new(b"dom", 4)
    assert t.stream == buf
    vals = np.array([1, 2, 3, 4, 0xDEADBEEF, 5], dtype="<u4").tobytes()
    t = t.observe(vals)
    buf += bytes([3, 2]) + _le8(6) + vals
    t, out = t.sample(3)
    buf += bytes([4, 2]) + _le8(3)
    sq = _ref_squeeze(buf, 12)
    buf += sq
    assert out == sq and t.stream == buf
    # A squeeze longer than one digest uses counters 0, 1, ...
    t, out = t.sample(20)
    buf += bytes([4, 2]) + _le8(20)
    sq = _ref_squeeze(buf, 80)
    assert out == sq
    # Scalar framing differs from slice framing.
    a = OT.OracleTranscript.new(b"x", 4).observe_scalar(b"\x01\x02\x03\x04").sample_scalar()[1]
    b = OT.OracleTranscript.new(b"x", 4).observe(b"\x01\x02\x03\x04").sample(1)[1]
    assert a != b


def test_transcript_state_is_streaming_midstate():
    t = OT.OracleTranscript.new(b"L" * 40, 4).observe(b"\x07" * 64)
    h, pend, tot = t.state()
    assert tot == len(t.stream) and pend == t.stream[tot - tot % 64:]
    assert OSHA.finalize(h, pend, tot) == hashlib.sha256(t.stream).digest()


# ------------------------------------------------------------------ fields
def test_babybear_montgomery_codec():
    assert BB.p == 0x78000001
    assert BB.to_words(1) == [(1 << 32) % BB.p]
    for a in [0, 1, 2, BB.p - 1, 123456789]:
        assert BB.from_words(BB.to_words(a)) == a
    # A non-reduced word w decodes to w * 2^-32 mod p.
    w = 0xFFFFFFFF
    assert BB.from_words([w]) == (w * pow(1 << 32, -1, BB.p)) % BB.p


def test_quartic_extension_is_a_field():
    p = BB.p
    assert p % 4 == 1
    # X^4 - 11 is irreducible over F_p (p = 1 mod 4) iff 11 is a non-square.
    assert pow(11, (p - 1) // 2, p) == p - 1
    x = (0, 1, 0, 0)
    assert BB4.pow(x, 4) == (11, 0, 0, 0)
    rng = random.Random(5)
    for _ in range(20):
        a = tuple(rng.randrange(p) for _ in range(4))
        b = tuple(rng.randrange(p) for _ in range(4))
        c = tuple(rng.randrange(p) for _ in range(4))
        assert BB4.mul(a, b) == BB4.mul(b, a)
        assert BB4.mul(BB4.mul(a, b), c) == BB4.mul(a, BB4.mul(b, c))
        assert BB4.mul(a, BB4.add(b, c)) == BB4.add(BB4.mul(a, b), BB4.mul(a, c))
        assert BB4.mul(a, BB4.inv(a)) == (1, 0, 0, 0)
    a = (3, 1, 4, 1)
    assert BB4.pow(a, p ** 4) == a  # Frobenius^4 is the identity


def test_goldilocks_closed_forms():
    p = GL.p
    assert p == 2 ** 64 - 2 ** 32 + 1
    assert GL.mul(1 << 32, 1 << 32) == (1 << 32) - 1  # 2^64 = 2^32 - 1
    assert GL.mul(1 << 48, 1 << 48) == p - 1           # 2^96 = -1
    assert GL.from_words([p + 5]) == 5


# ----------------------------------------------------- reference KATs
def test_suffix_eq_tables_kat():
    # zorch/sumcheck/eq/testing/eq_poly_test.py:31-37: w=[0,1].
    t = OS.suffix_eq_tables(BB, [0, 1])
    assert t[0] == [0, 1] and t[1] == [0, 1, 0, 0]


def test_msb_fold_kat():
    # eq_poly_test.py:39-51: r=3 fold of 1..32 gives 49..64.
    P = list(range(1, 33))
    got = OS._fold(BB, P, 3)
    assert got == list(range(49, 65))


def test_round_poly_degree1_and_2_kat():
    # prover_test.py:27-46 restated: s(0)=sum lo, s(1)=sum hi; degree-2 by definition.
    rng = random.Random(11)
    f = [rng.randrange(1 << 30) for _ in range(8)]
    pf = OS.prove_product(BB, BB, [f], sum(f) % BB.p, OT.OracleTranscript.new(b"", 4), 1,
                          rounds=1)
    assert pf.msgs[0] == [sum(f[:4]) % BB.p, sum(f[4:]) % BB.p]
    a = [rng.randrange(1 << 30) for _ in range(8)]
    b = [rng.randrange(1 << 30) for _ in range(8)]
    pf = OS.prove_product(BB, BB, [a, b], 0, OT.OracleTranscript.new(b"", 4), 2, rounds=1)
    for u in range(3):
        fa = [(a[j] + u * (a[j + 4] - a[j])) for j in range(4)]
        fb = [(b[j] + u * (b[j + 4] - b[j])) for j in range(4)]
        assert pf.msgs[0][u] == sum(x * y for x, y in zip(fa, fb)) % BB.p


# ------------------------------------------------- brute force / identities
def _rand_el(F, rng):
    return tuple(rng.randrange(F.p) for _ in range(4)) if F.degree == 4 else rng.randrange(F.p)


def _mle_univariate(F, tabs, prefix, t, x_rest, n):
    """prod_k T_k(prefix, t, x_rest) by multilinear extension on the hypercube."""
    out = F.one()
    for tab in tabs:
        pt = list(prefix) + [F.from_int(t)] + [F.from_int(b) for b in x_rest]
        out = F.mul(out, OS.eval_mle(F, tab, pt))
    return out


@pytest.mark.parametrize("F,n,k", [(BB, 3, 2), (BB4, 3, 2), (GL, 3, 2), (BB, 2, 3), (BB4, 2, 1)])
def test_product_messages_brute_force(F, n, k):
    rng = random.Random(n * 10 + k)
    tabs = [[_rand_el(F, rng) for _ in range(1 << n)] for _ in range(k)]
    claim = F.zero()
    for x in range(1 << n):
        pr = F.one()
        for t in tabs:
            pr = F.mul(pr, t[x])
        claim = F.add(claim, pr)
    T = F if F.degree == 1 else BB
    pf = OS.prove_product(F, T, tabs, claim, OT.OracleTranscript.new(b"bf", T.item_bytes), k)
    for i in range(n):
        prefix = pf.point[:i]
        rest = n - i - 1
        for t in range(k + 1):
            s = F.zero()
            for xr in range(1 << rest):
                bits = [(xr >> (rest - 1 - j)) & 1 for j in range(rest)]
                s = F.add(s, _mle_univariate(F, tabs, prefix, t, bits, n))
            assert pf.msgs[i][t] == s
    # Identities: s(0)+s(1) chains the claim; final value = prod_k MLE_k(point).
    c = claim
    for i, msg in enumerate(pf.msgs):
        assert F.add(msg[0], msg[1]) == c
        c = OS.lagrange_eval(F, msg, pf.point[i])
    want = F.one()
    for t in tabs:
        want = F.mul(want, OS.eval_mle(F, t, pf.point))
    assert c == pf.value == want


def test_eq_messages_match_product_over_eq_table():
    # Anchor: eq-poly messages equal a product sumcheck over [A, B, eq(w, .)]
    # driven by the same challenges (eq_poly_test.py:53-72).
    F, n = BB4, 4
    rng = random.Random(3)
    A = [rng.randrange(BB.p) for _ in range(1 << n)]
    B = [rng.randrange(BB.p) for _ in range(1 << n)]
    w = [_rand_el(F, rng) for _ in range(n)]
    eqt = OS.eq_table(F, w)
    claim = F.zero()
    for x in range(1 << n):
        claim = F.add(claim, F.mul(eqt[x], F.mul(A[x], B[x])))
    pf = OS.prove_eq(F, BB, [A, B], w, claim, OT.OracleTranscript.new(b"eq", 4), 2)
    state = [[F.embed(v) for v in t] for t in (A, B, eqt)]
    for i in range(n):
        h = len(state[0]) // 2
        for t in range(4):
            s = F.zero()
            for j in range(h):
                pr = F.one()
                for tab in state:
                    pr = F.mul(pr, F.add(tab[j], F.mul(F.from_int(t), F.sub(tab[j + h], tab[j]))))
                s = F.add(s, pr)
            assert pf.msgs[i][t] == s
        state = [OS._fold(F, tab, pf.point[i]) for tab in state]
    # Final: value = eq(w, r) * A(r) * B(r).
    e = F.one()
    for wi, ri in zip(w, pf.point):
        e = F.mul(e, OS.eq_factor(F, ri, wi))
    assert pf.value == F.mul(e, F.mul(OS.eval_mle(F, A, pf.point), OS.eval_mle(F, B, pf.point)))


def test_reduce_domain_equals_lagrange():
    rng = random.Random(9)
    for F in (BB, BB4, GL):
        for npts in (2, 3, 4, 5):
            msg = [_rand_el(F, rng) for _ in range(npts)]
            r = _rand_el(F, rng)
            assert OS.eval_coeffs(F, OS.natural_to_coeffs(F, msg), r) == \
                OS.lagrange_eval(F, msg, r)


def test_verifier_accepts_and_rejects():
    inst = build(RECIPES["product_bb_n4"])
    res, pf = run_py(inst)
    F = BB
    claim = BB.from_words([int(inst.claim.words.reshape(-1)[0])])
    t0 = OT.OracleTranscript.new(inst.label, 4)
    point, value, _, ok = OS.verify(F, F, claim, pf.msgs, t0, 2)
    assert ok and point == pf.point and value == pf.value
    _, _, _, ok = OS.verify(F, F, F.add(claim, 1), pf.msgs, t0, 2)
    assert not ok
    bad = [list(m) for m in pf.msgs]
    bad[2][0] = F.add(bad[2][0], 1)
    _, _, _, ok = OS.verify(F, F, claim, bad, t0, 2)
    assert not ok
    with pytest.raises(ValueError):
        OS.verify(F, F, claim, pf.msgs, t0, 3)
    with pytest.raises(ValueError):
        OS.verify(F, F, claim, [], t0, 2)


# ------------------------------------------------ fixtures and C oracle
@pytest.mark.parametrize("name", sorted(RECIPES))
def test_fixture_reproduces_and_c_oracle_agrees(name):
    fx = json.loads((GOLDEN / f"{name}.json").read_text())
    inst = build(fx["recipe"])
    py, _ = run_py(inst)
    results = [py]
    if inst.transcript_kind == "sha256":  # the C oracle restates SHA-256 only
        results.append(run_c(inst, threads=2))
    for res in results:
        assert res.msgs.tolist() == fx["msgs"]
        assert res.point.tolist() == fx["point"]
        assert res.value.tolist() == fx["value"]
        assert res.final.tolist() == fx["final"]
        assert res.state == fixture_state(fx)
