"""Round-message wire forms: evaluation domains (the compressed Û domain,
explicit node sets, the two-point compressed domains), ascending coefficients
and the compressed [c_0, c_2] product wire.

CPU tests pin the oracle's restatement (oracle/sumcheck.py) with the
reference's own small-integer KATs restated on BabyBear
(zorch/sumcheck/testing/domain_test.py, eq_poly_test.py, verifier_test.py,
gruen_test.py), the eq-poly == product-over-[P, eq] anchor on Û, MLE brute
force of every sample and closed-form final claims. GPU tests compare the
device prover with that oracle bit for bit, whole-stage and round by round,
batched and served, and check the verifiers accept.
"""

from __future__ import annotations

import math
import random

import numpy as np
import pytest

from oracle import fields as OF
from oracle import sumcheck as OS
from zorch_amd import _lib
from zorch_amd.fields import array, babybear_mont, babybearx4_mont, goldilocks, inv
from zorch_amd.sumcheck.domain import EvalDomain, compressed_domain, natural_domain, uhat_domain

BB, BB4 = OF.BABYBEAR, OF.BABYBEAR4


def _bb(vals):
    return [BB.from_int(v) for v in vals]


# ------------------------------------------------------------ oracle pins
def test_to_coeffs_reference_kats():
    """domain_test.py:52-136 restated on BabyBear, through the oracle and
    through the package's host EvalDomain.to_coeffs."""
    half = BB.inv(BB.from_int(2))
    cases = [
        (_bb([0, 1, 2]), None, _bb([1, 6, 17]), _bb([1, 2, 3])),
        (None, 0, _bb([3, 1, 6]), _bb([1, 2, 3])),
        ([BB.zero(), half], None, [3, BB.add(3, BB.mul(5, half))], _bb([3, 5])),
        (_bb([0, 2, 4]), None, _bb([1, 17, 57]), _bb([1, 2, 3])),
        (_bb([0]), -1, _bb([3, 5]), _bb([3, 5])),
    ]
    for nodes, inf, vals, want in cases:
        assert OS.to_coeffs(BB, nodes, inf, vals) == want
        dom = EvalDomain(array(nodes, babybear_mont) if nodes is not None else None, inf)
        got = dom.to_coeffs(array(vals, babybear_mont))
        assert got.to_ints() == want


def test_interp_routes_agree():
    """Gauss-Jordan on the Vandermonde system and the Lagrange expansion of
    the natural map agree on random values (two independent routes)."""
    rng = random.Random(5)
    for n in range(1, 7):
        ys = [rng.randrange(BB.p) for _ in range(n)]
        assert OS.interp(BB, _bb(range(n)), ys) == OS.natural_to_coeffs(BB, ys)
        ys4 = [tuple(rng.randrange(BB.p) for _ in range(4)) for _ in range(n)]
        assert OS.interp(BB4, [BB4.from_int(i) for i in range(n)], ys4) == \
            OS.natural_to_coeffs(BB4, ys4)


def test_product_sample_kats():
    """domain_test.py:40-48 (Û_2 of two factors), 92-106 ({0,2,4}),
    137-151 (compressed domains), 183-188 (one factor's coefficients)."""
    a, b = _bb([1, 2, 3, 4]), _bb([5, 6, 7, 8])
    assert OS._product_samples(BB, [a, b], [None, BB.zero()]) == _bb([8, 17])
    assert OS._product_samples(BB, [_bb([1, 3]), _bb([2, 5])], _bb([0, 2, 4])) == \
        _bb([2, 40, 126])
    assert OS._product_samples(BB, [a, b], [BB.zero(), None]) == _bb([17, 8])
    assert OS._product_samples(BB, [a, b], [BB.one(), None]) == _bb([53, 8])
    msg = OS._wire_message(BB, ("coeffs",),
                           lambda pts: OS._product_samples(BB, [_bb(range(1, 9))], pts), 1)
    assert msg == _bb([10, 16])


def _mle_sample(F, tables, x, weights=None):
    """s(x) = sum_y prod_k MLE_k(x, y) (x None: the leading coefficient, from
    the natural samples by finite differences) — hypercube brute force through
    eval_mle, independent of the per-pair sampling."""
    n = OS.log2_strict(len(tables[0]))
    if x is None:
        deg = len(tables) + (1 if weights is not None else 0)
        vals = [_mle_sample(F, tables, F.from_int(t), weights) for t in range(deg + 1)]
        lead = F.zero()
        for i, v in enumerate(vals):
            c = math.comb(deg, i) * (-1 if (deg - i) & 1 else 1)
            lead = F.add(lead, F.mul(F.from_int(c), v))
        return F.mul(lead, F.inv(F.from_int(math.factorial(deg))))
    acc = F.zero()
    for y in range(1 << (n - 1)):
        bits = [F.from_int((y >> (n - 2 - i)) & 1) for i in range(n - 1)]
        prod = F.one()
        for t in tables:
            prod = F.mul(prod, OS.eval_mle(F, t, [x] + bits))
        if weights is not None:
            prod = F.mul(prod, OS.eval_mle(F, weights, [x] + bits))
        acc = F.add(acc, prod)
    return acc


def _rand_tables(F, k, n, seed):
    rng = random.Random(seed)
    return [[rng.randrange(F.p if F.degree == 1 else BB.p) for _ in range(1 << n)]
            for _ in range(k)]


@pytest.mark.parametrize("wire", [
    ("domain", _bb([0, 2, 4]), None), ("domain", _bb([1]), -1), ("domain", _bb([0, 2]), 0),
    ("compressed",), ("coeffs",)])
def test_product_wire_messages_brute_force(wire):
    """Every round's message against MLE brute force at the proof's own
    challenges; coefficient messages through their values."""
    from oracle.transcript import OracleTranscript
    tabs = _rand_tables(BB, 2, 4, 7)
    claim = BB.zero()
    for j in range(16):
        claim = BB.add(claim, BB.mul(tabs[0][j], tabs[1][j]))
    pf = OS.prove_product_wire(BB, BB, tabs, claim, OracleTranscript.new(b"w", 4), 2, wire)
    cur = [list(t) for t in tabs]
    for msg, r in zip(pf.msgs, pf.point):
        if wire[0] == "domain":
            pts = OS._points(BB, wire[1], wire[2])
            assert msg == [_mle_sample(BB, cur, x) for x in pts]
        elif wire[0] == "compressed":
            assert msg == [_mle_sample(BB, cur, BB.zero()), _mle_sample(BB, cur, None)]
        else:
            for t in range(3):
                assert OS.eval_coeffs(BB, msg, BB.from_int(t)) == \
                    _mle_sample(BB, cur, BB.from_int(t))
        cur = [OS._fold(BB, t, r) for t in cur]
    if wire[0] in ("compressed", "coeffs"):
        # Honest claims reduce to s(r): the final claim is prod_k MLE_k(point).
        assert pf.value == BB.mul(OS.eval_mle(BB, tabs[0], pf.point),
                                  OS.eval_mle(BB, tabs[1], pf.point))


def test_uhat_eq_messages_match_product_over_eq():
    """eq_poly_test.py:53-72: EqPolyRound's Û_d messages equal a product
    sumcheck over [P_1, ..., P_d, eq(w, .)] on Û_{d+1} with its last sample
    dropped, at the same challenges."""
    from oracle.transcript import OracleTranscript
    d, n = 3, 4
    P = [[BB.from_int(q * 16 + j + 1) for j in range(16)] for q in range(d)]
    w = _bb([1, 0, 1, 0])
    eq = OS.eq_table(BB, w)
    claim = BB.zero()
    for j in range(16):
        claim = BB.add(claim, BB.mul(eq[j], BB.mul(P[0][j], BB.mul(P[1][j], P[2][j]))))
    uhat = ("domain", [BB.zero(), BB.from_int(2)], 0)
    pf = OS.prove_eq_wire(BB, BB, P, w, claim, OracleTranscript.new(b"u", 4), d, uhat)
    ref = [list(t) for t in P] + [eq]
    pts4 = OS._points(BB, _bb([0, 2, 3]), 0)
    for msg, r in zip(pf.msgs, pf.point):
        assert msg == OS._product_samples(BB, ref, pts4)[:-1]
        ref = [OS._fold(BB, t, r) for t in ref]


@pytest.mark.parametrize("kind", ["coeffs", "natural"])
def test_eq_wire_final_claim_closed_form(kind):
    """eq_poly_test.py:86-111: the coefficient form (and the natural values)
    reduce the claim to eq(w, r) prod_k P_k(r); the first coefficient
    message evaluates to the brute-force s_0."""
    from oracle.transcript import OracleTranscript
    tabs = _rand_tables(BB, 2, 4, 11)
    rng = random.Random(3)
    w = [tuple(rng.randrange(BB.p) for _ in range(4)) for _ in range(4)]
    F = BB4
    eq = OS.eq_table(F, w)
    claim = F.zero()
    for j in range(16):
        claim = F.add(claim, F.mul(eq[j], F.embed(BB.mul(tabs[0][j], tabs[1][j]))))
    wire = ("coeffs",) if kind == "coeffs" else ("values",)
    pf = OS.prove_eq_wire(F, BB, tabs, w, claim, OracleTranscript.new(b"e", 4), 2, wire)
    ptab = [[F.embed(x) for x in t] for t in tabs]
    expect = F.mul(OS.eval_mle(F, eq, pf.point),
                   F.mul(OS.eval_mle(F, ptab[0], pf.point), OS.eval_mle(F, ptab[1], pf.point)))
    assert pf.value == expect
    if kind == "coeffs":
        for t in range(4):
            assert OS.eval_coeffs(F, pf.msgs[0], F.from_int(t)) == \
                _mle_sample(F, ptab, F.from_int(t), eq)


def test_eq_mass_scales_messages():
    """A bound eq mass m scales the messages by m (the eq factor l_i is linear
    in it): checked on round 0, before the challenges diverge."""
    from oracle.transcript import OracleTranscript
    tabs = _rand_tables(BB, 2, 3, 13)
    w = [BB4.from_int(i + 2) for i in range(3)]
    m = (5, 6, 7, 8)
    one = OS.prove_eq_wire(BB4, BB, tabs, w, BB4.zero(), OracleTranscript.new(b"m", 4), 2,
                           ("values",))
    scaled = OS.prove_eq_wire(BB4, BB, tabs, w, BB4.zero(), OracleTranscript.new(b"m", 4), 2,
                              ("values",), eq_mass=m)
    assert scaled.msgs[0] == [BB4.mul(m, v) for v in one.msgs[0]]


def test_gruen_round_coeffs_recovers_polynomial():
    """gruen_test.py:29-40: round_coeffs on p's values recovers p, for a p
    with a zero at eq_root(z) (host helpers)."""
    from zorch_amd.sumcheck import gruen
    f = babybear_mont
    z = array(9, f)
    b = gruen.eq_root(z)
    # p(x) = (x - b)(3 + 2x + x^2): degree 3, extra point 1/2.
    def p(x):
        return (x - b) * (array(3, f) + array(2, f) * x + x * x)
    half = inv(array(2, f))
    got = gruen.round_coeffs(p(array(0, f)), p(array(0, f)) + p(array(1, f)), [half],
                             [p(half)], z)
    # Coefficients of (x - b)(3 + 2x + x^2) = -3b + (3 - 2b) x + (2 - b) x^2 + x^3.
    want = [array(0, f) - array(3, f) * b, array(3, f) - array(2, f) * b, array(2, f) - b,
            array(1, f)]
    assert [g.to_ints() for g in got] == [w_.to_ints() for w_ in want]
    with pytest.raises(ValueError):
        gruen.round_coeffs(array(1, f), array(1, f), [array(1, f)], [], z)


# ------------------------------------------------------------ host ABI
def _args(k=2, n=4, mode=_lib.ZS_MODE_PRODUCT, wire=None, rounds=None, tf=_lib.ZS_F_BABYBEAR,
          cf=_lib.ZS_F_BABYBEAR):
    claim = np.zeros(4, np.uint32)
    return _lib.ProveArgsC(table_field=tf, chal_field=cf, mode=mode, degree=k, num_tables=k,
                           rounds=rounds or n, tables=1, tables_on_device=1, refold_rounds=0,
                           eq_point=1 if mode == _lib.ZS_MODE_EQ else None,
                           claim=claim.ctypes.data, wire=wire, eq_mass=None), claim


def _width(args):
    from zorch_amd.runtime import message_width
    return message_width(args)


def test_message_widths_and_structural_errors():
    import ctypes
    a, keep = _args()
    assert _width(a) == 3
    a, keep = _args(mode=_lib.ZS_MODE_EQ, cf=_lib.ZS_F_BABYBEAR4)
    assert _width(a) == 4
    for kind, want in [(_lib.ZS_WIRE_COEFFS, 3), (_lib.ZS_WIRE_COMPRESSED, 2)]:
        w = _lib.WireC(kind=kind)
        a, keep = _args(wire=ctypes.pointer(w))
        assert _width(a) == want
    dom, nodes = uhat_domain(2, babybear_mont).wire(babybear_mont)
    a, keep = _args(wire=ctypes.pointer(dom))
    assert _width(a) == 2
    # compressed needs two product factors
    w = _lib.WireC(kind=_lib.ZS_WIRE_COMPRESSED)
    a, keep = _args(k=3, wire=ctypes.pointer(w))
    with pytest.raises(ValueError, match="exactly 2 factors"):
        _width(a)
    # a domain with more samples than the polynomial's values
    big, nodes = EvalDomain(array([0, 1, 2, 3], babybear_mont)).wire(babybear_mont)
    a, keep = _args(wire=ctypes.pointer(big))
    with pytest.raises(ValueError, match="samples"):
        _width(a)
    # table lengths are 32-bit in the kernels: at most 31 rounds
    a, keep = _args(rounds=_lib.ZS_MAX_ROUNDS)
    assert _width(a) == 3
    for rounds in (32, 33):
        a, keep = _args(rounds=rounds)
        with pytest.raises(ValueError, match="rounds must be in"):
            _width(a)


# ------------------------------------------------------------ GPU parity
def _wires_for(inst):
    """(name, wire description) pairs that apply to an instance."""
    base = inst.chal
    out = [("uhat", ("domain", uhat_domain(inst.degree, base))),
           ("coeffs", ("coeffs",)),
           ("nodes", ("domain", EvalDomain(array([0, 2, 4][:inst.degree + 1], base)))),
           ("inf_last", ("domain", compressed_domain(1, base)))]
    if inst.mode == "product" and inst.degree == 2:
        out.append(("compressed", ("compressed",)))
    return out


def _kind_domain(wire_py):
    if wire_py[0] == "domain":
        return _lib.ZS_WIRE_DOMAIN, wire_py[1]
    if wire_py[0] == "coeffs":
        return _lib.ZS_WIRE_COEFFS, None
    if wire_py[0] == "compressed":
        return _lib.ZS_WIRE_COMPRESSED, None
    return _lib.ZS_WIRE_VALUES, None


def _device_stage(inst, wire_py, tables=None):
    from zorch_amd import workloads as W
    from zorch_amd.sumcheck._device import prove_stage
    kind, dom = _kind_domain(wire_py)
    mode = _lib.ZS_MODE_EQ if inst.mode == "zerocheck" else _lib.ZS_MODE_PRODUCT
    return prove_stage(inst.tables if tables is None else tables, inst.claim, inst.chal, mode,
                       inst.degree, W.new_transcript(inst), inst.z, wire_kind=kind, domain=dom)


def _same_as_oracle(res, ref, words=True):
    assert res.msgs.words.tolist() == ref.msgs.tolist()
    assert res.point.words.tolist() == ref.point.tolist()
    assert res.value.words.reshape(-1).tolist() == ref.value.tolist()
    assert res.final.words.tolist() == ref.final.tolist()
    assert res.transcript.stream_state() == tuple(ref.state)


def _instances():
    from zorch_amd import workloads as W
    insts = [W.zerocheck_bb4(6, (201, 202), 203, 31), W.product_bb(5, (204, 205), 32),
             W.product_gl(6, (206, 207), 33), W.product_bb_ext(5, (208, 209), 34),
             W.zerocheck_bb4(5, (210, 211, 212), 213, 35), W.product_bb(4, (214,), 36),
             W.zerocheck_bb4(1, (215, 216), 217, 37)]
    b3 = W.zerocheck_bb4(6, (218, 219), 220, 38)
    b3.transcript_kind = "blake3"
    insts.append(b3)
    return insts


@pytest.mark.gpu
def test_device_wire_forms_match_oracle():
    from tests.oracle_glue import run_py_wire
    for inst in _instances():
        for name, wire_py in _wires_for(inst):
            if wire_py[0] == "domain" and wire_py[1].num_samples > inst.degree + 1 + (
                    inst.mode == "zerocheck"):
                continue
            ref, _ = run_py_wire(inst, wire_py)
            res = _device_stage(inst, wire_py)
            _same_as_oracle(res, ref)


@pytest.mark.gpu
@pytest.mark.parametrize("tune", [dict(), dict(tail_pairs=8, chain_pairs=64),
                                  dict(tail_pairs=4, chain_pairs=0), dict(lookahead=0,
                                                                          tail_pairs=16)])
def test_device_wire_forms_every_kernel_family(tune):
    """n = 10 with launch shapes that send rounds through rounds 0-2 from the
    input, pass, chained and tail kernels (plain and look-ahead)."""
    from tests.oracle_glue import run_py_wire
    from zorch_amd import workloads as W
    from zorch_amd.runtime import default_context
    ctx = default_context()
    old = ctx.get_tune()
    insts = [W.zerocheck_bb4(10, (231, 232), 233, 41), W.product_bb(10, (234, 235), 42),
             W.product_gl(10, (236, 237), 43)]
    try:
        ctx.set_tune(**tune)
        for inst in insts:
            for name, wire_py in _wires_for(inst):
                ref, _ = run_py_wire(inst, wire_py)
                _same_as_oracle(_device_stage(inst, wire_py), ref)
    finally:
        ctx.set_tune(**old)


@pytest.mark.gpu
def test_round_objects_match_oracle():
    """The drop-in round objects under fold_rounds and the stage fast path:
    StandardRound(domain), CompressedProductRound, CoeffsProductRound,
    EqPolyRound (Û by default), EqPolyCoeffsRound, and the verifiers."""
    from tests.oracle_glue import run_py_wire
    from zorch_amd import workloads as W
    from zorch_amd.challenge import ChallengePolicy
    from zorch_amd.prove import fold_rounds
    from zorch_amd.round import RunningClaim
    from zorch_amd.fields import zeros
    from zorch_amd.sumcheck.eq.eq_poly import EqPolyCoeffsRound, EqPolyRound
    from zorch_amd.sumcheck.prover import (CoeffsProductRound, CompressedProductRound,
                                           ProductSummand, StandardRound, initial_claim)
    from zorch_amd.sumcheck.stage import SumClaim, SumcheckProver, SumcheckVerifier, \
        SumcheckWitness
    from zorch_amd.sumcheck.verifier import CoeffsSumcheckRound, CompressedCoeffsSumcheckRound
    from zorch_amd.verify import verify

    def run(rnd, inst):
        t0 = W.new_transcript(inst)
        return fold_rounds(rnd, initial_claim(inst.tables, inst.claim, inst.n), t0, inst.n)

    def check(carry, t, msgs, ref):
        assert [m.words.tolist() for m in msgs] == ref.msgs.tolist()
        assert carry.claim.point.words.tolist() == ref.point.tolist()
        assert carry.claim.value.words.reshape(-1).tolist() == ref.value.tolist()
        assert t.stream_state() == tuple(ref.state)

    z = W.zerocheck_bb4(7, (241, 242), 243, 51)
    pol = ChallengePolicy(z.chal)
    ref, _ = run_py_wire(z, ("domain", uhat_domain(2, z.chal)))
    check(*run(EqPolyRound(ProductSummand(2), z.z, challenges=pol), z), ref)
    ref, _ = run_py_wire(z, ("coeffs",))
    carry, t, msgs = run(EqPolyCoeffsRound(ProductSummand(2), z.z, challenges=pol), z)
    check(carry, t, msgs, ref)
    # The coefficient form verifies round by round (eq_poly_test.py:86-111).
    from zorch_amd.fields import stack
    pt, val, _, ok = verify(CoeffsSumcheckRound(3, pol), z.claim, stack(msgs),
                            W.new_transcript(z))
    assert ok and pt.words.tolist() == ref.point.tolist()
    assert val.words.reshape(-1).tolist() == ref.value.tolist()
    # _round_poly / _round_coeffs: the first message without advancing.
    rnd = EqPolyRound(ProductSummand(2), z.z, challenges=pol)
    m0, _ = rnd._round_coeffs(z.tables)
    assert m0.words.tolist() == ref.msgs[0].tolist()

    p = W.product_bb(7, (244, 245), 52)
    pol = ChallengePolicy(p.chal)
    for rnd, wire_py, verifier in [
            (CompressedProductRound(pol), ("compressed",), CompressedCoeffsSumcheckRound(pol)),
            (CoeffsProductRound(ProductSummand(2), challenges=pol), ("coeffs",),
             CoeffsSumcheckRound(2, pol)),
            (StandardRound(ProductSummand(2), uhat_domain(2, p.chal), challenges=pol),
             ("domain", uhat_domain(2, p.chal)), None),
            (StandardRound(ProductSummand(2), natural_domain(2, p.chal), challenges=pol),
             ("values",), None)]:
        ref, _ = run_py_wire(p, wire_py)
        check(*run(rnd, p), ref)
        # The stage fast path (one zs_prove call) equals the rounds.
        res = SumcheckProver(rnd).prove(SumClaim(p.claim, p.n), SumcheckWitness(p.tables),
                                        W.new_transcript(p))
        assert res.reduction_proof.words.tolist() == ref.msgs.tolist()
        assert res.reduced_claim.value.words.reshape(-1).tolist() == ref.value.tolist()
        if verifier is not None:
            out = SumcheckVerifier(verifier).verify(SumClaim(p.claim, p.n),
                                                    res.reduction_proof, W.new_transcript(p))
            assert out.ok
            assert out.reduced_claim.value.words.tolist() == \
                res.reduced_claim.value.words.tolist()
    with pytest.raises(ValueError, match="exactly 2 factors"):
        three = W.product_bb(4, (246, 247, 248), 53)
        SumcheckProver(CompressedProductRound(ChallengePolicy(three.chal))).prove(
            SumClaim(three.claim, 4), SumcheckWitness(three.tables), W.new_transcript(three))


@pytest.mark.gpu
def test_eq_round_honours_the_state_eq_mass():
    """An EqPolyState tuple (tables, mass) starts from that eq mass."""
    from tests.oracle_glue import run_py_wire
    from zorch_amd import workloads as W
    from zorch_amd.challenge import ChallengePolicy
    from zorch_amd.prove import fold_rounds
    from zorch_amd.sumcheck.eq.eq_poly import EqPolyRound
    from zorch_amd.sumcheck.prover import ProductSummand, initial_claim
    z = W.zerocheck_bb4(6, (251, 252), 253, 61)
    mass = array((3, 1, 4, 1), babybearx4_mont)
    ref, _ = run_py_wire(z, ("values",), eq_mass=mass)
    rnd = EqPolyRound(ProductSummand(2), z.z, natural_domain(3, z.chal),
                      challenges=ChallengePolicy(z.chal))
    carry, t, msgs = fold_rounds(rnd, initial_claim((z.tables, mass), z.claim, z.n),
                                 W.new_transcript(z), z.n)
    assert [m.words.tolist() for m in msgs] == ref.msgs.tolist()
    assert carry.claim.value.words.reshape(-1).tolist() == ref.value.tolist()


@pytest.mark.gpu
def test_device_reference_kats():
    """The reference's integer KATs through the device: product_round_poly
    (domain_test.py:40-48), product_round_coeffs of one factor (:183-188)."""
    from zorch_amd.fields import stack
    from zorch_amd.sumcheck.domain import product_round_coeffs, product_round_poly
    a = array([1, 2, 3, 4], babybear_mont)
    b = array([5, 6, 7, 8], babybear_mont)
    assert product_round_poly(stack([a, b])).to_ints() == [8, 17]
    assert product_round_coeffs(array(list(range(1, 9)), babybear_mont).reshape(1, 8)) \
        .to_ints() == [10, 16]


@pytest.mark.gpu
def test_batched_and_served_wire_forms_equal_one_call():
    from zorch_amd import workloads as W
    from zorch_amd.serve import ProofServer
    from zorch_amd.sumcheck._device import make_job, prove_stage_batch
    insts, wires = [], []
    for j, n in enumerate([3, 8, 12, 14]):
        z = W.zerocheck_bb4(n, (261 + j, 271 + j), 281 + j, 291 + j)
        p = W.product_bb(n, (301 + j, 311 + j), 321 + j)
        for inst in (z, p):
            for name, wire_py in _wires_for(inst):
                insts.append(inst)
                wires.append(wire_py)
    singles = [_device_stage(i, w) for i, w in zip(insts, wires)]

    def job(i, w):
        kind, dom = _kind_domain(w)
        mode = _lib.ZS_MODE_EQ if i.mode == "zerocheck" else _lib.ZS_MODE_PRODUCT
        return make_job(i.tables, i.claim, i.chal, mode, i.degree, i.z, wire_kind=kind,
                        domain=dom)

    batch = prove_stage_batch([job(i, w) for i, w in zip(insts, wires)],
                              [W.new_transcript(i) for i in insts])
    for a, b in zip(batch, singles):
        assert a.msgs.words.tolist() == b.msgs.words.tolist()
        assert a.value.words.tolist() == b.value.words.tolist()
        assert a.transcript.stream_state() == b.transcript.stream_state()
    with ProofServer() as srv:
        ids = [srv.submit_job(job(i, w), W.new_transcript(i)) for i, w in zip(insts, wires)]
        for rid, b in zip(ids, singles):
            got, _ = srv.wait(rid)
            assert got.msgs.words.tolist() == b.msgs.words.tolist()
            assert got.value.words.tolist() == b.value.words.tolist()
            assert got.transcript.stream_state() == b.transcript.stream_state()
