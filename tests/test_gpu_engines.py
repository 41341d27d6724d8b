"""The round-cost engines on the GPU against the oracle (oracle/engines.py).

sqrt-space (``prove_sqrt_space``, ``SqrtSpaceRound``, ``zs_prove`` with
``refold_rounds``): messages, raw point, reduced value, final tables and
transcript state equal the oracle's restatement bit for bit on every wire
form, field pair and factor count, whole-stage and round by round; at full
size (natural values) they equal the C oracle's linear-time proof, which the
engine must reproduce (reference sqrt_space_test.py:52-111).
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import engines as OE
from tests.oracle_glue import oracle_inputs, oracle_wire, proof_result, run_c
from zorch_amd import _lib
from zorch_amd import workloads as W
from zorch_amd.fields import array
from zorch_amd.sumcheck.domain import EvalDomain, natural_domain, uhat_domain

pytestmark = pytest.mark.gpu


def _same(res, ref):
    assert res.msgs.words.tolist() == ref.msgs.tolist()
    assert res.point.words.tolist() == ref.point.tolist()
    assert res.value.words.reshape(-1).tolist() == ref.value.tolist()
    assert res.final.words.tolist() == ref.final.tolist()
    assert res.transcript.stream_state() == tuple(ref.state)


def _kind_domain(wire_py):
    if wire_py[0] == "domain":
        return _lib.ZS_WIRE_DOMAIN, wire_py[1]
    if wire_py[0] == "coeffs":
        return _lib.ZS_WIRE_COEFFS, None
    if wire_py[0] == "compressed":
        return _lib.ZS_WIRE_COMPRESSED, None
    return _lib.ZS_WIRE_VALUES, None


def _sqrt_device(inst, wire_py, refold=None, tables=None):
    from zorch_amd.sumcheck._device import prove_stage
    kind, dom = _kind_domain(wire_py)
    return prove_stage(inst.tables if tables is None else tables, inst.claim, inst.chal,
                       _lib.ZS_MODE_PRODUCT, inst.degree, W.new_transcript(inst),
                       wire_kind=kind, domain=dom,
                       refold_rounds=inst.n // 2 if refold is None else refold)


def _sqrt_oracle(inst, wire_py):
    tf, cf, trf, tabs, claim, tr = oracle_inputs(inst)
    _, wire = oracle_wire(inst, wire_py)
    pf = OE.prove_sqrt_space(cf, trf, tabs, claim, tr, inst.degree, wire)
    return proof_result(inst, cf, pf)


def _sqrt_instances():
    insts = [W.product_bb(6, (301, 302), 41), W.product_bb_ext(7, (303, 304), 42),
             W.product_gl(6, (305, 306), 43), W.product_bb(5, (307, 308, 309), 44),
             W.product_bb_ext(6, (310, 311, 312, 313), 45), W.product_bb(4, (314,), 46),
             W.product_bb(1, (315, 316), 47), W.product_bb(2, (317, 318), 48),
             W.product_bb_ext(3, (319, 320), 49)]
    b3 = W.product_bb_ext(6, (321, 322), 50)
    b3.transcript_kind = "blake3"
    insts.append(b3)
    return insts


def _wires(inst):
    out = [("values",), ("domain", uhat_domain(inst.degree, inst.chal)), ("coeffs",)]
    if inst.degree == 2:
        out.append(("compressed",))
    return out


def test_sqrt_space_matches_oracle_every_wire():
    for inst in _sqrt_instances():
        for wire_py in _wires(inst):
            print(f"sqrt-space {inst.name} k={inst.degree} {wire_py[0]}", flush=True)
            _same(_sqrt_device(inst, wire_py), _sqrt_oracle(inst, wire_py))


@pytest.mark.parametrize("tune", [dict(max_grid=1), dict(max_grid=4), dict(max_grid=64),
                                  dict(tail_pairs=0, chain_pairs=0), dict(lookahead=0)])
def test_sqrt_space_launch_shapes(tune):
    """Few blocks force many pair groups a block and prefix sums spread over
    slices; results never depend on the launch shapes."""
    from zorch_amd.runtime import default_context
    ctx = default_context()
    saved = ctx.get_tune()
    try:
        ctx.set_tune(**tune)
        for inst in (W.product_bb_ext(12, (331, 332), 51), W.product_gl(11, (333, 334), 52),
                     W.product_bb(13, (335, 336, 337), 53)):
            wire_py = ("domain", uhat_domain(inst.degree, inst.chal))
            print(f"sqrt-space shapes {tune} {inst.name}", flush=True)
            _same(_sqrt_device(inst, wire_py), _sqrt_oracle(inst, wire_py))
    finally:
        ctx.set_tune(**saved)


def test_refold_rounds_any_count_equals_dense():
    """Any number of refold rounds up to n/2 sends the dense prover's proof."""
    inst = W.product_bb_ext(10, (341, 342), 54)
    wire_py = ("domain", uhat_domain(2, inst.chal))
    dense = _sqrt_device(inst, wire_py, refold=0)
    for h in range(1, 6):
        print(f"refold rounds {h}", flush=True)
        res = _sqrt_device(inst, wire_py, refold=h)
        assert res.msgs.words.tolist() == dense.msgs.words.tolist()
        assert res.value.words.tolist() == dense.value.words.tolist()
        assert res.final.words.tolist() == dense.final.words.tolist()
        assert res.transcript.stream_state() == dense.transcript.stream_state()


@pytest.mark.parametrize("name", ["product_bb_bb4_n22", "product_gl_n22"])
def test_sqrt_space_full_size_equals_c_oracle(name):
    """n = 22 with natural-domain messages: the C oracle's linear-time proof."""
    inst = (W.product_bb_ext(22, (1, 2), 0) if "bb4" in name else W.product_gl(22))
    print(f"sqrt-space full size {name}", flush=True)
    res = _sqrt_device(inst, ("values",))
    ref = run_c(inst)
    assert res.msgs.words.tolist() == ref.msgs.tolist()
    assert res.point.words.tolist() == ref.point.tolist()
    assert res.value.words.reshape(-1).tolist() == ref.value.tolist()
    assert res.transcript.stream_state()[:3] == (tuple(ref.h), ref.pending, ref.total)


def test_prove_sqrt_space_api_and_rounds():
    """prove_sqrt_space (U-hat by default) and the per-round SqrtSpaceRound for
    l//2 rounds then StandardRound on the same state give the oracle's proof;
    a device-resident input gives the same."""
    from zorch_amd.challenge import ChallengePolicy
    from zorch_amd.prove import fold_rounds
    from zorch_amd.runtime import DeviceTables
    from zorch_amd.sumcheck.prover import ProductSummand, StandardRound, initial_claim
    from zorch_amd.sumcheck.sqrt_space import SqrtSpaceRound, prove_sqrt_space
    for inst in (W.product_bb_ext(8, (351, 352), 55), W.product_bb(7, (353, 354, 355), 56)):
        wire_py = ("domain", uhat_domain(inst.degree, inst.chal))
        ref = _sqrt_oracle(inst, wire_py)
        pol = ChallengePolicy(inst.chal)
        carry, t, msgs = prove_sqrt_space(inst.tables, inst.claim, W.new_transcript(inst),
                                          challenges=pol)
        assert [m.words.tolist() for m in msgs] == ref.msgs.tolist()
        assert carry.state.shape == (inst.degree, 1)
        assert carry.claim.value.words.reshape(-1).tolist() == ref.value.tolist()
        assert carry.claim.point.words.tolist() == ref.point.tolist()
        assert t.stream_state() == tuple(ref.state)
        dev = DeviceTables.from_host(inst.tables)
        carry2, t2, msgs2 = prove_sqrt_space(dev, inst.claim, W.new_transcript(inst),
                                             challenges=pol)
        assert [m.words.tolist() for m in msgs2] == ref.msgs.tolist()
        dom = uhat_domain(inst.degree, inst.chal)
        summand = ProductSummand(inst.degree)
        half = inst.n // 2
        c = initial_claim(inst.tables, inst.claim, inst.n)
        c, tr, m1 = fold_rounds(SqrtSpaceRound(summand, dom, pol), c, W.new_transcript(inst), half)
        c, tr, m2 = fold_rounds(StandardRound(summand, dom, challenges=pol), c, tr,
                                inst.n - half)
        assert [m.words.tolist() for m in m1 + m2] == ref.msgs.tolist()
        assert c.claim.value.words.reshape(-1).tolist() == ref.value.tolist()
        assert tr.stream_state() == tuple(ref.state)
        assert c.state.read_tables().words.reshape(inst.degree, -1).tolist() == \
            ref.final.tolist()


# ----------------------------------------------------------- univariate skip
def _skip_oracle(inst, skip):
    """Oracle skip proof: (msg0 storage words, tail result with r0 first)."""
    tf, cf, trf, tabs, claim, tr = oracle_inputs(inst)
    msg0, pf = OE.prove_univariate_skip(tf, cf, trf, tabs, tf.from_words(
        [int(w) for w in W.product_claim(inst.tables, inst.tables.field).words.reshape(-1)]), skip, tr,
        inst.degree)
    words = None if msg0 is None else [tf.to_words(m) for m in msg0]
    return words, proof_result(inst, cf, pf)


def _skip_device(inst, skip, tables=None):
    from zorch_amd.challenge import ChallengePolicy
    from zorch_amd.sumcheck.univariate_skip import prove_univariate_skip
    return prove_univariate_skip(inst.tables if tables is None else tables,
                                 W.product_claim(inst.tables, inst.tables.field), skip,
                                 W.new_transcript(inst), challenges=ChallengePolicy(inst.chal))


def _same_skip(dev, ref_msg0, ref):
    carry, t, msgs = dev
    assert msgs[0].words.reshape(-1).tolist() == [w for m in ref_msg0 for w in m]
    assert [m.words.tolist() for m in msgs[1:]] == ref.msgs.tolist()
    assert carry.claim.point.words.tolist() == ref.point.tolist()
    assert carry.claim.value.words.reshape(-1).tolist() == ref.value.tolist()
    assert carry.state.words.reshape(len(ref.final), -1).tolist() == ref.final.tolist()
    assert t.stream_state() == tuple(ref.state)


def _skip_instances():
    return [W.product_bb_ext(7, (401, 402), 61), W.product_bb(6, (403, 404), 62),
            W.product_gl(6, (405, 406), 63), W.product_bb_ext(6, (407, 408, 409), 64),
            W.product_bb_ext(5, (410,), 65), W.product_bb(6, (411, 412, 413, 414), 66)]


@pytest.mark.parametrize("skip", [1, 2, 3, 4])
def test_univariate_skip_matches_oracle(skip):
    for inst in _skip_instances():
        print(f"skip {skip} {inst.name} k={inst.degree}", flush=True)
        ref_msg0, ref = _skip_oracle(inst, skip)
        _same_skip(_skip_device(inst, skip), ref_msg0, ref)


def test_univariate_skip_all_rounds_and_blake3():
    inst = W.product_bb_ext(4, (421, 422), 67)
    ref_msg0, ref = _skip_oracle(inst, 4)   # every variable skipped: no tail
    _same_skip(_skip_device(inst, 4), ref_msg0, ref)
    b3 = W.product_bb_ext(8, (423, 424), 68)
    b3.transcript_kind = "blake3"
    ref_msg0, ref = _skip_oracle(b3, 3)
    _same_skip(_skip_device(b3, 3), ref_msg0, ref)


def test_univariate_skip_full_size_round_trip():
    """n = 22 (skip 3 and 5): the subgroup sum of msg0 is the claim, the
    verifier accepts, and the reduced value is the product of the bound
    factors; a resident input gives the same proof."""
    from zorch_amd.challenge import ChallengePolicy
    from zorch_amd.runtime import DeviceTables
    from zorch_amd.sumcheck.domain import subgroup_sum
    from zorch_amd.sumcheck.univariate_skip import verify_univariate_skip
    inst = W.product_bb_ext(22, (1, 2), 0)
    dev = DeviceTables.from_host(inst.tables)
    base_claim = W.product_claim(inst.tables, inst.tables.field)
    for skip in (3, 5):
        print(f"skip full size {skip}", flush=True)
        carry, t, msgs = _skip_device(inst, skip)
        assert bool((subgroup_sum(msgs[0], skip) == base_claim).all())
        value, t2, point, ok = verify_univariate_skip(
            base_claim, msgs, skip, 22, W.new_transcript(inst), 2,
            challenges=ChallengePolicy(inst.chal))
        assert ok
        assert value.words.tolist() == carry.claim.value.words.tolist()
        assert point.words.tolist() == carry.claim.point.words.tolist()
        assert t2.stream_state() == t.stream_state()
        fin = carry.state
        assert (fin[0, 0] * fin[1, 0]).words.tolist() == value.words.tolist()
        carry2, _, msgs2 = _skip_device(inst, skip, tables=dev)
        assert [m.words.tolist() for m in msgs2] == [m.words.tolist() for m in msgs]


def test_skip_round0_then_standard_tail_and_stages():
    """skip_round0 + StandardRound under fold_rounds equals the one-call skip;
    the prover and verifier stages agree; skip 0 is the plain run."""
    from zorch_amd.challenge import ChallengePolicy
    from zorch_amd.prove import fold_rounds
    from zorch_amd.round import RunningClaim
    from zorch_amd.fields import zeros
    from zorch_amd.sumcheck.prover import ProductSummand, StandardRound
    from zorch_amd.sumcheck.stage import SumClaim, SumcheckWitness
    from zorch_amd.sumcheck.univariate_skip import (UnivariateSkipProver, UnivariateSkipVerifier,
                                                    round0_message, skip_round0)
    from zorch_amd.sumcheck.domain import subgroup_sum
    inst = W.product_bb_ext(9, (431, 432), 69)
    pol = ChallengePolicy(inst.chal)
    base_claim = W.product_claim(inst.tables, inst.tables.field)
    ref_carry, ref_t, ref_msgs = _skip_device(inst, 2)
    start = RunningClaim(base_claim, zeros((1 + 7,), inst.chal), 0)
    carry, t, msg0 = skip_round0(inst.tables, start, 2, W.new_transcript(inst),
                                 ProductSummand(2), pol)
    assert msg0.words.tolist() == ref_msgs[0].words.tolist()
    carry, t, tail = fold_rounds(StandardRound(ProductSummand(2), challenges=pol), carry, t, 7)
    assert [m.words.tolist() for m in tail] == [m.words.tolist() for m in ref_msgs[1:]]
    assert carry.claim.value.words.tolist() == ref_carry.claim.value.words.tolist()
    assert carry.claim.point.words.tolist() == ref_carry.claim.point.words.tolist()
    assert t.stream_state() == ref_t.stream_state()
    m0, coeffs_z = round0_message(inst.tables, 2, ProductSummand(2))
    assert m0.words.tolist() == msg0.words.tolist()
    assert coeffs_z.shape == (2, 1 << 7, 4)
    assert bool((subgroup_sum(m0, 2) == base_claim).all())
    prover = UnivariateSkipProver(2, ProductSummand(2), challenges=pol)
    verifier = UnivariateSkipVerifier(2, ProductSummand(2), challenges=pol)
    res = prover.prove(SumClaim(base_claim, 9), SumcheckWitness(inst.tables),
                       W.new_transcript(inst))
    vr = verifier.verify(SumClaim(base_claim, 9), res.reduction_proof, W.new_transcript(inst))
    assert vr.ok
    assert vr.reduced_claim.value.words.tolist() == res.reduced_claim.value.words.tolist()
    _, pt = res.transcript.sample(1)
    _, vt = vr.transcript.sample(1)
    assert pt.words.tolist() == vt.words.tolist()
    with pytest.raises(ValueError):
        UnivariateSkipProver(0, ProductSummand(2), challenges=pol)


# ---------------------------------------------------------------- small value
def _sv_instances():
    from zorch_amd.fields import babybearx4_mont
    return [W.zerocheck_bb4(6, (501, 502), 503, 71, transcript_field=babybearx4_mont),
            W.zerocheck_bb4(7, (504, 505, 506), 507, 72, transcript_field=babybearx4_mont),
            W.zerocheck_bb4(8, (508, 509, 510, 511), 512, 73, transcript_field=babybearx4_mont)]


def _sv_oracle(inst, l0):
    tf, cf, trf, tabs, claim, tr = oracle_inputs(inst)
    z = [cf.from_words([int(w) for w in inst.z.words[i]]) for i in range(inst.n)]
    pf = OE.prove_eq_small_value(cf, trf, tabs, z, l0, claim, tr)
    return proof_result(inst, cf, pf, eq=True)


def _sv_device(inst, l0):
    from zorch_amd.challenge import ChallengePolicy
    from zorch_amd.sumcheck.eq.small_value import prove_eq_poly_small_value
    return prove_eq_poly_small_value(inst.tables, inst.z, l0, inst.claim, W.new_transcript(inst),
                                     challenges=ChallengePolicy(inst.chal))


def test_small_value_matches_oracle():
    for inst in _sv_instances():
        for l0 in range(1, inst.n - inst.n // 2 + 1):
            if (inst.degree + 1) ** l0 > 256:
                continue
            print(f"small-value {inst.name} k={inst.degree} l0={l0}", flush=True)
            ref = _sv_oracle(inst, l0)
            carry, t, msgs = _sv_device(inst, l0)
            assert [m.words.tolist() for m in msgs] == ref.msgs.tolist()
            assert carry.claim.point.words.tolist() == ref.point.tolist()
            assert carry.claim.value.words.reshape(-1).tolist() == ref.value.tolist()
            fin, mass = carry.state
            got_fin = fin.words.reshape(inst.degree, -1).tolist() + [mass.words.tolist()]
            assert got_fin == ref.final.tolist()
            assert t.stream_state() == tuple(ref.state)


def test_small_value_equals_eq_poly_rounds_at_size():
    """n = 20: the small-value engine's messages equal EqPolyRound's on the
    compressed domain (the dense device prover with the U-hat wire)."""
    from zorch_amd.fields import babybearx4_mont
    from zorch_amd.sumcheck._device import prove_stage
    inst = W.zerocheck_bb4(20, (1, 2), 3, 0, transcript_field=babybearx4_mont)
    dense = prove_stage(inst.tables, inst.claim, inst.chal, _lib.ZS_MODE_EQ, 2,
                        W.new_transcript(inst), inst.z, wire_kind=_lib.ZS_WIRE_DOMAIN,
                        domain=uhat_domain(2, inst.chal))
    for l0 in (1, 2, 3, 4, 5):
        print(f"small-value n=20 l0={l0}", flush=True)
        carry, t, msgs = _sv_device(inst, l0)
        assert [m.words.tolist() for m in msgs] == dense.msgs.words.tolist()
        assert carry.claim.value.words.tolist() == dense.value.words.tolist()
        assert t.stream_state() == dense.transcript.stream_state()
