"""GPU parity: the HIP prover against the oracle, bit for bit.

Small sizes compare with the committed golden fixtures (Python oracle);
larger sizes, up to the full benchmark shapes, compare with the C oracle run
on the same seeded inputs; full-size runs also check size-independent
properties (the verifier accepts; the reduced value equals the product of the
final folded tables).
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from tests.golden.recipes import RECIPES, build, fixture_state
from tests.oracle_glue import run_c
from zorch_amd import workloads as W
from zorch_amd.challenge import ChallengePolicy
from zorch_amd.fields import FieldArray, array, babybear_mont, babybearx4_mont, goldilocks
from zorch_amd.prove import fold_rounds
from zorch_amd.sha256_field_transcript import Sha256FieldTranscript
from zorch_amd.sumcheck.domain import natural_domain
from zorch_amd.sumcheck.eq.eq_poly import EqPolyRound
from zorch_amd.sumcheck.eq.stage import EqPolyProver, EqPolyVerifier, EqPolyWitness, EqSumClaim
from zorch_amd.sumcheck.prover import ProductSummand, StandardRound, initial_claim
from zorch_amd.sumcheck.stage import SumClaim, SumcheckProver, SumcheckVerifier, SumcheckWitness
from zorch_amd.sumcheck.verifier import SumcheckRound

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def _prove(inst, tables=None):
    t0 = W.new_transcript(inst)
    pol = ChallengePolicy(inst.chal)
    tables = inst.tables if tables is None else tables
    if inst.mode == "zerocheck":
        p = EqPolyProver(ProductSummand(inst.degree), challenges=pol)
        return p.prove(EqSumClaim(inst.z, inst.claim, inst.n), EqPolyWitness(tables), t0)
    p = SumcheckProver(StandardRound(ProductSummand(inst.degree), challenges=pol))
    return p.prove(SumClaim(inst.claim, inst.n), SumcheckWitness(tables), t0)


def _verify(inst, res):
    t0 = W.new_transcript(inst)
    pol = ChallengePolicy(inst.chal)
    if inst.mode == "zerocheck":
        v = EqPolyVerifier(ProductSummand(inst.degree), challenges=pol)
        return v.verify(EqSumClaim(inst.z, inst.claim, inst.n), res.reduction_proof, t0)
    v = SumcheckVerifier(SumcheckRound(inst.degree, pol))
    return v.verify(SumClaim(inst.claim, inst.n), res.reduction_proof, t0)


def _words(fa: FieldArray):
    return fa.words.reshape(fa.words.shape[0], -1).tolist() if fa.ndim else fa.words.tolist()


def _assert_matches(res, ref_msgs, ref_point, ref_value, ref_state):
    assert res.reduction_proof.words.tolist() == ref_msgs
    assert res.reduced_claim.point.words.tolist() == ref_point
    assert res.reduced_claim.value.words.reshape(-1).tolist() == ref_value
    assert res.transcript.stream_state() == tuple(ref_state)


@pytest.mark.parametrize("name", sorted(RECIPES))
def test_stage_matches_golden_fixture(name):
    fx = json.loads((GOLDEN / f"{name}.json").read_text())
    inst = build(fx["recipe"])
    res = _prove(inst)
    _assert_matches(res, fx["msgs"], fx["point"], fx["value"], fixture_state(fx))
    assert _verify(inst, res).ok


@pytest.mark.parametrize("name", ["product_bb_n4", "product_bb_k3_n4", "product_bb_bb4_n6",
                                  "zerocheck_bb4_n5", "zerocheck_bb4_t4_n5", "product_gl_n5",
                                  "product_bb_n1", "zerocheck_bb4_n2", "zerocheck_bb4_n9_b3",
                                  "product_gl_n6_b3", "zerocheck_bb4_t4_n5_b3"])
def test_round_by_round_matches_golden_fixture(name):
    """The drop-in per-round path under fold_rounds equals the whole stage."""
    fx = json.loads((GOLDEN / f"{name}.json").read_text())
    inst = build(fx["recipe"])
    t0 = W.new_transcript(inst)
    pol = ChallengePolicy(inst.chal)
    if inst.mode == "zerocheck":
        # The natural domain, as EqPolyProver passes it (eq/stage.py:70-76).
        rnd = EqPolyRound(ProductSummand(inst.degree), inst.z,
                          natural_domain(inst.degree + 1, inst.chal), challenges=pol)
    else:
        rnd = StandardRound(ProductSummand(inst.degree), challenges=pol)
    carry, t, msgs = fold_rounds(rnd, initial_claim(inst.tables, inst.claim, inst.n), t0, inst.n)
    assert [m.words.tolist() for m in msgs] == fx["msgs"]
    assert carry.claim.point.words.tolist() == fx["point"]
    assert carry.claim.value.words.reshape(-1).tolist() == fx["value"]
    assert carry.claim.index == inst.n
    final = carry.state.read_tables()
    assert final.words[:, 0].tolist() == fx["final"][: inst.degree]
    assert t.stream_state() == fixture_state(fx)
    # A consumed carry cannot be replayed.
    with pytest.raises(ValueError):
        rnd(initial_claim(carry.state, inst.claim, inst.n), t0)


def _compare_with_c(inst, threads=16):
    res = _prove(inst)
    ref = run_c(inst, threads=threads)
    _assert_matches(res, ref.msgs.tolist(), ref.point.tolist(), ref.value.tolist(), ref.state)
    return res


@pytest.mark.parametrize("n", [11, 12, 13, 14, 16, 20])
def test_zerocheck_matches_c_oracle(n):
    inst = W.zerocheck_bb4(n, (21 + n, 22 + n), 23 + n, n)
    res = _compare_with_c(inst)
    assert _verify(inst, res).ok


@pytest.mark.parametrize("n", [12, 13, 18])
def test_product_bb_matches_c_oracle(n):
    _compare_with_c(W.product_bb(n, (31 + n, 32 + n), n))


@pytest.mark.parametrize("n", [12, 13, 20])
def test_product_gl_matches_c_oracle(n):
    _compare_with_c(W.product_gl(n, (41 + n, 42 + n), n))


@pytest.mark.parametrize("n", [12, 17])
def test_product_ext_challenges_match_c_oracle(n):
    _compare_with_c(W.product_bb_ext(n, (51 + n, 52 + n), n))


def test_more_factors_match_c_oracle():
    _compare_with_c(W.product_bb(13, (61, 62, 63), 1))
    _compare_with_c(W.product_bb(12, (64, 65, 66, 67), 2))
    _compare_with_c(W.zerocheck_bb4(13, (68,), 69, 3))
    _compare_with_c(W.zerocheck_bb4(12, (70, 71, 72, 73), 74, 4))


def test_zerocheck_n24_full_size():
    """The benchmark shape: parity with the C oracle and the verifier accepts."""
    inst = W.zerocheck_bb4(24)
    res = _compare_with_c(inst)
    assert _verify(inst, res).ok


def test_product_gl_n27_full_size():
    inst = W.product_gl(27)
    res = _compare_with_c(inst)
    assert _verify(inst, res).ok


def test_product_bb_n16_full_size_and_determinism():
    inst = W.product_bb(16)
    a = _compare_with_c(inst)
    b = _prove(inst)
    assert a.reduction_proof.words.tolist() == b.reduction_proof.words.tolist()


def _prove_with_tune(inst, **tune):
    from zorch_amd.runtime import default_context
    ctx = default_context()
    old = ctx.get_tune()
    try:
        ctx.set_tune(**tune)
        return _prove(inst)
    finally:
        ctx.set_tune(**old)


@pytest.mark.parametrize("tune", [
    dict(chain_pairs=1 << 20, max_grid=8),                   # few blocks, many pairs each
    dict(chain_pairs=1 << 20, tail_pairs=0),                 # chain runs to the last round
    dict(chain_pairs=1 << 20, tail_pairs=64, pairs_per_thread=4, max_grid=4096),
    dict(chain_pairs=1 << 12, tail_pairs=256, max_grid=64),
])
def test_chained_rounds_equal_one_launch_per_round(tune):
    """Rounds run in one persistent launch equal one launch per round, bit
    for bit, whatever the launch shapes; the C oracle anchors one of them."""
    insts = [W.zerocheck_bb4(17, (91, 92), 93, 11), W.product_gl(16, (94, 95), 12),
             W.product_bb(15, (96, 97, 98), 13), W.zerocheck_bb4(16, (99, 100), 101, 14)]
    insts[3].transcript_kind = "blake3"
    for inst in insts:
        ref = _prove_with_tune(inst, chain_pairs=0)
        got = _prove_with_tune(inst, **tune)
        assert got.reduction_proof.words.tolist() == ref.reduction_proof.words.tolist()
        assert got.reduced_claim.point.words.tolist() == ref.reduced_claim.point.words.tolist()
        assert got.reduced_claim.value.words.tolist() == ref.reduced_claim.value.words.tolist()
        assert got.transcript.stream_state() == ref.transcript.stream_state()
    _compare_with_c(insts[0])


@pytest.mark.parametrize("tune", [
    dict(tail_pairs=2), dict(tail_pairs=64), dict(tail_pairs=1024), dict(tail_pairs=4096),
    # chained rounds too: to the tail, to the last round, over few blocks
    dict(tail_pairs=64, chain_pairs=1 << 20),
    dict(tail_pairs=0, chain_pairs=1 << 20),
    dict(tail_pairs=8, chain_pairs=1 << 20, max_grid=4),
    dict(tail_pairs=256, chain_pairs=1 << 12, pairs_per_thread=4),
    # rounds 0-2 from the input (round 2 a pass round)
    dict(tail_pairs=16, chain_pairs=0),
    dict(tail_pairs=8, chain_pairs=64),
    dict(tail_pairs=0, chain_pairs=0, max_grid=4),
])
def test_lookahead_equals_plain_rounds(tune):
    """Tails and chained launches that fold each round while the previous
    message is hashed, and rounds 0-2 computed from the input, equal the
    round-by-round kernels, bit for bit, over fields, factor counts, both
    modes, both transcripts and every start and end round; the C oracle
    anchors one of them."""
    insts = [W.zerocheck_bb4(15, (111, 112), 113, 21), W.product_gl(14, (114, 115), 22),
             W.product_bb(13, (116, 117, 118), 23), W.product_bb(12, (119, 120, 121, 122), 24),
             W.zerocheck_bb4(13, (123, 124, 125), 126, 25), W.product_bb_ext(12, (127,), 26),
             W.zerocheck_bb4(14, (128, 129), 130, 27)]
    insts[6].transcript_kind = "blake3"
    for inst in insts:
        ref = _prove_with_tune(inst, **tune, lookahead=0)
        got = _prove_with_tune(inst, **tune, lookahead=1)
        assert got.reduction_proof.words.tolist() == ref.reduction_proof.words.tolist()
        assert got.reduced_claim.point.words.tolist() == ref.reduced_claim.point.words.tolist()
        assert got.reduced_claim.value.words.tolist() == ref.reduced_claim.value.words.tolist()
        assert got.transcript.stream_state() == ref.transcript.stream_state()
    _compare_with_c(insts[0])


def test_chained_rounds_round_by_round():
    """The per-round drop-in path runs chained single rounds; it equals the
    whole stage."""
    inst = W.zerocheck_bb4(17, (102, 103), 104, 15)
    whole = _prove(inst)
    t0 = W.new_transcript(inst)
    rnd = EqPolyRound(ProductSummand(inst.degree), inst.z, natural_domain(inst.degree + 1, inst.chal),
                      challenges=ChallengePolicy(inst.chal))
    carry, t, msgs = fold_rounds(rnd, initial_claim(inst.tables, inst.claim, inst.n), t0, inst.n)
    assert [m.words.tolist() for m in msgs] == whole.reduction_proof.words.tolist()
    assert carry.claim.value.words.tolist() == whole.reduced_claim.value.words.tolist()
    assert t.stream_state() == whole.transcript.stream_state()


def test_device_resident_tables_equal_host_tables():
    from zorch_amd.runtime import DeviceTables
    inst = W.zerocheck_bb4(14, (81, 82), 83, 9)
    dev = DeviceTables.from_host(inst.tables)
    assert dev.to_host().words.tolist() == inst.tables.words.tolist()
    a = _prove(inst)
    b = _prove(inst, dev)
    assert a.reduction_proof.words.tolist() == b.reduction_proof.words.tolist()
    assert a.reduced_claim.point.words.tolist() == b.reduced_claim.point.words.tolist()
    dev.free()


def test_structural_errors_raise_value_error():
    inst = W.product_bb(6)
    t0 = Sha256FieldTranscript.new(inst.label, babybear_mont)
    with pytest.raises(ValueError):  # degree != number of tables
        SumcheckProver(StandardRound(ProductSummand(3), challenges=ChallengePolicy(
            babybear_mont))).prove(SumClaim(inst.claim, 6), SumcheckWitness(inst.tables), t0)
    with pytest.raises(ValueError):  # extension challenges over... a base policy on an ext transcript
        SumcheckProver(StandardRound(ProductSummand(2), challenges=ChallengePolicy(
            babybear_mont))).prove(SumClaim(inst.claim, 6), SumcheckWitness(inst.tables),
                                   Sha256FieldTranscript.new(b"", babybearx4_mont))
    z = W.ext_point(3, 5)
    with pytest.raises(ValueError):  # equality point length
        EqPolyProver(ProductSummand(2), challenges=ChallengePolicy(babybearx4_mont)).prove(
            EqSumClaim(z, inst.claim.astype(babybearx4_mont), 6), EqPolyWitness(inst.tables), t0)
