"""GPU parity of batched proving (zs_prove_batch).

A batch mixes kinds (field pairs, product / zerocheck, factor counts) and
sizes; every member must equal the C oracle on the same inputs and the
one-proof-per-call path, bit for bit, whatever the launch shapes. Members
whose round 2 is a pass round take rounds 0-2 from the input (two launches
of their own in steps 0 and 1) unless the lookahead setting is off.
"""

from __future__ import annotations

import pytest

from tests.oracle_glue import run_c
from zorch_amd import _lib
from zorch_amd import workloads as W
from zorch_amd.runtime import DeviceTables, default_context
from zorch_amd.sumcheck._device import make_job, prove_stage, prove_stage_batch

pytestmark = pytest.mark.gpu


def _mode(inst):
    return _lib.ZS_MODE_EQ if inst.mode == "zerocheck" else _lib.ZS_MODE_PRODUCT


def _t0(inst):
    return W.new_transcript(inst)


def _batch(insts, tables=None):
    tables = tables or [i.tables for i in insts]
    jobs = [make_job(t, i.claim, i.chal, _mode(i), i.degree, i.z) for i, t in zip(insts, tables)]
    return prove_stage_batch(jobs, [_t0(i) for i in insts])


def _single(inst):
    return prove_stage(inst.tables, inst.claim, inst.chal, _mode(inst), inst.degree, _t0(inst),
                       inst.z)


def _same(a, b):
    assert a.msgs.words.tolist() == b.msgs.words.tolist()
    assert a.point.words.tolist() == b.point.words.tolist()
    assert a.value.words.tolist() == b.value.words.tolist()
    assert a.final.words.tolist() == b.final.words.tolist()
    assert a.transcript.stream_state() == b.transcript.stream_state()


def _matches_oracle(res, inst):
    ref = run_c(inst, threads=16)
    assert res.msgs.words.tolist() == ref.msgs.tolist()
    assert res.point.words.tolist() == ref.point.tolist()
    assert res.value.words.reshape(-1).tolist() == ref.value.tolist()
    h, pend, tot = res.transcript.stream_state()
    assert list(h) == list(ref.h) and pend == ref.pending and tot == ref.total


def _mixed():
    """Every supported kind, sizes from one round up to multi-block rounds."""
    insts = []
    for j, n in enumerate([1, 2, 5, 9, 12, 13, 14, 16]):
        insts.append(W.zerocheck_bb4(n, (100 + j, 200 + j), 300 + j, 400 + j))
        insts.append(W.product_bb_ext(n, (500 + j, 600 + j), 700 + j))
    insts.append(W.product_bb(13, (801, 802), 803))
    insts.append(W.product_bb(6, (804, 805, 806), 807))
    insts.append(W.product_gl(14, (811, 812), 813))
    insts.append(W.zerocheck_bb4(12, (821, 822, 823), 824, 825))
    # BLAKE3 transcripts in the same groups as SHA-256 ones.
    for j, n in enumerate([3, 13]):
        z = W.zerocheck_bb4(n, (830 + j, 840 + j), 850 + j, 860 + j)
        z.transcript_kind = "blake3"
        insts.append(z)
    return insts


def test_mixed_batch_matches_c_oracle():
    insts = _mixed()
    for res, inst in zip(_batch(insts), insts):
        if inst.transcript_kind == "sha256":
            _matches_oracle(res, inst)
        else:  # the C oracle restates SHA-256 only: compare with the one-call path
            _same(res, _single(inst))


@pytest.mark.parametrize("tune", [dict(tail_pairs=0, max_grid=4, pairs_per_thread=1),
                                  dict(tail_pairs=64, max_grid=4096, pairs_per_thread=4),
                                  dict(tail_pairs=1 << 20, max_grid=2048, pairs_per_thread=1),
                                  # tails split over several steps
                                  dict(tail_pairs=1 << 20, max_grid=512, pairs_per_thread=1,
                                       step_rounds=1),
                                  dict(tail_pairs=1024, max_grid=512, pairs_per_thread=2,
                                       step_rounds=3),
                                  # every member's rounds one step at a time (no rounds
                                  # 0-2 from the input)
                                  dict(tail_pairs=64, max_grid=4096, pairs_per_thread=4,
                                       lookahead=0)])
def test_batch_independent_of_launch_shapes(tune):
    ctx = default_context()
    insts = _mixed()
    ref = [_single(i) for i in insts]
    old = ctx.get_tune()
    try:
        ctx.set_tune(**tune)
        got = _batch(insts)
    finally:
        ctx.set_tune(**old)
    for a, b in zip(got, ref):
        _same(a, b)


def test_batch_spec_members_equal_single_proofs():
    """A slice of the 512-instance spec (sizes 12..18) with device-resident
    tables: each member equals its one-call proof and the verifier accepts."""
    spec = W.batch_spec(count=48, seed=4, lo=12, hi=18)
    insts = [W.batch_instance(spec, i) for i in range(48)]
    dev = [DeviceTables.from_host(i.tables) for i in insts]
    got = _batch(insts, dev)
    for res, inst in zip(got, insts):
        _same(res, _single(inst))
    for d in dev:
        d.free()


def test_batch_of_one_and_empty_batch():
    inst = W.zerocheck_bb4(10, (901, 902), 903, 904)
    (res,) = _batch([inst])
    _same(res, _single(inst))
    assert prove_stage_batch([], []) == []


def test_batch_512_spec_at_full_size():
    """The batched configuration as stated: 512 instances, n ~ U{12..22} from
    seed 4, product or zerocheck with probability 1/2, BabyBear^4 challenges,
    device-resident tables, one zs_prove_batch. Every member equals its
    one-call proof and the verifier accepts it; a seeded sample, the n >= 21
    members among it, equals the C oracle."""
    import numpy as np

    from zorch_amd.challenge import ChallengePolicy
    from zorch_amd.sumcheck.eq.stage import EqPolyVerifier, EqSumClaim
    from zorch_amd.sumcheck.prover import ProductSummand
    from zorch_amd.sumcheck.stage import SumClaim, SumcheckVerifier
    from zorch_amd.sumcheck.verifier import SumcheckRound

    spec = W.batch_spec(count=512)
    insts = W.batch_instances(spec)
    ctx = default_context()
    dev = [DeviceTables.from_host(i.tables, ctx) for i in insts]
    got = _batch(insts, dev)
    for res, inst, d in zip(got, insts, dev):
        single = prove_stage(d, inst.claim, inst.chal, _mode(inst), inst.degree, _t0(inst), inst.z)
        _same(res, single)
        pol = ChallengePolicy(inst.chal)
        if inst.mode == "zerocheck":
            ok = EqPolyVerifier(ProductSummand(inst.degree), challenges=pol).verify(
                EqSumClaim(inst.z, inst.claim, inst.n), res.msgs, _t0(inst)).ok
        else:
            ok = SumcheckVerifier(SumcheckRound(inst.degree, pol)).verify(
                SumClaim(inst.claim, inst.n), res.msgs, _t0(inst)).ok
        assert ok, inst.name
    rng = np.random.default_rng(4)
    big = [i for i, inst in enumerate(insts) if inst.n >= 21]
    sample = set(rng.choice(big, size=min(6, len(big)), replace=False).tolist())
    sample |= set(rng.choice(len(insts), size=6, replace=False).tolist())
    for i in sorted(sample):
        _matches_oracle(got[i], insts[i])
    for d in dev:
        d.free()
