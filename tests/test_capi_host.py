"""The C ABI library on the host (no GPU compute): it loads, exports every
symbol include/zorch_amd.h declares, and its host transcript and verifier
agree with the oracle."""

from __future__ import annotations

import hashlib
import json
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import blake3 as OB3
from oracle import transcript as OT
from tests.golden.recipes import RECIPES, build, fixture_state
from zorch_amd import _lib
from zorch_amd import workloads as W
from zorch_amd.challenge import ChallengePolicy
from zorch_amd.fields import (
    FieldArray,
    array,
    babybear_mont,
    babybearx4_mont,
    from_storage,
    goldilocks,
)
from zorch_amd.blake3_field_transcript import Blake3FieldTranscript
from zorch_amd.sha256_field_transcript import Sha256FieldTranscript
from zorch_amd.sumcheck.stage import SumClaim, SumcheckVerifier
from zorch_amd.sumcheck.eq.stage import EqPolyVerifier, EqSumClaim
from zorch_amd.sumcheck.prover import ProductSummand
from zorch_amd.sumcheck.verifier import SumcheckRound
from zorch_amd.transcript import reinterpret_challenge, sample_challenge

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"


def test_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "zorch_amd.h").read_text()
    declared = set(re.findall(r"\b(zs_[a-z0-9_]+)\s*\(", header))
    lib = _lib.lib()
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert declared == set(_lib.SIGNATURES), declared ^ set(_lib.SIGNATURES)
    assert b"gfx950" in lib.zs_version()


def _oracle_item(dtype):
    return 16 if dtype is babybearx4_mont else (8 if dtype is goldilocks else 4)


@pytest.mark.parametrize("kind", ["sha256", "blake3"])
@pytest.mark.parametrize("dtype", [babybear_mont, babybearx4_mont, goldilocks])
def test_host_transcript_matches_oracle(dtype, kind):
    item = _oracle_item(dtype)
    cls, ocls = ((Blake3FieldTranscript, OT.OracleBlake3Transcript) if kind == "blake3"
                 else (Sha256FieldTranscript, OT.OracleTranscript))
    t = cls.new(b"domain-label", dtype)
    o = ocls.new(b"domain-label", item)
    assert t.stream_state() == o.state()
    vals = array(list(range(1, 7)), dtype)
    t = t.observe(vals)
    o = o.observe(vals.tobytes())
    t, s = t.sample(5)
    o, so = o.sample(5)
    assert s.tobytes() == so
    t = t.observe_label(b"lbl").observe_bytes(b"\x00\x01\x02")
    o = o.observe_label(b"lbl").observe_bytes(b"\x00\x01\x02")
    t, s1 = t.sample_scalar()
    o, so1 = o.sample_scalar()
    assert s1.tobytes() == so1
    t = t.observe_scalar(vals)
    o = o.observe_scalar(vals.tobytes())
    t, s2 = t.sample(0)
    o, so2 = o.sample(0)
    assert s2.shape == (0,) and so2 == b""
    assert t.stream_state() == o.state()
    # Long streams: chunk boundaries and subtree merges (BLAKE3), many blocks.
    rng = np.random.default_rng(len(kind) + item)
    for _ in range(40):
        blob = rng.bytes(int(rng.integers(0, 400)))
        t, o = t.observe_bytes(blob), o.observe_bytes(blob)
        t, a = t.sample(int(rng.integers(1, 9)))
        o, b = o.sample(a.shape[0])
        assert a.tobytes() == b
    assert t.stream_state() == o.state()
    assert t.digest() == (OB3.blake3(o.stream) if kind == "blake3"
                          else hashlib.sha256(o.stream).digest())


def test_transcripts_are_immutable_values():
    t0 = Sha256FieldTranscript.new(b"x", babybear_mont)
    t1, a = t0.sample(2)
    t2, b = t0.sample(2)
    assert a.tobytes() == b.tobytes() and t1 == t2 and t0 != t1


def test_challenge_policy_packing():
    # Extension policy over a base transcript: 4 words regrouped as one element
    # (zorch/testing/transcript_test.py:191-197).
    t = Sha256FieldTranscript.new(b"p", babybear_mont)
    _, ch = ChallengePolicy(babybearx4_mont).sample(t)
    _, raw = t.sample(4)
    assert ch.field is babybearx4_mont and ch.tobytes() == raw.tobytes()
    assert ChallengePolicy(babybearx4_mont).limbs_over(babybearx4_mont) == 1
    with pytest.raises(ValueError):
        ChallengePolicy(babybear_mont).limbs_over(babybearx4_mont)
    with pytest.raises(ValueError):
        sample_challenge(t, babybearx4_mont, 8)
    with pytest.raises(ValueError):
        sample_challenge(t, babybear_mont, 0)
    with pytest.raises(ValueError):
        reinterpret_challenge(raw[:2], babybearx4_mont)


def _fixture(name):
    fx = json.loads((GOLDEN / f"{name}.json").read_text())
    inst = build(fx["recipe"])
    return fx, inst


@pytest.mark.parametrize("name", sorted(RECIPES))
def test_host_verifier_accepts_oracle_proofs(name):
    fx, inst = _fixture(name)
    c = inst.chal
    proof = from_storage(np.array(fx["msgs"], dtype=c.word_dtype), c)
    t0 = W.new_transcript(inst)
    pol = ChallengePolicy(c)
    if inst.mode == "zerocheck":
        v = EqPolyVerifier(ProductSummand(inst.degree), challenges=pol)
        res = v.verify(EqSumClaim(inst.z, inst.claim, inst.n), proof, t0)
    else:
        v = SumcheckVerifier(SumcheckRound(inst.degree, pol))
        res = v.verify(SumClaim(inst.claim, inst.n), proof, t0)
    assert res.ok
    assert res.reduced_claim.point.words.reshape(inst.n, -1).tolist() == fx["point"]
    assert res.reduced_claim.value.words.reshape(-1).tolist() == fx["value"]
    assert res.transcript.stream_state() == fixture_state(fx)


def test_host_verifier_rejects_tampering_and_malformed():
    fx, inst = _fixture("product_bb_n4")
    c = inst.chal
    proof = from_storage(np.array(fx["msgs"], dtype=np.uint32), c)
    t0 = Sha256FieldTranscript.new(inst.label, c)
    v = SumcheckVerifier(SumcheckRound(2, ChallengePolicy(c)))
    assert not v.verify(SumClaim(inst.claim + array(1, c), 4), proof, t0).ok
    bad = proof.at_set((2, 0), proof[2, 0] + array(1, c))
    assert not v.verify(SumClaim(inst.claim, 4), bad, t0).ok
    with pytest.raises(ValueError):
        v.verify(SumClaim(inst.claim, 5), proof, t0)
    with pytest.raises(ValueError):
        SumcheckVerifier(SumcheckRound(3, ChallengePolicy(c))).verify(
            SumClaim(inst.claim, 4), proof, t0)
    with pytest.raises(ValueError):
        SumcheckRound(0, ChallengePolicy(c))
    with pytest.raises(ValueError):
        from zorch_amd.verify import verify
        verify(SumcheckRound(2, ChallengePolicy(c)), inst.claim,
               from_storage(np.zeros((0, 3), np.uint32), c), t0)


def test_field_array_arithmetic_matches_oracle():
    from oracle import fields as OF
    rng = np.random.default_rng(1)
    for f, of in ((babybear_mont, OF.BABYBEAR), (goldilocks, OF.GOLDILOCKS)):
        a = [int(x) for x in rng.integers(0, f.p, 50, dtype=np.uint64)]
        b = [int(x) for x in rng.integers(0, f.p, 50, dtype=np.uint64)]
        A, B = array(a, f), array(b, f)
        assert (A * B).to_ints() == [of.mul(x, y) for x, y in zip(a, b)]
        assert (A + B).to_ints() == [of.add(x, y) for x, y in zip(a, b)]
        assert (A - B).to_ints() == [of.sub(x, y) for x, y in zip(a, b)]
    a = [tuple(int(v) for v in rng.integers(0, babybear_mont.p, 4)) for _ in range(20)]
    b = [tuple(int(v) for v in rng.integers(0, babybear_mont.p, 4)) for _ in range(20)]
    A, B = array(a, babybearx4_mont), array(b, babybearx4_mont)
    assert (A * B).to_ints() == [OF.BABYBEAR4.mul(x, y) for x, y in zip(a, b)]


def test_workload_claims_match_c_oracle():
    import ctypes
    from tests.oracle_glue import oref
    from zorch_amd import workloads as W
    for inst in (W.product_bb(8), W.zerocheck_bb4(8), W.product_gl(8)):
        out = np.zeros(inst.chal.words, dtype=inst.chal.word_dtype)
        z = np.ascontiguousarray(inst.z.words) if inst.z is not None else np.zeros(1, np.uint32)
        tabs = np.ascontiguousarray(inst.tables.words)
        ids = {"babybear_mont": 1, "babybearx4_mont": 2, "goldilocks": 3}
        oref().oref_claim(ids[inst.tables.field.name], ids[inst.chal.name],
                          1 if inst.mode == "zerocheck" else 0, 2, inst.n, tabs.ctypes.data,
                          z.ctypes.data, out.ctypes.data, 2)
        assert out.tolist() == inst.claim.words.reshape(-1).tolist()
