"""Jagged LogUp-GKR with every layer's sumcheck on the GPU
(``zs_prove_jagged_layer``), against the oracle (oracle/jagged.py).

* ``prove_jagged_layer``: the round polynomials, the sampled challenges, the
  openings and the transcript state equal the oracle's bit for bit, for three
  field pairs, over layouts with odd, saturated and even segments, one
  segment (no batch variables), a one-row-variable layer, a wide random
  layout and a deep one whose first round runs multi-block.
* ``JaggedGkrLayerRound`` over a jagged pyramid: every layer proof and the
  final carry equal the oracle chain's; the jagged verifier accepts and
  reaches the same carry.
* A layer of 64 segments with up to 2^12 rows each (about 2^17 live rows,
  2^18 virtual): the verifier accepts and the openings are the virtual
  planes' MLEs at the bound point, computed on the host.
"""

from __future__ import annotations

import random
import zlib

import numpy as np
import pytest

from oracle import fields as OF
from oracle import jagged as OJ
from oracle.transcript import OracleTranscript
from zorch_amd.challenge import ChallengePolicy
from zorch_amd.fields import FieldArray, as_field, concatenate, from_storage, fsum, ones, zeros
from zorch_amd.logup_gkr import (JaggedGkrLayer, JaggedGkrLayerRound, JaggedVerifierLayerRound,
                                 RoundWidthCaps, bind_output, build_jagged_pyramid,
                                 extract_jagged_outputs, jagged_device_pyramid,
                                 prove_jagged_layer)
from zorch_amd.logup_gkr.circuit import _interleave
from zorch_amd.logup_gkr.prover import eval_mle
from zorch_amd.round import prove_rounds, verify_rounds
from zorch_amd.sha256_field_transcript import Sha256FieldTranscript

pytestmark = pytest.mark.gpu

PAIRS = [("babybearx4_mont", "babybear_mont"), ("babybear_mont", "babybear_mont"),
         ("goldilocks", "goldilocks")]
OFIELD = {"babybear_mont": OF.BABYBEAR, "babybearx4_mont": OF.BABYBEAR4,
          "goldilocks": OF.GOLDILOCKS}


def _fa(F, field, vals) -> FieldArray:
    f = as_field(field)
    words = np.array([F.to_words(v) for v in vals], dtype=f.word_dtype).reshape(len(vals), f.words)
    return from_storage(words, f)


def _words(fa: FieldArray) -> list:
    return fa.normalized().words.reshape(-1).tolist()


def _owords(F, vals) -> list:
    return [int(w) for v in vals for w in F.to_words(v)]


def _caps(counts, nrv):
    width = sum(rc + rc % 2 for rc in counts)
    return RoundWidthCaps(elements=max(width + (-width % 4), sum(counts) + (-sum(counts) % 4)),
                          eq_row=1 << nrv, interaction=max(4, len(counts)))


def _layers(F, field, rng, counts):
    w = sum(counts)
    vals = [[F.from_int(rng.randrange(1 << 62)) for _ in range(w)] for _ in range(4)]
    return (OJ.JaggedLayer(*vals, tuple(counts)),
            JaggedGkrLayer(*(_fa(F, field, v) for v in vals), tuple(counts)))


# Rounds of more than 2048 pairs run as multi-block launches, the rest in one
# block: "deep" (8 segments of 512..1024 rows) takes both.
LAYOUTS = [((3, 1, 5, 2), 3), ((1, 1, 1, 1), 3), ((7,), 3), ((2, 4, 1, 3, 8, 6, 5, 1), 4),
           ((2, 1), 1), ("wide", 7), ("deep", 10)]


def _counts(spec, nrv, rng):
    if spec == "wide":
        return tuple(rng.randrange(1, (1 << nrv) + 1) for _ in range(32))
    if spec == "deep":
        return tuple(rng.randrange(1 << (nrv - 1), (1 << nrv) + 1) for _ in range(8))
    return spec


@pytest.mark.parametrize("cf,tf", PAIRS)
@pytest.mark.parametrize("spec,nrv", LAYOUTS)
def test_layer_matches_oracle(cf, tf, spec, nrv):
    F, T = OFIELD[cf], OFIELD[tf]
    rng = random.Random(zlib.crc32(repr((cf, spec, nrv)).encode()))
    counts = _counts(spec, nrv, rng)
    olayer, dlayer = _layers(F, cf, rng, counts)
    niv = olayer.num_batch_variables
    lam = F.from_int(rng.randrange(1 << 62))
    claim = F.from_int(rng.randrange(1 << 62))
    z = [F.from_int(rng.randrange(1 << 62)) for _ in range(niv + nrv)]
    label = b"jagged-gpu"
    opoint, t_or, oproof = OJ.prove_jagged_layer(F, T, olayer, lam, claim, z,
                                                 OracleTranscript.new(label, T.item_bytes))
    point, t_dev, proof = prove_jagged_layer(
        dlayer, _fa(F, cf, [lam])[0], _fa(F, cf, [claim])[0], _fa(F, cf, z),
        Sha256FieldTranscript.new(label, tf), challenges=ChallengePolicy(cf),
        caps=_caps(counts, nrv))
    assert _words(proof.round_polys) == _owords(F, [c for p in oproof.round_polys for c in p])
    assert _words(point) == _owords(F, opoint)
    assert _words(proof.numerator_0) == _owords(F, [oproof.n0])
    assert _words(proof.numerator_1) == _owords(F, [oproof.n1])
    assert _words(proof.denominator_0) == _owords(F, [oproof.d0])
    assert _words(proof.denominator_1) == _owords(F, [oproof.d1])
    assert t_dev.stream_state() == tuple(t_or.state())


@pytest.mark.parametrize("cf", ["babybearx4_mont", "goldilocks"])
def test_device_pyramid_equals_host_circuit(cf):
    """zs_jagged_transition layer by layer against the host circuit, over
    odd, saturated and over-padded segments."""
    F = OFIELD[cf]
    rng = random.Random(12)
    counts = tuple(rng.randrange(1, 40) for _ in range(16))
    _, dlayer = _layers(F, cf, rng, counts)
    sched = OJ.jagged_fold_schedules(counts)
    host = build_jagged_pyramid(dlayer, sched)
    dev = jagged_device_pyramid(dlayer, sched, cf)
    for h, d in zip(host, dev):
        back = d.to_host()
        assert back.row_counts == h.row_counts
        for a, b in ((back.numerator_0, h.numerator_0), (back.numerator_1, h.numerator_1),
                     (back.denominator_0, h.denominator_0),
                     (back.denominator_1, h.denominator_1)):
            assert _words(a) == _words(b)


@pytest.mark.parametrize("cf,tf", PAIRS)
@pytest.mark.parametrize("on_device", [False, True])
def test_chain_matches_oracle_and_verifies(cf, tf, on_device):
    F, T = OFIELD[cf], OFIELD[tf]
    rng = random.Random(11)
    counts = (3, 1, 5, 2, 6, 1, 8, 4)
    olayer, dlayer = _layers(F, cf, rng, counts)
    sched = OJ.jagged_fold_schedules(counts)
    ch = ChallengePolicy(cf)
    label = b"jagged-chain"
    layers = (jagged_device_pyramid(dlayer, sched, cf) if on_device
              else build_jagged_pyramid(dlayer, sched))
    floor = layers[-1].to_host() if on_device else layers[-1]
    output = extract_jagged_outputs(floor)
    caps = _caps(counts, len(layers) - 1)
    carry, t = bind_output(output, Sha256FieldTranscript.new(label, tf), ch)
    final, t_dev, proofs = prove_rounds(
        [JaggedGkrLayerRound(layer, ch, caps=caps) for layer in reversed(layers[:-1])], carry, t)
    _, (onum, oden), oproofs, ocarry, t_or = OJ.prove_jagged_gkr(
        F, T, olayer, OracleTranscript.new(label, T.item_bytes), sched)
    assert _words(output.numerator) == _owords(F, onum)
    assert _words(output.denominator) == _owords(F, oden)
    for p, q in zip(proofs, oproofs):
        assert _words(p.round_polys) == _owords(F, [c for m in q.round_polys for c in m])
        assert _words(p.point) == _owords(F, q.point)
        assert _words(p.lam) == _owords(F, [q.lam]) and _words(p.claim) == _owords(F, [q.claim])
    assert _words(final[0]) == _owords(F, [ocarry[0]])
    assert _words(final[1]) == _owords(F, [ocarry[1]])
    assert _words(final[2]) == _owords(F, ocarry[2])
    assert t_dev.stream_state() == tuple(t_or.state())
    carry, t = bind_output(output, Sha256FieldTranscript.new(label, tf), ch)
    vfinal, t_ver, ok = verify_rounds([JaggedVerifierLayerRound(ch) for _ in proofs], carry,
                                      proofs, t)
    assert ok
    for a, b in zip(vfinal, final):
        assert _words(a) == _words(b)
    assert t_ver.stream_state() == t_dev.stream_state()


def _virtual(layer: JaggedGkrLayer, nrv: int):
    f = layer.numerator_0.field
    rows = 1 << nrv
    out = []
    for a, neutral in ((layer.numerator_0, zeros), (layer.numerator_1, zeros),
                       (layer.denominator_0, ones), (layer.denominator_1, ones)):
        parts, pos = [], 0
        for rc in layer.row_counts:
            parts.append(a[pos:pos + rc])
            if rows > rc:
                parts.append(neutral((rows - rc,), f))
            pos += rc
        out.append(concatenate(parts))
    return out


@pytest.mark.parametrize("cf,tf", PAIRS[:2])
def test_wide_layer_openings_are_virtual_mles(cf, tf):
    f = as_field(cf)
    rng = np.random.default_rng(5)
    nrv, nseg = 12, 64
    counts = tuple(int(c) for c in rng.integers(1, (1 << nrv) + 1, size=nseg))
    w = sum(counts)
    planes = [FieldArray(rng.integers(0, OF.BABYBEAR.p, size=(w, f.words), dtype=np.uint64)
                         .astype(f.word_dtype), f).normalized() for _ in range(4)]
    layer = JaggedGkrLayer(*planes, counts)
    niv = layer.num_batch_variables
    z = FieldArray(rng.integers(0, OF.BABYBEAR.p, size=(niv + nrv, f.words), dtype=np.uint64)
                   .astype(f.word_dtype), f).normalized()
    lam = FieldArray(rng.integers(0, OF.BABYBEAR.p, size=(f.words,), dtype=np.uint64)
                     .astype(f.word_dtype), f).normalized()
    vn0, vn1, vd0, vd1 = _virtual(layer, nrv)
    from zorch_amd.logup_gkr.prover import expand_eq_to_hypercube, logup_combine
    eq = expand_eq_to_hypercube(z, ones((), f))
    claim = fsum(logup_combine(lam, eq, vn0, vd1, vn1, vd0))
    ch = ChallengePolicy(cf)
    point, _, proof = prove_jagged_layer(layer, lam, claim, z,
                                         Sha256FieldTranscript.new(b"wide", tf), challenges=ch,
                                         caps=_caps(counts, nrv))
    for got, plane in ((proof.numerator_0, vn0), (proof.numerator_1, vn1),
                       (proof.denominator_0, vd0), (proof.denominator_1, vd1)):
        assert _words(got) == _words(eval_mle(plane, point))
    # the coefficient-form replay accepts and ends on the combine at the point
    from zorch_amd.logup_gkr.verifier import eval_eq
    from zorch_amd.sumcheck.verifier import CoeffsSumcheckRound
    from zorch_amd.verify import verify
    vpoint, final, _, ok = verify(CoeffsSumcheckRound(3, ch), claim, proof.round_polys,
                                  Sha256FieldTranscript.new(b"wide", tf))
    assert ok
    assert _words(vpoint.normalized()[::-1]) == _words(point)
    want = logup_combine(lam, eval_eq(z, point), proof.numerator_0, proof.denominator_1,
                         proof.numerator_1, proof.denominator_0)
    assert _words(final) == _words(want)
