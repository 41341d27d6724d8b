"""Dense LogUp-GKR with every layer's sumcheck on the GPU, against the oracle
(oracle/logup.py).

* ``prove_gkr`` (the layer rounds' ``zs_prove`` of ``ZS_MODE_LOGUP`` stages):
  every layer's round polynomials, bound point and openings, the final carry
  and the transcript state equal the oracle's chain bit for bit, for three
  field pairs and several (batch, row) shapes; the verifier chain accepts.
* ``LogupSumcheckRound`` under ``fold_rounds`` (``zs_state_round``): the
  messages and challenges equal the oracle's round by round on an arbitrary
  five-table state, and the folded tables read back equal the oracle's.
* ``LogUpGkrProver`` / ``LogUpGkrVerifier`` (the stage pair): the input-layer
  claim and the layer proofs equal the oracle chain's and the verifier
  derives the same claim.
* At 2^20 entries a layer (beyond the oracle's reach): the verifier accepts
  and the final carry is the input layer's MLEs at the final point, computed
  on the host (the reference roundtrip's closing check, verifier.py:12-15).
"""

from __future__ import annotations

import random

import numpy as np
import pytest

from oracle import fields as OF
from oracle import logup as OL
from oracle.sumcheck import fs_hop
from oracle.transcript import OracleTranscript
from zorch_amd.challenge import ChallengePolicy
from zorch_amd.fields import FieldArray, as_field, encode, from_storage, stack
from zorch_amd.logup_gkr import (GkrLayer, LogUpGkrProver, LogUpGkrVerifier, LogUpOutputClaim,
                                 LogupSumcheckRound, extract_outputs, build_pyramid, prove_gkr,
                                 verify_gkr)
from zorch_amd.logup_gkr.circuit import _interleave
from zorch_amd.logup_gkr.prover import eval_mle
from zorch_amd.prove import fold_rounds
from zorch_amd.sha256_field_transcript import Sha256FieldTranscript

pytestmark = pytest.mark.gpu

# (challenge / layer field, transcript field)
PAIRS = [("babybearx4_mont", "babybear_mont"), ("babybear_mont", "babybear_mont"),
         ("goldilocks", "goldilocks")]
OFIELD = {"babybear_mont": OF.BABYBEAR, "babybearx4_mont": OF.BABYBEAR4,
          "goldilocks": OF.GOLDILOCKS}


def _to_fa(F, field, vals) -> FieldArray:
    f = as_field(field)
    words = np.array([F.to_words(v) for v in vals], dtype=f.word_dtype).reshape(len(vals), f.words)
    return from_storage(words, f)


def _words(fa: FieldArray) -> list:
    return fa.normalized().words.reshape(-1).tolist()


def _owords(F, vals) -> list:
    out = []
    for v in vals:
        out += [int(w) for w in F.to_words(v)]
    return out


def _layers(F, field, rng, nb, nr):
    w = 1 << (nb + nr)
    vals = [[F.from_int(rng.randrange(1 << 62)) for _ in range(w)] for _ in range(4)]
    olayer = OL.Layer(vals[0], vals[1], vals[2], vals[3], nb)
    dlayer = GkrLayer(*(_to_fa(F, field, v) for v in vals), nb)
    return olayer, dlayer


@pytest.mark.parametrize("cf,tf", PAIRS)
@pytest.mark.parametrize("nb,nr", [(0, 1), (1, 3), (2, 5), (0, 9)])
def test_prove_gkr_matches_oracle(cf, tf, nb, nr):
    F, T = OFIELD[cf], OFIELD[tf]
    rng = random.Random(7 * nb + nr)
    olayer, dlayer = _layers(F, cf, rng, nb, nr)
    label = b"logup-gpu-%d-%d" % (nb, nr)
    ch = ChallengePolicy(cf)
    _, output, proofs, carry, t_dev = prove_gkr(dlayer, Sha256FieldTranscript.new(label, tf), ch)
    _, (onum, oden), oproofs, ocarry, t_or = OL.prove_gkr(
        F, T, olayer, OracleTranscript.new(label, T.item_bytes))
    assert _words(output.numerator) == _owords(F, onum)
    assert _words(output.denominator) == _owords(F, oden)
    assert len(proofs) == len(oproofs)
    for p, q in zip(proofs, oproofs):
        assert _words(p.round_polys) == _owords(F, [v for m in q.round_polys for v in m])
        assert _words(p.point) == _owords(F, q.point)
        assert _words(stack([p.numerator_0, p.numerator_1, p.denominator_0, p.denominator_1])) \
            == _owords(F, [q.n0, q.n1, q.d0, q.d1])
    assert _words(carry[0]) == _owords(F, [ocarry[0]])
    assert _words(carry[1]) == _owords(F, [ocarry[1]])
    assert _words(carry[2]) == _owords(F, ocarry[2])
    assert t_dev.stream_state() == tuple(t_or.state())
    vcarry, t_ver, ok = verify_gkr(output, proofs, Sha256FieldTranscript.new(label, tf), ch)
    assert ok
    assert t_ver.stream_state() == t_dev.stream_state()


@pytest.mark.parametrize("cf", ["babybearx4_mont", "goldilocks"])
def test_device_pyramid_equals_host_circuit(cf):
    """zs_logup_transition layer by layer against the host circuit, and
    zs_logup_eq against the host eq expansion."""
    from zorch_amd.fields import ones
    from zorch_amd.logup_gkr import build_pyramid
    from zorch_amd.logup_gkr.prover import device_pyramid, expand_eq_to_hypercube
    F = OFIELD[cf]
    rng = random.Random(21)
    _, dlayer = _layers(F, cf, rng, 2, 7)
    host = build_pyramid(dlayer)
    dev = device_pyramid(dlayer, cf)
    assert len(dev) == len(host)
    for h, d in zip(host, dev):
        back = d.to_host()
        for a, b in ((back.numerator_0, h.numerator_0), (back.numerator_1, h.numerator_1),
                     (back.denominator_0, h.denominator_0), (back.denominator_1, h.denominator_1)):
            assert _words(a) == _words(b)
    pt = _to_fa(F, cf, [F.from_int(rng.randrange(1 << 62)) for _ in range(9)])
    dev[0].set_eq(pt)
    assert _words(dev[0].tables.to_host()[0]) == _words(expand_eq_to_hypercube(pt, ones((), pt.field)))


def test_multi_block_round_matches_oracle():
    """A layer of 2^13 entries: round 0 (4096 pairs) runs as a multi-block
    launch, the later rounds in one block."""
    cf, tf = PAIRS[0]
    F, T = OFIELD[cf], OFIELD[tf]
    rng = random.Random(13)
    olayer, dlayer = _layers(F, cf, rng, 0, 13)
    ch = ChallengePolicy(cf)
    _, output, proofs, carry, t_dev = prove_gkr(dlayer, Sha256FieldTranscript.new(b"mb", tf), ch)
    _, _, oproofs, ocarry, t_or = OL.prove_gkr(F, T, olayer, OracleTranscript.new(b"mb", T.item_bytes))
    for p, q in zip(proofs, oproofs):
        assert _words(p.round_polys) == _owords(F, [v for m in q.round_polys for v in m])
    assert _words(carry[2]) == _owords(F, ocarry[2])
    assert t_dev.stream_state() == tuple(t_or.state())


@pytest.mark.parametrize("cf,tf", PAIRS)
def test_stage_pair_matches_oracle(cf, tf):
    F, T = OFIELD[cf], OFIELD[tf]
    rng = random.Random(77)
    nb, nr = 1, 4
    olayer, dlayer = _layers(F, cf, rng, nb, nr)
    ch = ChallengePolicy(cf)
    out = extract_outputs(build_pyramid(dlayer)[-1])
    claim = LogUpOutputClaim(out, nr)
    res = LogUpGkrProver(ch).prove(claim, dlayer, Sha256FieldTranscript.new(b"stage", tf))
    _, _, oproofs, ocarry, t_or = OL.prove_gkr(F, T, olayer, OracleTranscript.new(b"stage", T.item_bytes))
    assert len(res.reduction_proof.layers) == nr == len(oproofs)
    for p, q in zip(res.reduction_proof.layers, oproofs):
        assert _words(p.round_polys) == _owords(F, [v for m in q.round_polys for v in m])
    ic = res.reduced_claim
    assert _words(ic.numerator) == _owords(F, [ocarry[0]])
    assert _words(ic.denominator) == _owords(F, [ocarry[1]])
    assert _words(ic.point) == _owords(F, ocarry[2])
    assert res.transcript.stream_state() == tuple(t_or.state())
    v = LogUpGkrVerifier(ch).verify(claim, res.reduction_proof,
                                    Sha256FieldTranscript.new(b"stage", tf))
    assert v.ok
    assert _words(v.reduced_claim.numerator) == _words(ic.numerator)
    assert _words(v.reduced_claim.point) == _words(ic.point)


@pytest.mark.parametrize("cf,tf", PAIRS)
def test_logup_round_under_fold_rounds_matches_oracle(cf, tf):
    F, T = OFIELD[cf], OFIELD[tf]
    rng = random.Random(31)
    n = 6
    state = [[F.from_int(rng.randrange(1 << 62)) for _ in range(1 << n)] for _ in range(5)]
    lam = F.from_int(rng.randrange(1 << 62))
    rnd = LogupSumcheckRound(_to_fa(F, cf, [lam])[0], ChallengePolicy(cf))
    dstate = stack([_to_fa(F, cf, t) for t in state])
    assert _words(rnd._round_poly(dstate)) == _owords(F, OL.logup_round_poly(F, lam, state))
    final, t_dev, msgs = fold_rounds(rnd, dstate, Sha256FieldTranscript.new(b"rounds", tf), n)
    t_or = OracleTranscript.new(b"rounds", T.item_bytes)
    ost = state
    for i in range(n):
        omsg = OL.logup_round_poly(F, lam, ost)
        t_or, _, r = fs_hop(t_or, F, T, omsg)
        assert _words(msgs[i].round_poly) == _owords(F, omsg)
        assert _words(msgs[i].challenge) == _owords(F, [r])
        h = len(ost[0]) // 2
        ost = [[F.add(t[j], F.mul(r, F.sub(t[j + h], t[j]))) for j in range(h)] for t in ost]
    assert t_dev.stream_state() == tuple(t_or.state())
    assert _words(final.read_tables()) == _owords(F, [t[0] for t in ost])


@pytest.mark.parametrize("cf,tf", PAIRS[:2])
def test_large_layer_roundtrip_closes_on_the_leaf(cf, tf):
    f = as_field(cf)
    rng = np.random.default_rng(3)
    nb, nr = 2, 18
    w = 1 << (nb + nr)
    vals = [rng.integers(0, OF.BABYBEAR.p, size=(w, f.words), dtype=np.uint64) for _ in range(4)]
    mles = [FieldArray(encode(f, v), f) for v in vals]
    layer = GkrLayer(*mles, nb)
    ch = ChallengePolicy(cf)
    _, output, proofs, carry, _ = prove_gkr(layer, Sha256FieldTranscript.new(b"big", tf), ch)
    vcarry, _, ok = verify_gkr(output, proofs, Sha256FieldTranscript.new(b"big", tf), ch)
    assert ok
    num_eval, den_eval, point = carry
    assert point.shape[0] == nb + nr + 1
    assert _words(num_eval) == _words(eval_mle(_interleave(mles[0], mles[1]), point))
    assert _words(den_eval) == _words(eval_mle(_interleave(mles[2], mles[3]), point))
    assert _words(vcarry[0]) == _words(num_eval)
