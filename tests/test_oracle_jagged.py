"""The oracle's jagged LogUp-GKR, pinned without the device.

Every check is anchored on the brute-force virtual dense view (each segment
extended to 2^nrv rows with the neutral fraction n = 0, d = 1), as the
reference's own tests are (zorch/logup_gkr/testing/jagged_prover_test.py:
60-175, jagged_verifier_test.py:79-113, jagged_circuit_test.py:96-147):

* every round polynomial equals, at t = 0..3, the LogUp summand summed over
  the virtual hypercube with the earlier challenges bound LSB first (a dense
  fold of the virtual planes and of eq(z, .)), so the closed-form virtual
  mass, the re-padding and the coefficient interpolation are all pinned;
* s(0) + s(1) threads the claims; the final openings are the virtual planes'
  MLEs at the bound point; the final claim is the combine at eq(z, point);
* the transition of uniform even segments equals the dense transition;
  re-padding keeps each segment's fraction sum; slots past a segment's
  folded rows are neutral;
* prove -> verify accepts, the chain closes on the input layer's virtual
  leaf MLEs, and a tampered polynomial or opening is rejected;
* a planted slip in the oracle (the virtual mass at u = 0 without its eq
  factor) fails the brute-force check.
"""

from __future__ import annotations

import random

import pytest

from oracle import fields as OF
from oracle import jagged as OJ
from oracle import logup as OL
from oracle.sumcheck import eq_table, eval_coeffs, eval_mle
from oracle.transcript import OracleTranscript

FIELDS = [(OF.BABYBEAR4, OF.BABYBEAR), (OF.BABYBEAR, OF.BABYBEAR), (OF.GOLDILOCKS, OF.GOLDILOCKS),
          (OF.KOALABEAR, OF.KOALABEAR)]


def _rand(F, rng, n):
    return [F.from_int(rng.randrange(1 << 62)) for _ in range(n)]


def _layer(F, rng, counts):
    w = sum(counts)
    return OJ.JaggedLayer(_rand(F, rng, w), _rand(F, rng, w), _rand(F, rng, w), _rand(F, rng, w),
                          tuple(counts))


def _brute_round_poly(F, lam, planes, z, chals):
    """s(t), t = 0..3, over the virtual hypercube with `chals` bound LSB first."""
    tabs = [eq_table(F, z)] + [list(p) for p in planes]
    for r in chals:
        tabs = [[F.add(t[2 * j], F.mul(r, F.sub(t[2 * j + 1], t[2 * j])))
                 for j in range(len(t) // 2)] for t in tabs]
    out = []
    for t in range(4):
        tt = F.from_int(t)
        acc = F.zero()
        for j in range(len(tabs[0]) // 2):
            v = [F.add(x[2 * j], F.mul(tt, F.sub(x[2 * j + 1], x[2 * j]))) for x in tabs]
            eq, n0, n1, d0, d1 = v
            acc = F.add(acc, OL.logup_combine(F, lam, eq, n0, d1, n1, d0))
        out.append(acc)
    return out


def _virtual_claim(F, lam, planes, z):
    eq = eq_table(F, z)
    n0, n1, d0, d1 = planes
    acc = F.zero()
    for i in range(len(eq)):
        acc = F.add(acc, OL.logup_combine(F, lam, eq[i], n0[i], d1[i], n1[i], d0[i]))
    return acc


def _check_layer(F, T, counts, nrv, seed):
    rng = random.Random(seed)
    layer = _layer(F, rng, counts)
    niv = layer.num_batch_variables
    lam = _rand(F, rng, 1)[0]
    z = _rand(F, rng, niv + nrv)
    planes = OJ.virtual_planes(F, layer, nrv)
    claim = _virtual_claim(F, lam, planes, z)
    t0 = OracleTranscript.new(b"jagged-oracle", T.item_bytes)
    point, _, proof = OJ.prove_jagged_layer(F, T, layer, lam, claim, z, t0)
    chals = point[::-1]
    running = claim
    for k, poly in enumerate(proof.round_polys):
        want = _brute_round_poly(F, lam, planes, z, chals[:k])
        assert [eval_coeffs(F, poly, F.from_int(t)) for t in range(4)] == want, k
        assert F.add(eval_coeffs(F, poly, F.zero()), eval_coeffs(F, poly, F.one())) == running
        running = eval_coeffs(F, poly, chals[k])
    assert proof.claim == claim and proof.lam == lam
    for got, plane in zip((proof.n0, proof.n1, proof.d0, proof.d1), planes):
        assert got == eval_mle(F, plane, point)
    assert running == OL.logup_combine(F, lam, OL.eval_eq(F, z, point), proof.n0, proof.d1,
                                       proof.n1, proof.d0)


@pytest.mark.parametrize("F,T", FIELDS)
@pytest.mark.parametrize("counts,nrv", [((3, 1, 5, 2), 3), ((1, 1, 1, 1), 3), ((7,), 3),
                                        ((2, 4, 1, 3, 8, 6, 5, 1), 4), ((2, 1), 1)])
def test_round_polys_match_the_virtual_hypercube(F, T, counts, nrv):
    _check_layer(F, T, counts, nrv, seed=sum(counts) * 31 + nrv)


def test_planted_virtual_mass_slip_is_caught(monkeypatch):
    F, T = OF.BABYBEAR4, OF.BABYBEAR
    def slipped(F_, ez, eh, es, eq_adj, pad_adj, z, claim):
        one = F_.one()
        corr = F_.sub(pad_adj, es)
        s0 = F_.mul(F_.add(ez, corr), eq_adj)  # the virtual mass at u = 0 without (1 - z)
        sh = F_.mul(F_.mul(F_.add(eh, F_.mul(corr, F_.from_int(4))), F_.inv(F_.from_int(8))),
                    eq_adj)
        b = F_.mul(F_.sub(one, z), F_.inv(F_.sub(one, F_.mul(F_.from_int(2), z))))
        return OJ.interp(F_, [F_.zero(), one, F_.inv(F_.from_int(2)), b],
                         [s0, F_.sub(claim, s0), sh, F_.zero()])

    monkeypatch.setattr(OJ, "round_coeffs", slipped)
    with pytest.raises(AssertionError):
        _check_layer(F, T, (3, 1, 5, 2), 3, seed=5)


def test_uniform_even_segments_match_the_dense_transition():
    F = OF.BABYBEAR4
    rng = random.Random(1)
    layer = _layer(F, rng, (4, 4, 4, 4))
    dense = OL.Layer(layer.n0, layer.n1, layer.d0, layer.d1, 2)
    got = OJ.jagged_layer_transition(F, layer, (2, 2, 2, 2))
    want = OL.layer_transition(F, dense)
    assert (got.n0, got.n1, got.d0, got.d1) == (want.n0, want.n1, want.d0, want.d1)


def _seg_frac_sums(F, layer):
    out, pos = [], 0
    for rc in layer.row_counts:
        acc = F.zero()
        for i in range(pos, pos + rc):
            acc = F.add(acc, F.mul(layer.n0[i], F.inv(layer.d0[i])))
            acc = F.add(acc, F.mul(layer.n1[i], F.inv(layer.d1[i])))
        out.append(acc)
        pos += rc
    return out


def test_odd_segments_and_post_padding_keep_fraction_sums():
    F = OF.BABYBEAR4
    rng = random.Random(2)
    layer = _layer(F, rng, (3, 1, 5, 2))
    nxt = OJ.jagged_layer_transition(F, layer, (4, 1, 4, 2))
    assert _seg_frac_sums(F, nxt) == _seg_frac_sums(F, layer)
    # segment 0 folds 2 rows into an out count of 4: rows 2, 3 are neutral
    assert nxt.n0[2:4] == [F.zero()] * 2 and nxt.d1[2:4] == [F.one()] * 2
    with pytest.raises(ValueError):
        OJ.jagged_layer_transition(F, layer, (1, 1, 3, 1))


@pytest.mark.parametrize("F,T", FIELDS[:3])
def test_chain_roundtrip_closes_on_the_virtual_leaf(F, T):
    rng = random.Random(7)
    first = _layer(F, rng, (3, 1, 5, 2))
    t0 = OracleTranscript.new(b"jagged-chain", T.item_bytes)
    layers, (num, den), proofs, carry, t_p = OJ.prove_jagged_gkr(F, T, first, t0)
    assert len(proofs) == len(layers) - 1
    vcarry, t_v, ok = OJ.verify_jagged_gkr(F, T, num, den, proofs, t0)
    assert ok and vcarry == carry and t_v.stream == t_p.stream
    num_eval, den_eval, point = carry
    nrv = len(point) - first.num_batch_variables - 1
    n0, n1, d0, d1 = OJ.virtual_planes(F, first, nrv)
    assert num_eval == eval_mle(F, OJ.interleave(n0, n1), point)
    assert den_eval == eval_mle(F, OJ.interleave(d0, d1), point)
    bad = [OJ.JaggedLayerProof(p.lam, p.claim, [list(m) for m in p.round_polys], p.point, p.n0,
                               p.n1, p.d0, p.d1) for p in proofs]
    bad[1].round_polys[0][0] = F.add(bad[1].round_polys[0][0], F.one())
    assert not OJ.verify_jagged_gkr(F, T, num, den, bad, t0)[2]
    bad = [OJ.JaggedLayerProof(p.lam, p.claim, p.round_polys, p.point, p.n0, p.n1, p.d0, p.d1)
           for p in proofs]
    bad[0].n0 = F.add(bad[0].n0, F.one())
    assert not OJ.verify_jagged_gkr(F, T, num, den, bad, t0)[2]
