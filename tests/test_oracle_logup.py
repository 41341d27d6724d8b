"""The oracle's dense LogUp-GKR, pinned without the device.

* The round polynomial equals the combine summed over the table halves folded
  at each node (the reference's naive-cubic test, zorch/logup_gkr/testing/
  prover_test.py:75-87), and s(0) + s(1) is the hypercube sum (:89-94).
* The circuit's output fractions add up to the input's (each transition
  preserves n0/d0 + n1/d1, checked with field inverses).
* prove -> verify accepts; the final carry is the input layer's interleaved
  numerator and denominator MLEs at the final point (the closing check of
  the reference's roundtrip, verifier.py:12-15); a tampered round polynomial
  or opening is rejected.
"""

from __future__ import annotations

import random

import pytest

from oracle import fields as OF
from oracle import logup as OL
from oracle.sumcheck import eval_mle
from oracle.transcript import OracleTranscript


def _rand(F, rng, n):
    return [F.from_int(rng.randrange(1 << 62)) for _ in range(n)]


def _layer(F, rng, nb, nr):
    w = 1 << (nb + nr)
    return OL.Layer(_rand(F, rng, w), _rand(F, rng, w), _rand(F, rng, w), _rand(F, rng, w), nb)


FIELDS = [(OF.BABYBEAR4, OF.BABYBEAR), (OF.BABYBEAR, OF.BABYBEAR), (OF.GOLDILOCKS, OF.GOLDILOCKS)]


@pytest.mark.parametrize("F,T", FIELDS)
def test_round_poly_matches_naive_cubic(F, T):
    rng = random.Random(20)
    state = [_rand(F, rng, 8) for _ in range(5)]
    lam = F.from_int(7)
    msg = OL.logup_round_poly(F, lam, state)
    assert len(msg) == 4
    for u in range(4):
        uf = F.from_int(u)
        acc = F.zero()
        for j in range(4):
            v = [F.add(x[j], F.mul(uf, F.sub(x[j + 4], x[j]))) for x in state]
            acc = F.add(acc, OL.logup_combine(F, lam, *v))
        assert msg[u] == acc
    total = F.zero()
    for j in range(8):
        total = F.add(total, OL.logup_combine(F, lam, *[x[j] for x in state]))
    assert F.add(msg[0], msg[1]) == total


def _frac_sum(F, nums, dens):
    acc = F.zero()
    for n, d in zip(nums, dens):
        acc = F.add(acc, F.mul(n, F.inv(d)))
    return acc


@pytest.mark.parametrize("F,T", FIELDS[:2])
def test_pyramid_preserves_the_fraction_sum(F, T):
    rng = random.Random(5)
    first = _layer(F, rng, 1, 3)
    layers = OL.build_pyramid(F, first)
    assert [l.num_row_variables for l in layers] == [3, 2, 1, 0]
    want = F.add(_frac_sum(F, first.n0, first.d0), _frac_sum(F, first.n1, first.d1))
    for layer in layers:
        got = F.add(_frac_sum(F, layer.n0, layer.d0), _frac_sum(F, layer.n1, layer.d1))
        assert got == want
    num, den = OL.extract_outputs(F, layers[-1])
    assert len(num) == 4 and _frac_sum(F, num, den) == want
    with pytest.raises(ValueError):
        OL.layer_transition(F, layers[-1])


def _interleave(a, b):
    out = []
    for x, y in zip(a, b):
        out += [x, y]
    return out


@pytest.mark.parametrize("F,T", FIELDS)
@pytest.mark.parametrize("nb,nr", [(0, 3), (1, 2), (2, 3)])
def test_prove_verify_roundtrip_and_leaf_claim(F, T, nb, nr):
    rng = random.Random(100 * nb + nr)
    first = _layer(F, rng, nb, nr)
    t0 = OracleTranscript.new(b"logup-oracle", T.item_bytes)
    layers, (num, den), proofs, carry, t_p = OL.prove_gkr(F, T, first, t0)
    assert len(proofs) == len(layers) - 1
    carry_v, t_v, ok = OL.verify_gkr(F, T, num, den, proofs, t0)
    assert ok and carry_v == carry and t_v.stream == t_p.stream
    num_eval, den_eval, point = carry
    assert len(point) == nb + nr + 1
    assert num_eval == eval_mle(F, _interleave(first.n0, first.n1), point)
    assert den_eval == eval_mle(F, _interleave(first.d0, first.d1), point)


def test_tampering_is_rejected():
    F, T = OF.BABYBEAR4, OF.BABYBEAR
    rng = random.Random(9)
    first = _layer(F, rng, 1, 2)
    t0 = OracleTranscript.new(b"logup-oracle", T.item_bytes)
    _, (num, den), proofs, _, _ = OL.prove_gkr(F, T, first, t0)
    bad = [OL.LayerProof([list(m) for m in p.round_polys], list(p.point), p.n0, p.n1, p.d0, p.d1)
           for p in proofs]
    bad[-1].round_polys[0][2] = F.add(bad[-1].round_polys[0][2], F.one())
    assert not OL.verify_gkr(F, T, num, den, bad, t0)[2]
    bad = [OL.LayerProof(p.round_polys, p.point, p.n0, p.n1, p.d0, p.d1) for p in proofs]
    bad[0].n1 = F.add(bad[0].n1, F.one())
    assert not OL.verify_gkr(F, T, num, den, bad, t0)[2]
