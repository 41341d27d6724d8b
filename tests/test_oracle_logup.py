"""The oracle's dense LogUp-GKR, pinned without the device.

* The round polynomial equals the combine summed over the table halves folded
  at each node (the reference's naive-cubic test, zorch/logup_gkr/testing/
  prover_test.py:75-87), and s(0) + s(1) is the hypercube sum (:89-94).
* The circuit's output fractions add up to the input's (each transition
  preserves n0/d0 + n1/d1, checked with field inverses).
* The circuit's outputs hash to the reference's golden vectors
  (zorch/logup_gkr/testing/golden_vector_test.py:21-41): KoalaBear layers
  drawn as the reference's ``random_first_layer(42, ...)`` draws them
  (numpy PCG64 integers in [0, 2^30), seeds 42..45;
  zorch/testkit/random_field.py:46-47), outputs as canonical u32 words.
* prove -> verify accepts; the final carry is the input layer's interleaved
  numerator and denominator MLEs at the final point (the closing check of
  the reference's roundtrip, verifier.py:12-15); a tampered round polynomial
  or opening is rejected.
"""

from __future__ import annotations

import hashlib
import random

import numpy as np
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


# (batch variables, row variables) -> SHA-256 of the output MLEs' canonical
# u32 words, numerator then denominator (golden_vector_test.py:21-27).
_GOLDENS = {
    (2, 3): "09f0bbf7f0cc0223e004a83f43c58f656463d2c157f32d627a182602da27fe93",
    (2, 4): "a899f692b4ded58fd8218b37c29515ca3f3086588f5ddc6fa16e46e5e7ddb32b",
    (2, 5): "a1568612f53abfbaed673c49fdfbfdec4fd321a04e7104a62cd0897a9a47413b",
    (3, 4): "918a9179e246a562ee180d2e00f84febddfb53f3ca566de619800c5d06c535c2",
}


def _draw(seed, n):
    return [int(v) for v in np.random.default_rng(seed).integers(0, 1 << 30, size=(n,),
                                                                   dtype=np.int64)]


@pytest.mark.parametrize("nb,nr", sorted(_GOLDENS))
def test_circuit_outputs_match_reference_golden_vectors(nb, nr):
    F = OF.KOALABEAR
    w = 1 << (nb + nr)
    first = OL.Layer(_draw(42, w), _draw(43, w), _draw(44, w), _draw(45, w), nb)
    num, den = OL.extract_outputs(F, OL.build_pyramid(F, first)[-1])
    words = np.array(list(num) + list(den), dtype="<u4").tobytes()
    assert hashlib.sha256(words).hexdigest() == _GOLDENS[(nb, nr)]
    # A planted slip in the transition (the cross term's second product
    # dropped) moves the hash.
    bad = OL.Layer(first.n0, first.n1, first.d0, first.d1, nb)
    layers = [bad]
    while layers[-1].num_row_variables > 0:
        l = layers[-1]
        h = len(l.n0) // 2
        ev = lambda a: [a[2 * i] for i in range(h)]
        od = lambda a: [a[2 * i + 1] for i in range(h)]
        n = [F.mul(a, b) for a, b in zip(l.n0, l.d1)]
        d = [F.mul(a, b) for a, b in zip(l.d0, l.d1)]
        layers.append(OL.Layer(ev(n), od(n), ev(d), od(d), nb))
    num, den = OL.extract_outputs(F, layers[-1])
    words = np.array(list(num) + list(den), dtype="<u4").tobytes()
    assert hashlib.sha256(words).hexdigest() != _GOLDENS[(nb, nr)]
