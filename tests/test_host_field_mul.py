"""Host BabyBear^4 products (zorch_amd.fields) against the oracle's quartic
field (oracle/fields.py), on each of the three code paths: one element,
a small array (the sixteen limb products collected by a weight table) and a
large array (limb by limb), with limbs at 0 and p - 1 among the inputs."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import fields as OF
from zorch_amd import fields as Z


@pytest.mark.parametrize("shape", [(), (5,), (3, 4), (Z._BB4_SMALL + 3,)])
def test_bb4_mul_matches_oracle(shape):
    f = Z.as_field("babybearx4_mont")
    rng = np.random.default_rng(len(shape) * 7 + sum(shape))
    a = rng.integers(0, Z.BB_P, size=shape + (4,), dtype=np.uint64)
    b = rng.integers(0, Z.BB_P, size=shape + (4,), dtype=np.uint64)
    a.reshape(-1, 4)[0] = Z.BB_P - 1
    b.reshape(-1, 4)[0, 1:] = 0
    got = Z.c_mul(f, a, b)
    assert got.shape == a.shape and got.dtype == np.uint64
    F = OF.BABYBEAR4
    for x, y, z in zip(a.reshape(-1, 4), b.reshape(-1, 4), got.reshape(-1, 4)):
        want = F.mul(tuple(int(v) for v in x), tuple(int(v) for v in y))
        assert tuple(int(v) for v in z) == tuple(want)


def test_bb4_mul_broadcasts_a_scalar():
    f = Z.as_field("babybearx4_mont")
    rng = np.random.default_rng(9)
    s = rng.integers(0, Z.BB_P, size=(4,), dtype=np.uint64)
    v = rng.integers(0, Z.BB_P, size=(6, 4), dtype=np.uint64)
    got = Z.c_mul(f, s, v)
    for i in range(6):
        assert (got[i] == Z.c_mul(f, s, v[i])).all()
