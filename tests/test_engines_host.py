"""Structural checks of the engine entry points (no GPU): malformed requests
are rejected with ZS_EINVAL before any device work, as the reference raises
ValueError on them (sqrt_space.py, univariate_skip.py:172-174,
small_value.py:164-168), and the host helpers of the skip's subgroup agree
with the oracle.
"""

from __future__ import annotations

import ctypes
import random

import numpy as np
import pytest

from oracle import engines as OE
from oracle import fields as OF
from zorch_amd import _lib
from zorch_amd.fields import array, babybear_mont, babybearx4_mont, goldilocks
from zorch_amd.runtime import message_width
from zorch_amd.sha256_field_transcript import Sha256FieldTranscript
from zorch_amd.sumcheck.domain import (UnivariateSkipDomain, subgroup_evals, subgroup_sum,
                                       subgroup_to_coeffs)


def _args(k=2, n=6, mode=_lib.ZS_MODE_PRODUCT, tf=_lib.ZS_F_BABYBEAR, cf=_lib.ZS_F_BABYBEAR4,
          refold=0, wire=None):
    keep = np.zeros(64, np.uint32)
    a = _lib.ProveArgsC(table_field=tf, chal_field=cf, mode=mode, degree=k, num_tables=k,
                        rounds=n, tables=keep.ctypes.data, tables_on_device=0,
                        refold_rounds=refold,
                        eq_point=keep.ctypes.data if mode == _lib.ZS_MODE_EQ else None,
                        claim=keep.ctypes.data, wire=wire, eq_mass=None)
    return a, keep


def test_refold_rounds_bounds():
    a, keep = _args(n=6, refold=3)
    assert message_width(a) == 3
    a, keep = _args(n=6, refold=4)
    with pytest.raises(ValueError, match="refold_rounds must be at most"):
        message_width(a)
    a, keep = _args(n=6, refold=1, mode=_lib.ZS_MODE_EQ)
    with pytest.raises(ValueError, match="product summand"):
        message_width(a)


def _skip(args, skip, field=babybear_mont):
    t = Sha256FieldTranscript.new(b"x", field)
    st = t.state.copy()
    proof = _lib.ProofC()
    msg0 = np.zeros(512, np.uint64)
    rc = _lib.lib().zs_prove_univariate_skip(None, ctypes.byref(args), skip, ctypes.byref(st),
                                             msg0.ctypes.data, ctypes.byref(proof))
    return rc, _lib.lib().zs_last_error().decode()


def test_univariate_skip_structural_errors():
    a, keep = _args(n=6)
    assert _skip(a, 0)[0] == _lib.ZS_EINVAL and "skip_rounds must be in" in _skip(a, 0)[1]
    assert "skip_rounds must be in" in _skip(a, 7)[1]
    assert "null context" in _skip(a, 2)[1]  # well formed: only the device is missing
    a, keep = _args(n=6, mode=_lib.ZS_MODE_EQ)
    assert "product summand" in _skip(a, 2)[1]
    a, keep = _args(n=6, tf=_lib.ZS_F_BABYBEAR4, cf=_lib.ZS_F_BABYBEAR4)
    assert "base-field work" in _skip(a, 2, babybearx4_mont)[1]
    a, keep = _args(n=6)
    assert "transcript must be over the tables' field" in _skip(a, 2, babybearx4_mont)[1]
    a, keep = _args(n=6, refold=1)
    assert "refold_rounds must be 0" in _skip(a, 2)[1]


def _sv(args, l0, field=babybearx4_mont):
    t = Sha256FieldTranscript.new(b"x", field)
    st = t.state.copy()
    proof = _lib.ProofC()
    rc = _lib.lib().zs_prove_small_value(None, ctypes.byref(args), l0, ctypes.byref(st),
                                         ctypes.byref(proof))
    return rc, _lib.lib().zs_last_error().decode()


def test_small_value_structural_errors():
    eq = _lib.ZS_MODE_EQ
    a, keep = _args(n=6, mode=eq)
    assert "l_0 must be in [1, 3]" in _sv(a, 0)[1]
    assert "l_0 must be in [1, 3]" in _sv(a, 4)[1]
    assert "null context" in _sv(a, 3)[1]
    assert "one transcript word" in _sv(a, 2, babybear_mont)[1]
    a, keep = _args(n=6)
    assert "eq-weighted" in _sv(a, 2)[1]
    a, keep = _args(k=1, n=6, mode=eq)
    assert "degree >= 2" in _sv(a, 2)[1]
    a, keep = _args(n=1, mode=eq)
    assert "at least 2 rounds" in _sv(a, 1)[1]
    w = _lib.WireC(kind=_lib.ZS_WIRE_COEFFS)
    a, keep = _args(n=6, mode=eq, wire=ctypes.pointer(w))
    assert "wire must be null" in _sv(a, 2)[1]


@pytest.mark.parametrize("field,of", [(babybear_mont, OF.BABYBEAR), (goldilocks, OF.GOLDILOCKS)])
def test_host_subgroup_helpers_match_oracle(field, of):
    rng = random.Random(3)
    for skip in (1, 2, 3):
        vals = [rng.randrange(of.p) for _ in range(1 << skip)]
        coeffs = subgroup_to_coeffs(array(vals, field))
        assert coeffs.to_ints() == OE.subgroup_to_coeffs(of, vals)
        ev = subgroup_evals(coeffs, 4 << skip)
        assert ev.to_ints() == OE.subgroup_evals(of, OE.subgroup_to_coeffs(of, vals), 4 << skip)
        assert subgroup_sum(coeffs, skip).to_ints() == OE.subgroup_sum(of, coeffs.to_ints(), skip)
        nodes = UnivariateSkipDomain(skip, field).nodes().to_ints()
        w = OE.subgroup_root(of, skip)
        assert nodes == [pow(w, j, of.p) for j in range(1 << skip)]
