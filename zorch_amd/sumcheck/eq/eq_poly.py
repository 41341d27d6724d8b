"""EqPoly round (drop-in for reference zorch/sumcheck/eq/eq_poly.py:114-221).

``EqPolyRound(summand, w, challenges=C)`` is one eq-weighted round under
``fold_rounds``: s_i(t) = l_i(t) * sum_x eq_l[x_hi] eq_r[x_lo] prod_k P_k(t, x)
on the natural domain {0..degree+1}, then fold and eq-mass update, on the GPU.
Host helpers build eq tables for small checks.
"""

from __future__ import annotations

from zorch_amd import _lib
from zorch_amd.fields import FieldArray, concatenate, ones, stack
from zorch_amd.runtime import DeviceFoldState, DeviceTables
from zorch_amd.sumcheck.prover import FoldingClaim, ProductSummand, _advance_device


def expand_hypercube_step(state: FieldArray, coord: FieldArray, *, msb: bool = False) -> FieldArray:
    """(2^k,) -> (2^(k+1),) adding a variable's (1-c)/c split (poly/eq.py:55-62)."""
    high = state * coord
    low = state - high
    if msb:
        return concatenate([low, high])
    return stack([low, high], axis=1).reshape(-1)


def compute_eq_evaluations(w: FieldArray) -> list[FieldArray]:
    """Suffix eq tables [eq(w[-1:]), ..., eq(w)] with each new coordinate
    prepended as the MSB (eq_poly.py:46-51)."""
    state = ones((1,), w.field)
    tables = []
    for j in range(w.shape[0] - 1, -1, -1):
        state = expand_hypercube_step(state, w[j], msb=True)
        tables.append(state)
    return tables


def eq_factor(t: FieldArray, z: FieldArray) -> FieldArray:
    """t z + (1 - t)(1 - z) (poly/eq.py:16-26)."""
    one = ones((), z.field)
    return t * z + (one - t) * (one - z)


class EqPolyRound:
    """One EqPoly variable-binding round, reused across all rounds."""

    def __init__(self, summand: ProductSummand, w: FieldArray, domain=None, *, challenges,
                 msb: bool = True):
        if domain is not None:
            raise ValueError("the device EqPoly round samples the natural domain")
        if not msb:
            raise ValueError("the device EqPoly round binds MSB-first")
        self.summand = summand
        self.w = w
        self.challenges = challenges
        self.l = int(w.shape[0])

    def _make_state(self, state, claim: FieldArray) -> DeviceFoldState:
        tables = state[0] if isinstance(state, tuple) else state
        if not isinstance(tables, (FieldArray, DeviceTables)):
            raise TypeError("state must be the stacked factor tables")
        return DeviceFoldState(tables, claim, self.challenges.dtype, _lib.ZS_MODE_EQ,
                               self.summand.degree, eq_point=self.w)

    def __call__(self, carry: FoldingClaim, transcript):
        return _advance_device(carry, transcript,
                               lambda st: self._make_state(st, carry.claim.value))
