"""EqPoly round (drop-in for reference zorch/sumcheck/eq/eq_poly.py:114-221).

``EqPolyRound(summand, w, domain=None, challenges=C)`` is one eq-weighted round
under ``fold_rounds``: s_i = l_i * t_i with
t_i(x) = sum_y eq_l[y_hi] eq_r[y_lo] prod_k P_k(x, y) and l_i the linear eq
factor of the bound variable, sampled on `domain` (the compressed
Û_degree = {inf, 0, 2, ..., degree-1} by default, as the reference; the
natural domain {0..degree+1} is what ``EqPolyProver`` passes), then the claim
reduction of that domain (reduce_domain), the fold and the eq-mass update, on
the GPU. ``EqPolyCoeffsRound`` sends the ascending coefficients of s_i instead
(the standalone-verifiable form of the reference's ``_round_coeffs``,
eq_poly.py:189-206, checked by ``CoeffsSumcheckRound``). Host helpers build eq
tables for small checks.
"""

from __future__ import annotations

from zorch_amd import _lib
from zorch_amd.fields import FieldArray, concatenate, ones, stack
from zorch_amd.runtime import DeviceFoldState, DeviceTables
from zorch_amd.sumcheck.domain import EvalDomain, uhat_domain
from zorch_amd.sumcheck.prover import (FoldingClaim, ProductSummand, _advance_device,
                                       _scratch_message, round_wire)


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


def eq_root(z: FieldArray) -> FieldArray:
    """The root in t of eq_factor(t, z): (1 - z) / (1 - 2 z) (poly/eq.py:29-40)."""
    from zorch_amd.fields import array, inv
    one = ones((), z.field)
    return (one - z) * inv(one - array(2, z.field) * z)


def _split_state(state):
    """(tables, eq mass) of an EqPolyState tuple, or tables alone."""
    if isinstance(state, tuple):
        if len(state) != 2:
            raise ValueError("an EqPoly state is (stacked factors, eq mass)")
        mass = state[1]
        if isinstance(mass, FieldArray):
            mass = mass.reshape(-1)[0] if mass.ndim else mass
        return state[0], mass
    return state, None


class EqPolyRound:
    """One EqPoly variable-binding round, reused across all rounds (the round
    index is the number of rounds bound). Messages are sampled on `domain`:
    Û_degree by default (eq_poly.py:130), or any explicit domain."""

    def __init__(self, summand: ProductSummand, w: FieldArray, domain: EvalDomain | None = None,
                 *, challenges, msb: bool = True):
        if not msb:
            raise ValueError("the device EqPoly round binds MSB-first")
        self.summand = summand
        self.w = w
        self.domain = domain if domain is not None else uhat_domain(summand.degree, w.field)
        self.challenges = challenges
        self.msb = msb
        self.l = int(w.shape[0])
        self.l_half = self.l // 2
        self._wire()  # rejects a node-less domain early

    def _wire(self) -> tuple[int, EvalDomain | None]:
        return round_wire(self.domain, self.summand.degree + 2)

    def _make_state(self, state, claim: FieldArray) -> DeviceFoldState:
        tables, mass = _split_state(state)
        if not isinstance(tables, (FieldArray, DeviceTables)):
            raise TypeError("state must be the stacked factor tables")
        kind, domain = self._wire()
        return DeviceFoldState(tables, claim, self.challenges.dtype, _lib.ZS_MODE_EQ,
                               self.summand.degree, eq_point=self.w, wire_kind=kind,
                               domain=domain, eq_mass=mass)

    def _round_poly(self, state) -> tuple[FieldArray, None]:
        """The round message sampled on this round's domain, without advancing
        any caller state. The device keeps its own fold cache, so the second
        element (the reference's host cache) is None."""
        return _scratch_message(self._make_state, state, self.challenges.dtype), None

    def _round_coeffs(self, state) -> tuple[FieldArray, None]:
        """Ascending coefficients of s_i (degree + 2 of them), without advancing
        any caller state (eq_poly.py:189-206)."""
        rnd = EqPolyCoeffsRound(self.summand, self.w, challenges=self.challenges)
        return _scratch_message(rnd._make_state, state, self.challenges.dtype), None

    def __call__(self, carry: FoldingClaim, transcript):
        return _advance_device(carry, transcript,
                               lambda st: self._make_state(st, carry.claim.value))


class EqPolyCoeffsRound(EqPolyRound):
    """EqPoly round sending the ascending coefficients of s_i = l_i t_i (the
    form ``CoeffsSumcheckRound(degree + 1)`` verifies; claim reduced by
    reduce_coeffs)."""

    def __init__(self, summand: ProductSummand, w: FieldArray, *, challenges, msb: bool = True):
        super().__init__(summand, w, None, challenges=challenges, msb=msb)

    def _wire(self) -> tuple[int, EvalDomain | None]:
        return _lib.ZS_WIRE_COEFFS, None
