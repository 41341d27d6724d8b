"""Sumcheck prover rounds (drop-in for reference zorch/sumcheck/prover.py).

``StandardRound(ProductSummand(d), challenges=C)`` is the reference's plain
materialised product round (prover.py:127-174): send s(t) =
sum_j prod_k (lo_k[j] + t (hi_k[j] - lo_k[j])) sampled on its domain (the
naturals t = 0..d by default), draw the challenge, reduce the claim, fold
MSB-first. ``CompressedProductRound`` sends [c_0, c_2] (prover.py:177-212)
and ``CoeffsProductRound`` the ascending coefficients. Here the tables live on
the GPU (``DeviceFoldState``) and each call is one fused fold+evaluate pass
plus the device Fiat-Shamir hop (``zs_state_round``); the message form is a
map the device finisher applies to the round's natural values.

Messages, claims and points are in the challenge field; base-field tables are
lifted into it (the reference's tested pattern for extension challenges,
zorch/sumcheck/testing/verifier_test.py:97-114).
"""

from __future__ import annotations

import operator
from dataclasses import dataclass
from functools import reduce
from typing import Any, Sequence

from zorch_amd import _lib
from zorch_amd.challenge import ChallengePolicy
from zorch_amd.fields import FieldArray, fsum, zeros
from zorch_amd.round import RunningClaim
from zorch_amd.runtime import DeviceFoldState, DeviceTables
from zorch_amd.sumcheck.domain import EvalDomain


@dataclass(frozen=True)
class FoldingClaim:
    """What a sumcheck round threads: the folding state and the claim
    (prover.py:54-74). `state` is the stacked tables (a host FieldArray or
    DeviceTables before round 1) or a DeviceFoldState."""

    state: Any
    claim: RunningClaim

    def advance(self, state: Any, reduced: FieldArray, challenge: FieldArray) -> "FoldingClaim":
        return FoldingClaim(state, self.claim.bind(reduced, challenge))


def initial_claim(state: Any, value: FieldArray, rounds: int) -> FoldingClaim:
    """Start a fold from `state` with nothing bound (prover.py:86-91)."""
    return FoldingClaim(state, RunningClaim(value, zeros((rounds,), value.dtype), 0))


@dataclass(frozen=True)
class ProductSummand:
    """The product summand s = sum_x prod_k P_k(x) (prover.py:97-124)."""

    degree: int

    def __post_init__(self) -> None:
        if self.degree < 1:
            raise ValueError("degree must be >= 1")

    def combine_scalars(self) -> tuple:
        return ()

    def combine(self, scalars: Sequence[Any], *factors: FieldArray) -> FieldArray:
        del scalars
        return reduce(operator.mul, factors)

    def _combine(self, *factors: FieldArray) -> FieldArray:
        return self.combine(self.combine_scalars(), *factors)


def round_wire(domain: EvalDomain | None, npts: int) -> tuple[int, EvalDomain | None]:
    """The device wire form of messages sampled on `domain` for a round
    polynomial of npts natural values: the natural domain itself (or None) is
    the plain values; any other domain is sampled from them on the device."""
    if domain is None:
        return _lib.ZS_WIRE_VALUES, None
    if domain.nodes is None:
        raise ValueError("a sampling domain needs an explicit node set")
    if domain.is_natural() and domain.nodes.shape[0] == npts:
        return _lib.ZS_WIRE_VALUES, None
    return _lib.ZS_WIRE_DOMAIN, domain


def _advance_device(carry: FoldingClaim, transcript, make_state) -> tuple:
    state = carry.state
    if not isinstance(state, DeviceFoldState):
        state = make_state(state)
        # A state made here starts from the claim's current index (a skip's
        # bound prism starts after r0).
        state.start_index = carry.claim.index
    if state.start_index + state.bound != carry.claim.index:
        raise ValueError("this carry was already advanced; a device fold state is consumed "
                         "by each round")
    msg, chal, reduced, transcript = state.round(transcript)
    return FoldingClaim(state, carry.claim.bind(reduced, chal)), transcript, msg


def _scratch_message(make_state, tables, chal_dtype) -> FieldArray:
    """A round's message for stacked tables without advancing any caller
    state: a scratch device state and a scratch transcript absorb the hop."""
    from zorch_amd.sha256_field_transcript import Sha256FieldTranscript

    st = make_state(tables, zeros((), chal_dtype))
    msg, _, _, _ = st.round(Sha256FieldTranscript.new(b"", chal_dtype.base))
    return msg


class _DeviceProductRound:
    """Shared device plumbing of the product rounds: the wire form decides the
    message; tables fold MSB-first on the device (prover.py:76-83)."""

    summand: ProductSummand
    challenges: ChallengePolicy

    def _stage_wire(self) -> tuple[int, EvalDomain | None]:
        raise NotImplementedError

    def _make_state(self, tables, claim: FieldArray) -> DeviceFoldState:
        if not isinstance(tables, (FieldArray, DeviceTables)):
            raise TypeError("state must be stacked tables (FieldArray or DeviceTables)")
        kind, domain = self._stage_wire()
        return DeviceFoldState(tables, claim, self.challenges.dtype, _lib.ZS_MODE_PRODUCT,
                               self.summand.degree, wire_kind=kind, domain=domain)

    def _round_poly(self, folded) -> FieldArray:
        """The round message for stacked tables, without advancing any caller
        state (a scratch transcript absorbs the hop)."""
        return _scratch_message(self._make_state, folded, self.challenges.dtype)

    def __call__(self, carry: FoldingClaim, transcript):
        return _advance_device(carry, transcript,
                               lambda tables: self._make_state(tables, carry.claim.value))


class StandardRound(_DeviceProductRound):
    """The plain materialised product round (prover.py:127-174): messages
    sampled on `domain` (the natural {0..degree} when None) and reduced by
    Lagrange evaluation, or by the domain's value-to-coefficient map
    (reduce_domain) for an explicit domain."""

    def __init__(self, summand: ProductSummand, domain: EvalDomain | None = None, *,
                 challenges: ChallengePolicy):
        self.summand = summand
        self.domain = domain
        self.challenges = challenges
        round_wire(domain, summand.degree + 1)  # rejects a node-less domain early

    def _stage_wire(self):
        return round_wire(self.domain, self.summand.degree + 1)


class CompressedProductRound(_DeviceProductRound):
    """Two-factor product round on the compressed coefficient wire [c_0, c_2]
    (prover.py:177-212): c_0 = sum f0 b0, c_2 = sum (f1 - f0)(b1 - b0); the
    linear coefficient stays off the wire and the claim reduces through
    s(0) + s(1) = claim (reduce_compressed). Split and fold are those of
    ``StandardRound(ProductSummand(2))``."""

    def __init__(self, challenges: ChallengePolicy) -> None:
        self.challenges = challenges
        self.summand = ProductSummand(2)

    def _make_state(self, tables, claim: FieldArray) -> DeviceFoldState:
        if hasattr(tables, "shape") and tables.shape[0] != 2:
            raise ValueError(f"compressed product round takes exactly 2 factors, got "
                             f"{tables.shape[0]}")
        return super()._make_state(tables, claim)

    def _stage_wire(self):
        return _lib.ZS_WIRE_COMPRESSED, None


class CoeffsProductRound(_DeviceProductRound):
    """Product round sending the ascending coefficients of its round
    polynomial (the wire ``CoeffsSumcheckRound`` checks; the message of
    ``product_round_coeffs``, domain.py:216-223), reduced by reduce_coeffs."""

    def __init__(self, summand: ProductSummand, *, challenges: ChallengePolicy) -> None:
        self.summand = summand
        self.challenges = challenges

    def _stage_wire(self):
        return _lib.ZS_WIRE_COEFFS, None


def product(factors: Sequence[FieldArray]) -> FieldArray:
    return reduce(operator.mul, factors)


__all__ = [
    "FoldingClaim", "initial_claim", "ProductSummand", "StandardRound", "CompressedProductRound",
    "CoeffsProductRound", "round_wire", "product", "fsum",
]
