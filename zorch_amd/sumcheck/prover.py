"""Sumcheck prover rounds (drop-in for reference zorch/sumcheck/prover.py).

``StandardRound(ProductSummand(d), challenges=C)`` is the reference's plain
materialised product round on the natural domain (prover.py:127-174): send
s(t) = sum_j prod_k (lo_k[j] + t (hi_k[j] - lo_k[j])) for t = 0..d, draw the
challenge, reduce the claim, fold MSB-first. Here the tables live on the GPU
(``DeviceFoldState``) and each call is one fused fold+evaluate pass plus the
device Fiat-Shamir hop (``zs_state_round``).

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


def _check_domain(domain: EvalDomain | None, degree: int) -> None:
    if domain is None:
        return
    if not domain.is_natural() or domain.nodes.shape[0] != degree + 1:
        raise ValueError("only the natural domain {0..degree} is supported by the device prover")


def _advance_device(carry: FoldingClaim, transcript, make_state) -> tuple:
    state = carry.state
    if not isinstance(state, DeviceFoldState):
        state = make_state(state)
    if state.bound != carry.claim.index:
        raise ValueError("this carry was already advanced; a device fold state is consumed "
                         "by each round")
    msg, chal, reduced, transcript = state.round(transcript)
    return FoldingClaim(state, carry.claim.bind(reduced, chal)), transcript, msg


class StandardRound:
    """The plain materialised product round on the natural domain."""

    def __init__(self, summand: ProductSummand, domain: EvalDomain | None = None, *,
                 challenges: ChallengePolicy):
        _check_domain(domain, summand.degree)
        self.summand = summand
        self.domain = domain
        self.challenges = challenges

    def _make_state(self, tables, claim: FieldArray) -> DeviceFoldState:
        if not isinstance(tables, (FieldArray, DeviceTables)):
            raise TypeError("state must be stacked tables (FieldArray or DeviceTables)")
        return DeviceFoldState(tables, claim, self.challenges.dtype, _lib.ZS_MODE_PRODUCT,
                               self.summand.degree)

    def _round_poly(self, folded) -> FieldArray:
        """s sampled on {0..degree} for stacked tables, without advancing any
        caller state (a scratch transcript absorbs the hop)."""
        from zorch_amd.sha256_field_transcript import Sha256FieldTranscript

        claim = zeros((), self.challenges.dtype)
        st = self._make_state(folded, claim)
        msg, _, _, _ = st.round(Sha256FieldTranscript.new(b"", self.challenges.dtype.base))
        return msg

    def __call__(self, carry: FoldingClaim, transcript):
        return _advance_device(carry, transcript,
                               lambda tables: self._make_state(tables, carry.claim.value))


def product(factors: Sequence[FieldArray]) -> FieldArray:
    return reduce(operator.mul, factors)


__all__ = [
    "FoldingClaim", "initial_claim", "ProductSummand", "StandardRound", "product", "fsum",
]
