"""Sqrt-space sumcheck (drop-in for reference zorch/sumcheck/sqrt_space.py).

The first l//2 rounds never fold the factors: each round refolds them over
the bound prefix with the running eq(r[<i], .) table and sends the round
polynomial of the refolded tables; the tail is a plain product sumcheck over
the tables folded at the boundary. The messages are those of the linear-time
prover on the same challenges (reference sqrt_space.py:1-20).

On the GPU (``zs_prove`` with ``refold_rounds = l // 2``) a refold round is
one streaming pass over the untouched input, a block spreading each pair's
prefix sum over its threads once pairs grow few; the last refold round stores
its refolded round table, the one the first dense round folds (2 * 2^(l-l//2)
entries a table: the only folded state the engine holds); the rounds after
it are the dense kernels. Messages may be sampled on any wire form of the
product rounds (the compressed U-hat_degree by default, as the reference).
"""

from __future__ import annotations

from typing import Any

from zorch_amd import _lib
from zorch_amd.challenge import ChallengePolicy
from zorch_amd.fields import FieldArray, ones, stack, zeros
from zorch_amd.round import RunningClaim
from zorch_amd.runtime import DeviceFoldState, DeviceTables, log2_strict
from zorch_amd.sumcheck._device import prove_stage
from zorch_amd.sumcheck.domain import EvalDomain, uhat_domain
from zorch_amd.sumcheck.prover import (FoldingClaim, ProductSummand, _advance_device,
                                       round_wire)

# (P_stacked, eq_evals) in the reference: the untouched factors and the eq
# table over the bound prefix. Here the device state carries both.
SqrtSpaceState = tuple


def compute_folded_evaluations(p_stacked: FieldArray, eq_evals: FieldArray) -> FieldArray:
    """p_k(r[<i], x') = sum_b eq(r[<i], b) p_k(b, x') over the bound prefix b on
    the high index bits, shape (d, 2^(l-i)) (reference sqrt_space.py:40-48).
    A host helper for small arrays; the engine refolds on the device."""
    d, width = p_stacked.shape[0], p_stacked.shape[1]
    i = log2_strict(eq_evals.shape[0])
    L = width >> i
    f = eq_evals.field
    rows = []
    for q in range(d):
        acc = zeros((L,), f)
        row = p_stacked[q].astype(f)
        for b in range(1 << i):
            acc = acc + row[b * L:(b + 1) * L] * eq_evals[b]
        rows.append(acc)
    return stack(rows)


def _check_summand(summand) -> None:
    if type(summand) is not ProductSummand:
        raise NotImplementedError("the device sqrt-space engine takes the product summand")


class SqrtSpaceRound:
    """One first-phase round (reference sqrt_space.py:51-96): the round
    polynomial of the refolded factors on `domain`, the challenge, the
    domain's claim reduction; the factors stay put. On the device the state
    is a folding state whose first l//2 rounds refold; rounds after those fold
    as ``StandardRound`` does, so the same state continues under it."""

    def __init__(self, summand: ProductSummand, domain: EvalDomain,
                 challenges: ChallengePolicy) -> None:
        _check_summand(summand)
        self.summand = summand
        self.domain = domain
        self.challenges = challenges
        round_wire(domain, summand.degree + 1)

    def _make_state(self, state, claim: FieldArray) -> DeviceFoldState:
        if isinstance(state, tuple):
            tables, eq_evals = state
            if eq_evals.shape[0] != 1:
                raise ValueError("a sqrt-space device state starts with nothing bound "
                                 "(eq_evals of length 1)")
        else:
            tables = state
        if not isinstance(tables, (FieldArray, DeviceTables)):
            raise TypeError("state must be the stacked factor tables")
        kind, domain = round_wire(self.domain, self.summand.degree + 1)
        half = log2_strict(tables.shape[1]) // 2
        return DeviceFoldState(tables, claim, self.challenges.dtype, _lib.ZS_MODE_PRODUCT,
                               self.summand.degree, wire_kind=kind, domain=domain,
                               refold_rounds=half)

    def __call__(self, carry: FoldingClaim, transcript):
        return _advance_device(carry, transcript,
                               lambda st: self._make_state(st, carry.claim.value))


def prove_sqrt_space(p_initial, claim: FieldArray, transcript, summand: ProductSummand | None = None,
                     domain: EvalDomain | None = None, *, challenges: ChallengePolicy):
    """Prove the sumcheck with a sqrt-space first phase and a standard second
    phase (reference sqrt_space.py:99-132): returns (carry, transcript,
    messages) with carry.state the final folded factors (d, 1) and carry.claim
    the reduced claim over all l challenges. One device call."""
    summand = summand or ProductSummand(degree=p_initial.shape[0])
    _check_summand(summand)
    domain = domain or uhat_domain(summand.degree, p_initial.dtype)
    l = log2_strict(p_initial.shape[1])
    kind, dom = round_wire(domain, summand.degree + 1)
    out = prove_stage(p_initial, claim, challenges.dtype, _lib.ZS_MODE_PRODUCT, summand.degree,
                      transcript, wire_kind=kind, domain=dom, refold_rounds=l // 2)
    final = out.final.reshape(p_initial.shape[0], 1)
    carry = FoldingClaim(final, RunningClaim(out.value, out.point, l))
    return carry, out.transcript, [out.msgs[i] for i in range(l)]


__all__ = ["SqrtSpaceState", "SqrtSpaceRound", "compute_folded_evaluations", "prove_sqrt_space"]
