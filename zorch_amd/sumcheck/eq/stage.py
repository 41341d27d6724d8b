"""Equality-factored sumcheck roles (drop-in for reference
zorch/sumcheck/eq/stage.py).

``EqPolyProver(ProductSummand(d), challenges=C).prove`` proves
sum_x eq(w, x) * prod_k P_k(x) with eq factored into suffix half tables
(never materialised) and messages on the natural domain {0..d+1}. The whole
stage is one device call (``zs_prove`` in eq mode).
"""

from __future__ import annotations

from dataclasses import dataclass

from zorch_amd import _lib
from zorch_amd.fields import FieldArray
from zorch_amd.stage import ProveResult, VerifyResult
from zorch_amd.sumcheck._device import prove_stage
from zorch_amd.sumcheck.prover import ProductSummand
from zorch_amd.sumcheck.stage import EvaluationClaim
from zorch_amd.sumcheck.verifier import SumcheckRound
from zorch_amd.verify import verify


@dataclass(frozen=True)
class EqSumClaim:
    """Stated total of eq(equality_point, x) times the summand over the cube
    (reference zorch/sumcheck/eq/stage.py:24-30)."""

    equality_point: FieldArray
    value: FieldArray
    rounds: int


@dataclass(frozen=True)
class EqPolyWitness:
    """The stacked summand factors of an ``EqSumClaim`` (the eq factor is not
    among them: it is derived from the claim's point)."""

    factors: object


def _check_claim(claim: EqSumClaim) -> None:
    if claim.equality_point.shape[0] != claim.rounds:
        raise ValueError(f"equality point needs {claim.rounds} coordinates, "
                         f"got {claim.equality_point.shape[0]}")


class EqPolyProver:
    """Prover role of the eq-weighted sumcheck: eq(w, x) stays a factored
    weight beside the summand tables (zorch/sumcheck/eq/stage.py:48-92)."""

    def __init__(self, summand: ProductSummand, *, challenges) -> None:
        if type(summand) is not ProductSummand:
            raise ValueError("the device prover supports the product summand")
        self.summand = summand
        self.challenges = challenges
        self.degree = summand.degree + 1
        self.verifier_round = SumcheckRound(self.degree, challenges)

    def prove(self, claim: EqSumClaim, witness: EqPolyWitness, transcript) -> ProveResult:
        _check_claim(claim)
        out = prove_stage(witness.factors, claim.value, self.challenges.dtype, _lib.ZS_MODE_EQ,
                          self.summand.degree, transcript, eq_point=claim.equality_point)
        return ProveResult(EvaluationClaim(out.point, out.value), out.msgs, out.transcript)


class EqPolyVerifier:
    """Verify equality-weighted sumcheck (zorch/sumcheck/eq/stage.py:95-125)."""

    def __init__(self, summand: ProductSummand, *, challenges) -> None:
        self.verifier_round = SumcheckRound(summand.degree + 1, challenges)

    def verify(self, claim: EqSumClaim, reduction_proof: FieldArray, transcript) -> VerifyResult:
        _check_claim(claim)
        if reduction_proof.shape[0] != claim.rounds:
            raise ValueError(f"expected {claim.rounds} sumcheck rounds, "
                             f"got {reduction_proof.shape[0]}")
        point, value, transcript, ok = verify(self.verifier_round, claim.value, reduction_proof,
                                              transcript)
        return VerifyResult(EvaluationClaim(point, value), transcript, ok)
