"""Sumcheck roles reducing a sum claim to an evaluation claim (drop-in for
reference zorch/sumcheck/stage.py).

``SumcheckProver.prove`` keeps the reference signature. When its round is the
plain product ``StandardRound`` on the natural domain and the claim covers
every variable, the whole stage runs as one device call (``zs_prove``): all
rounds are enqueued on one stream with the Fiat-Shamir hop on the GPU and a
single synchronisation at the end. Any other round object is driven by
``fold_rounds`` exactly as the reference does.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any

from zorch_amd import _lib
from zorch_amd.fields import FieldArray, stack
from zorch_amd.prove import fold_rounds
from zorch_amd.stage import ProveResult, VerifyResult
from zorch_amd.sumcheck._device import prove_stage
from zorch_amd.sumcheck.prover import ProductSummand, StandardRound, initial_claim
from zorch_amd.verify import verify


@dataclass(frozen=True)
class SumClaim:
    """Public claim that a polynomial sums to ``value`` over a Boolean cube."""

    value: FieldArray
    rounds: int


@dataclass(frozen=True)
class SumcheckWitness:
    """Dense factor tables witnessing a ``SumClaim`` (FieldArray or DeviceTables)."""

    state: Any


@dataclass(frozen=True)
class EvaluationClaim:
    """Claim that the reduced polynomial evaluates to ``value`` at ``point``."""

    point: FieldArray
    value: FieldArray


def _fast_path_ok(rnd, rounds: int, tables) -> bool:
    return (isinstance(rnd, StandardRound) and type(rnd.summand) is ProductSummand
            and hasattr(tables, "shape") and len(tables.shape) == 2
            and tables.shape[1] == 1 << rounds)


class SumcheckProver:
    """The prover role of dense sumcheck (zorch/sumcheck/stage.py:43-69)."""

    def __init__(self, prover_round) -> None:
        self.prover_round = prover_round

    def prove(self, claim: SumClaim, witness: SumcheckWitness, transcript) -> ProveResult:
        rnd = self.prover_round
        if _fast_path_ok(rnd, claim.rounds, witness.state):
            out = prove_stage(witness.state, claim.value, rnd.challenges.dtype,
                              _lib.ZS_MODE_PRODUCT, rnd.summand.degree, transcript)
            return ProveResult(EvaluationClaim(out.point, out.value), out.msgs, out.transcript)
        folded, transcript, messages = fold_rounds(
            rnd, initial_claim(witness.state, claim.value, claim.rounds), transcript,
            claim.rounds)
        reduced = folded.claim
        return ProveResult(EvaluationClaim(reduced.point, reduced.value), stack(messages),
                           transcript)


class SumcheckVerifier:
    """The verifier role of dense sumcheck (zorch/sumcheck/stage.py:72-92)."""

    def __init__(self, verifier_round) -> None:
        self.verifier_round = verifier_round

    def verify(self, claim: SumClaim, reduction_proof: FieldArray, transcript) -> VerifyResult:
        if reduction_proof.shape[0] != claim.rounds:
            raise ValueError(
                f"expected {claim.rounds} sumcheck rounds, got {reduction_proof.shape[0]}")
        point, value, transcript, ok = verify(self.verifier_round, claim.value,
                                              reduction_proof, transcript)
        return VerifyResult(EvaluationClaim(point, value), transcript, ok)
