"""Sumcheck roles reducing a sum claim to an evaluation claim (drop-in for
reference zorch/sumcheck/stage.py).

``SumcheckProver.prove`` keeps the reference signature. When its round is one
of the device product rounds (``StandardRound`` on any domain,
``CompressedProductRound``, ``CoeffsProductRound``) over the product summand
and the claim covers every variable, the whole stage runs as one device call
(``zs_prove``): all rounds are enqueued on one stream with the Fiat-Shamir hop
on the GPU and a single synchronisation at the end. Any other round object is
driven by ``fold_rounds`` exactly as the reference does.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any

from zorch_amd import _lib
from zorch_amd.fields import FieldArray, stack
from zorch_amd.prove import fold_rounds
from zorch_amd.stage import ProveResult, VerifyResult
from zorch_amd.sumcheck._device import prove_stage
from zorch_amd.sumcheck.prover import (CompressedProductRound, ProductSummand,
                                       _DeviceProductRound, initial_claim)
from zorch_amd.verify import verify


@dataclass(frozen=True)
class SumClaim:
    """The stated total ``value`` of the summand over {0,1}^rounds (the input
    of a sumcheck stage; reference zorch/sumcheck/stage.py:20-25)."""

    value: FieldArray
    rounds: int


@dataclass(frozen=True)
class SumcheckWitness:
    """The stacked factor tables (k x 2^rounds) whose product sums to a
    ``SumClaim``: a host FieldArray or device-resident DeviceTables."""

    state: Any


@dataclass(frozen=True)
class EvaluationClaim:
    """The stage's output: the bound challenges ``point`` (one per round) and
    the value the summand must take there."""

    point: FieldArray
    value: FieldArray


def _fast_path_ok(rnd, rounds: int, tables) -> bool:
    return (isinstance(rnd, _DeviceProductRound) and type(rnd.summand) is ProductSummand
            and hasattr(tables, "shape") and len(tables.shape) == 2
            and tables.shape[1] == 1 << rounds)


class SumcheckProver:
    """The prover role of dense sumcheck (zorch/sumcheck/stage.py:43-69)."""

    def __init__(self, prover_round) -> None:
        self.prover_round = prover_round

    def prove(self, claim: SumClaim, witness: SumcheckWitness, transcript) -> ProveResult:
        rnd = self.prover_round
        if _fast_path_ok(rnd, claim.rounds, witness.state):
            if isinstance(rnd, CompressedProductRound) and witness.state.shape[0] != 2:
                raise ValueError(f"compressed product round takes exactly 2 factors, got "
                                 f"{witness.state.shape[0]}")
            kind, domain = rnd._stage_wire()
            out = prove_stage(witness.state, claim.value, rnd.challenges.dtype,
                              _lib.ZS_MODE_PRODUCT, rnd.summand.degree, transcript,
                              wire_kind=kind, domain=domain)
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
