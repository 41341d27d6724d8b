"""Dense LogUp-GKR as a prover/verifier stage pair (drop-in for reference
zorch/logup_gkr/stage.py).

The stage reduces a public output claim (the circuit's output fractions and
the number of layers the verifier expects) to a claim on the input layer:
the interleaved numerator and denominator MLEs evaluated at one point. The
prover builds the pyramid on the device and proves every layer from the
floor out (each layer's sumcheck one ``ZS_MODE_LOGUP`` device stage on the
layer's own buffer); the verifier replays the same chain (stage.py:59-128). The layer count is the verifier's own: a
witness or proof of another height is rejected before any work.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Generic, TypeVar

from zorch_amd.challenge import ChallengePolicy
from zorch_amd.fields import FieldArray, as_field
from zorch_amd.logup_gkr.circuit import GkrLayer, LogUpGkrOutput
from zorch_amd.logup_gkr.prover import GkrLayerRound as ProverLayerRound
from zorch_amd.logup_gkr.prover import LayerProof, bind_output, device_pyramid
from zorch_amd.logup_gkr.verifier import GkrLayerRound as VerifierLayerRound
from zorch_amd.round import prove_rounds, verify_rounds
from zorch_amd.stage import ProveResult, VerifyResult

LayerProofT = TypeVar("LayerProofT")


@dataclass(frozen=True)
class LogUpOutputClaim:
    """The circuit's output fractions and the layer count the verifier expects."""

    output: LogUpGkrOutput
    layers: int


@dataclass(frozen=True)
class InputLayerClaim:
    """The input layer's numerator and denominator MLEs claimed at `point`."""

    numerator: FieldArray
    denominator: FieldArray
    point: FieldArray


@dataclass(frozen=True)
class GkrProof(Generic[LayerProofT]):
    """One layer proof per GKR layer, floor first."""

    layers: tuple


def _input_claim(carry) -> InputLayerClaim:
    numerator, denominator, point = carry
    return InputLayerClaim(numerator, denominator, point)


class LogUpGkrProver:
    """Prove an output claim down to an input-layer claim."""

    def __init__(self, challenges: ChallengePolicy) -> None:
        self.challenges = challenges

    def prove(self, claim: LogUpOutputClaim, witness: GkrLayer, transcript) -> ProveResult:
        if witness.num_row_variables != claim.layers:
            raise ValueError(f"claim expects {claim.layers} GKR layers, "
                             f"witness needs {witness.num_row_variables}")
        pyramid = device_pyramid(witness, as_field(self.challenges.dtype))
        carry, transcript = bind_output(claim.output, transcript, self.challenges)
        carry, transcript, proofs = prove_rounds(
            [ProverLayerRound(layer, self.challenges) for layer in reversed(pyramid[:-1])],
            carry, transcript)
        proofs = tuple(proofs)
        if len(proofs) != claim.layers:
            raise AssertionError(f"built {len(proofs)} GKR proofs for {claim.layers} layers")
        return ProveResult(_input_claim(carry), GkrProof(proofs), transcript)


class LogUpGkrVerifier:
    """Verify an output-to-input-layer LogUp-GKR reduction."""

    def __init__(self, challenges: ChallengePolicy) -> None:
        self.challenges = challenges

    def verify(self, claim: LogUpOutputClaim, reduction_proof: GkrProof, transcript) -> VerifyResult:
        if len(reduction_proof.layers) != claim.layers:
            raise ValueError(f"expected {claim.layers} GKR layers, "
                             f"got {len(reduction_proof.layers)}")
        carry, transcript = bind_output(claim.output, transcript, self.challenges)
        carry, transcript, ok = verify_rounds(
            [VerifierLayerRound(self.challenges) for _ in range(claim.layers)], carry,
            list(reduction_proof.layers), transcript)
        return VerifyResult(_input_claim(carry), transcript, ok)


__all__ = ["LogUpOutputClaim", "InputLayerClaim", "GkrProof", "LogUpGkrProver",
           "LogUpGkrVerifier", "LayerProof"]
