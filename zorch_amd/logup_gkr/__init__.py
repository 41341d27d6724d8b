"""Dense LogUp-GKR (drop-in for reference zorch/logup_gkr: circuit.py,
prover.py, verifier.py, stage.py): the fractional-sum circuit on host field arrays, each
layer's sumcheck on the GPU (``ZS_MODE_LOGUP`` stages of the C ABI)."""

from zorch_amd.logup_gkr.circuit import (GkrLayer, LogUpGkrOutput, build_pyramid,
                                         extract_outputs, layer_transition)
from zorch_amd.logup_gkr.prover import (GkrLayerRound, LayerProof, LogupSumcheckRound,
                                        LogupSummand, bind_output, fold_carry, logup_combine,
                                        prove_gkr)
from zorch_amd.logup_gkr.stage import (GkrProof, InputLayerClaim, LogUpGkrProver,
                                       LogUpGkrVerifier, LogUpOutputClaim)
from zorch_amd.logup_gkr.verifier import GkrLayerRound as GkrLayerVerifierRound
from zorch_amd.logup_gkr.verifier import verify_gkr

__all__ = ["GkrLayer", "LogUpGkrOutput", "build_pyramid", "extract_outputs", "layer_transition",
           "GkrLayerRound", "LayerProof", "LogupSumcheckRound", "LogupSummand", "bind_output",
           "fold_carry", "logup_combine", "prove_gkr", "GkrLayerVerifierRound", "verify_gkr",
           "GkrProof", "InputLayerClaim", "LogUpGkrProver", "LogUpGkrVerifier",
           "LogUpOutputClaim"]
