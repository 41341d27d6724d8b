"""LogUp-GKR (drop-in for reference zorch/logup_gkr: circuit.py, prover.py,
verifier.py, stage.py, jagged_prover.py, jagged_verifier.py): the
fractional-sum circuit on host field arrays, each layer's sumcheck on the GPU
(``ZS_MODE_LOGUP`` stages of the C ABI for dense layers,
``zs_prove_jagged_layer`` for jagged ones)."""

from zorch_amd.logup_gkr.circuit import (GkrLayer, JaggedGkrLayer, LogUpGkrOutput,
                                         build_jagged_pyramid, build_pyramid,
                                         extract_jagged_outputs, extract_outputs,
                                         jagged_layer_transition, layer_transition)
from zorch_amd.logup_gkr.jagged_prover import (JaggedDeviceLayer, JaggedGkrLayerRound,
                                               JaggedLayerProof, RoundWidthCaps,
                                               jagged_device_pyramid, prove_jagged_layer)
from zorch_amd.logup_gkr.jagged_verifier import JaggedGkrLayerRound as JaggedVerifierLayerRound
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
           "LogUpOutputClaim", "JaggedGkrLayer", "jagged_layer_transition",
           "build_jagged_pyramid", "extract_jagged_outputs", "JaggedGkrLayerRound",
           "JaggedLayerProof", "RoundWidthCaps", "prove_jagged_layer", "JaggedVerifierLayerRound",
           "JaggedDeviceLayer", "jagged_device_pyramid"]
