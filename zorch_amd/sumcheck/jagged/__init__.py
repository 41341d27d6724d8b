"""Scheme-agnostic parts of the jagged sumcheck (drop-in for reference
zorch/sumcheck/jagged): the round-width caps and the segment layout
recurrences. The LogUp-specific prover and verifier are in
``zorch_amd.logup_gkr.jagged_prover`` / ``jagged_verifier``."""

from zorch_amd.sumcheck.jagged.layout import _prepad_folded, _segment_gather_np
from zorch_amd.sumcheck.jagged.types import RoundWidthCaps

__all__ = ["RoundWidthCaps", "_prepad_folded", "_segment_gather_np"]
