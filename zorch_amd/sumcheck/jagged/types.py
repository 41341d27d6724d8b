"""Round-width caps of the jagged sumcheck (drop-in for reference
zorch/sumcheck/jagged/types.py:14-30).

The reference lays every round of a phase into one fixed-width buffer so
that one compiled kernel serves every round; the caps bound those widths.
The device prover here keeps each round's exact layout (segments back to
back, rows ceil(rc / 2^k)) and sizes its buffers from the row counts, so the
caps only carry the reference's contract: they are validated against the
layer (the same conditions, the same errors) and change nothing else.
"""

from __future__ import annotations

from dataclasses import dataclass


@dataclass(frozen=True)
class RoundWidthCaps:
    """`elements` bounds the row-phase width (sum of the row counts, a
    multiple of 4); `eq_row` the row-eq table (>= 2^nrv, even);
    `interaction` the batch-phase width (>= 2^niv, a multiple of 4)."""

    elements: int
    eq_row: int
    interaction: int

    def check(self, width: int, nrv: int, niv: int) -> None:
        """The reference's cap checks (jagged_prover.py:320-345, 565-570)."""
        if self.elements % 4:
            raise ValueError(f"elements cap {self.elements} must be a multiple of 4 (the "
                             "boundary handoff binds then pairs the row-width state, two "
                             "stride-2 halvings)")
        if self.eq_row % 2 or self.interaction % 4:
            raise ValueError(f"eq_row cap {self.eq_row} must be even and interaction cap "
                             f"{self.interaction} a multiple of 4 (each folds stride-2 "
                             "through its rounds)")
        if self.eq_row < (1 << nrv):
            raise ValueError(f"eq_row cap {self.eq_row} cannot hold the layer's row-eq table "
                             f"({1 << nrv})")
        if self.interaction < (1 << niv):
            raise ValueError(f"interaction cap {self.interaction} cannot hold the layer's "
                             f"interaction-eq table ({1 << niv})")
        if self.elements < width:
            raise ValueError(f"elements cap {self.elements} cannot hold the layer's row-phase "
                             f"plane width ({width}); widen the cap (or its ladder class) so "
                             "the fixed-width pad is non-negative")
