"""Evaluation domains (reference zorch/sumcheck/domain.py:77-137).

The dense prover samples round polynomials on the natural domain {0..d}
(``natural_domain``), the wire form ``SumcheckRound`` checks. Other domains
(the compressed Û, Gruen nodes, subgroups) belong to other wire forms and
are not provided.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any

from zorch_amd.fields import FieldArray, naturals


@dataclass(frozen=True)
class EvalDomain:
    nodes: FieldArray | None = None
    inf_index: int | None = None

    def is_natural(self) -> bool:
        if self.nodes is None or self.inf_index is not None:
            return False
        return bool((self.nodes == naturals(self.nodes.shape[0], self.nodes.field)).all())


def natural_domain(degree: int, dtype: Any) -> EvalDomain:
    """{0, 1, ..., degree} in the base field of `dtype`."""
    return EvalDomain(naturals(degree + 1, dtype))


def split_halves(arr: FieldArray):
    """MSB split of the last axis into contiguous halves."""
    half = arr.shape[-1] // 2
    idx_lo = (slice(None),) * (arr.ndim - 1) + (slice(0, half),)
    idx_hi = (slice(None),) * (arr.ndim - 1) + (slice(half, None),)
    return arr[idx_lo], arr[idx_hi]


def fold(arr: FieldArray, r: FieldArray) -> FieldArray:
    """Host MSB fold P0 + r (P1 - P0) (domain.py:226-233), for small arrays."""
    if r.field is not arr.field:
        arr = arr.astype(r.field)
    p0, p1 = split_halves(arr)
    return p0 + (p1 - p0) * r
