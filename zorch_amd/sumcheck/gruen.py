"""Gruen eq-factor round-poly assembly in coefficient form (drop-in for
reference zorch/sumcheck/gruen.py:75-157).

A round whose summand carries the bound variable's eq factor
eq_factor(t, z) = t z + (1 - t)(1 - z) gets two evaluations for free: the claim
identity s(1) = claim - s(0), and a zero at the factor's root
b = (1 - z) / (1 - 2 z). A degree-d round therefore materialises only s(0)
and d - 2 extra points; ``round_coeffs`` interpolates through
{0, 1, *extra, b} and returns ascending coefficients (the form
``CoeffsSumcheckRound`` checks). These are per-round scalar helpers (a few
field operations on the host); the table-scale work of an eq-weighted round
is the device prover's (``EqPolyCoeffsRound`` sends the same coefficients for
an honest claim, computed from the round's natural values).
"""

from __future__ import annotations

from typing import Sequence

from zorch_amd.fields import FieldArray, ones, stack, zeros
from zorch_amd.sumcheck.domain import EvalDomain
from zorch_amd.sumcheck.eq.eq_poly import eq_factor, eq_root
from zorch_amd.sumcheck.reduce import eval_coeffs


def interp_matrix(extra_ts: Sequence[FieldArray], z: FieldArray) -> FieldArray:
    """Value-to-coefficient matrix of the domain {0, 1, *extra_ts, eq_root(z)}:
    maps [s(0), s(1), *extra_ys, 0] to ascending coefficients, shape
    (degree + 1, degree + 1) with degree = len(extra_ts) + 2 (gruen.py:75-88)."""
    f = z.field
    xs = stack([zeros((), f), ones((), f), *[t.astype(f) for t in extra_ts], eq_root(z)])
    return EvalDomain(xs).coeff_matrix(f)


def round_coeffs_from_matrix(matrix: FieldArray, s_zero: FieldArray, claim: FieldArray,
                             extra_ys: Sequence[FieldArray]) -> FieldArray:
    """One round polynomial from a prebuilt ``interp_matrix``:
    values [s_zero, claim - s_zero, *extra_ys, 0] times the matrix (gruen.py:91-112)."""
    if matrix.shape[-1] != len(extra_ys) + 3:
        raise ValueError(
            f"matrix expects {matrix.shape[-1] - 3} extra evaluations "
            f"(domain {{0, 1, ..., b}}), got {len(extra_ys)}")
    f = claim.field
    ys = [s_zero.astype(f), claim - s_zero.astype(f), *[y.astype(f) for y in extra_ys],
          zeros((), f)]
    n = matrix.shape[0]
    rows = []
    for k in range(n):
        acc = zeros((), f)
        for j in range(n):
            acc = acc + matrix[k, j] * ys[j]
        rows.append(acc)
    return stack(rows)


def round_coeffs(s_zero: FieldArray, claim: FieldArray, extra_ts: Sequence[FieldArray],
                 extra_ys: Sequence[FieldArray], z: FieldArray) -> FieldArray:
    """Gruen-compressed round polynomial in coefficient form: the interpolant
    through {0, 1, *extra_ts, eq_root(z)} with values {s_zero, claim - s_zero,
    *extra_ys, 0}, degree len(extra_ts) + 2 (gruen.py:115-142)."""
    if len(extra_ts) != len(extra_ys):
        raise ValueError(
            f"extra_ts and extra_ys must pair up: got {len(extra_ts)} points "
            f"and {len(extra_ys)} evaluations")
    return round_coeffs_from_matrix(interp_matrix(extra_ts, z), s_zero, claim, extra_ys)


def fold_round_scalars(poly: FieldArray, r: FieldArray, mass: FieldArray,
                       z: FieldArray) -> tuple[FieldArray, FieldArray]:
    """(claim', mass'): the coefficient-form poly at r, and the bound eq mass
    times eq_factor(r, z) (gruen.py:145-157)."""
    return eval_coeffs(poly, r), mass * eq_factor(r, z)


__all__ = ["interp_matrix", "round_coeffs_from_matrix", "round_coeffs", "fold_round_scalars",
           "eq_root"]
