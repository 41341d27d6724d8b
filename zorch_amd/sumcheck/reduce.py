"""Per-round claim reductions (reference zorch/sumcheck/reduce.py:18-52) and
univariate helpers (zorch/poly/univariate.py:36-65), on host field arrays.
"""

from __future__ import annotations

from zorch_amd.fields import FieldArray, array, fsum, inv, stack


def require_width(msg: FieldArray, expected: int, kind: str = "evals") -> None:
    """Reject a malformed message before anything reads the claim."""
    if msg.shape[0] != expected:
        raise ValueError(f"round message must have {expected} {kind}, got {msg.shape[0]}")


def compute_lagrange_basis(r: FieldArray, n: int) -> FieldArray:
    """L_k(r) = prod_{j != k} (r - j) / (k - j) over the naturals {0..n-1},
    direct form (no division by r - node)."""
    f = r.field
    out = []
    for k in range(n):
        num = array(1, f)
        den = array(1, f.base)
        for j in range(n):
            if j == k:
                continue
            num = num * (r - array(j, f))
            den = den * array(k - j, f.base)
        out.append(num * inv(den).astype(f))
    return stack(out)


def eval_univariate(evals: FieldArray, x: FieldArray) -> FieldArray:
    """Value at x of the polynomial with values `evals` on {0..len-1}."""
    basis = compute_lagrange_basis(x, evals.shape[0])
    return fsum(evals * basis)


def reduce_evals(claim: FieldArray, msg: FieldArray, r: FieldArray, degree: int):
    """`msg` is s on the naturals {0..degree}: (s(r), claim == s(0) + s(1))."""
    require_width(msg, degree + 1)
    return eval_univariate(msg, r), bool(claim == msg[0] + msg[1])
