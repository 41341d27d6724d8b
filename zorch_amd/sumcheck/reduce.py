"""Per-round claim reductions (reference zorch/sumcheck/reduce.py:18-66) and
univariate helpers (zorch/poly/univariate.py:36-131), on host field arrays.

One reduction per wire form, each returning ``(reduced, ok)`` with ``ok`` the
round's own consistency check; both roles use them (the device prover applies
the same maps in its finisher).
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


def eval_coeffs(coeffs: FieldArray, x: FieldArray) -> FieldArray:
    """sum_i coeffs[i] x^i by Horner (univariate.py:109-131)."""
    out = coeffs[coeffs.shape[0] - 1]
    for i in range(coeffs.shape[0] - 2, -1, -1):
        out = out * x + coeffs[i]
    return out


def reduce_evals(claim: FieldArray, msg: FieldArray, r: FieldArray, degree: int):
    """`msg` is s on the naturals {0..degree}: (s(r), claim == s(0) + s(1))."""
    require_width(msg, degree + 1)
    return eval_univariate(msg, r), bool(claim == msg[0] + msg[1])


def reduce_coeffs(claim: FieldArray, msg: FieldArray, r: FieldArray, degree: int):
    """`msg` is s's ascending coefficients: s(0) = c_0 and s(1) = sum c."""
    require_width(msg, degree + 1, "coefficients")
    return eval_coeffs(msg, r), bool(claim == msg[0] + fsum(msg))


def reduce_domain(claim: FieldArray, msg: FieldArray, r: FieldArray, domain):
    """`msg` is s sampled at `domain`'s points: its coefficients through the
    domain's value-to-coefficient map, evaluated at r."""
    coeffs = domain.to_coeffs(msg)
    return eval_coeffs(coeffs, r), bool(claim == coeffs[0] + fsum(coeffs))


def reduce_compressed(claim: FieldArray, msg: FieldArray, r: FieldArray):
    """`msg` is [c_0, c_2]; c_1 = claim - 2 c_0 - c_2 spends the
    s(0) + s(1) = claim identity, so ``ok`` is constant true."""
    require_width(msg, 2, "coefficients [c_0, c_2]")
    c0, c2 = msg[0], msg[1]
    c1 = claim - c0 - c0 - c2
    return eval_coeffs(stack([c0, c1, c2]), r), True


def reduce_subgroup(claim: FieldArray, msg: FieldArray, r: FieldArray, skip_rounds: int,
                    degree: int):
    """The skip's round 0 (reduce.py:68-74): `msg` is s_0's ascending
    coefficients, checked against the subgroup sum claim == sum_{z in D} s_0(z)
    rather than the hypercube identity."""
    from zorch_amd.sumcheck.domain import subgroup_sum
    require_width(msg, degree * ((1 << skip_rounds) - 1) + 1, "coefficients")
    ok = bool(claim.astype(msg.field) == subgroup_sum(msg, skip_rounds)) \
        if claim.field.degree == 1 else bool(claim == subgroup_sum(msg, skip_rounds).astype(
            claim.field))
    return eval_coeffs(msg.astype(r.field), r), ok
