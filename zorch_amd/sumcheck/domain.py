"""Evaluation domains of round polynomials (reference zorch/sumcheck/domain.py).

An ``EvalDomain`` names where a round polynomial is sampled: finite nodes,
optionally with the value "at infinity" (the leading coefficient) inserted at
``inf_index`` (0 first: the compressed Û domain; -1 last: the two-point
``[s(node), s(inf)]`` domain). The device prover always computes the natural
values s(0..D) of a round and maps them onto the requested wire form in its
finisher (``zs_wire``, include/zorch_amd.h), so any domain here is a wire form
of the same kernels. ``to_coeffs`` is the host value-to-coefficient map the
claim reductions use (reference domain.py:72-93), a handful of field
operations per round.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Any

import numpy as np

from zorch_amd import _lib
from zorch_amd.fields import FieldArray, array, as_field, concatenate, inv, naturals, stack


def _poly_mul_linear(poly: list, root: FieldArray) -> list:
    """(sum_i poly[i] x^i) (x - root), coefficients ascending."""
    out = [array(0, root.field) for _ in range(len(poly) + 1)]
    for i, c in enumerate(poly):
        out[i] = out[i] - c * root
        out[i + 1] = out[i + 1] + c
    return out


def interp_coeffs(nodes: FieldArray, values: FieldArray) -> FieldArray:
    """Ascending coefficients of the interpolant through (nodes[i], values[i])
    (the inverse-Vandermonde map of reference domain.py:45-50 and
    zorch/poly/univariate.py:68-106, by Lagrange expansion)."""
    f = values.field
    m = nodes.shape[0]
    if m == 0:
        return values
    xs = [nodes[i].astype(f) for i in range(m)]
    out = [array(0, f) for _ in range(m)]
    for j in range(m):
        poly = [array(1, f)]
        den = array(1, f)
        for k in range(m):
            if k == j:
                continue
            poly = _poly_mul_linear(poly, xs[k])
            den = den * (xs[j] - xs[k])
        if bool((den == array(0, f)).all()):
            raise ValueError("interpolation nodes must be distinct")
        scale = values[j] * inv(den)
        for i in range(m):
            out[i] = out[i] + poly[i] * scale
    return stack(out)


@dataclass(frozen=True)
class EvalDomain:
    """Sample points of a round polynomial and the map from samples to
    ascending coefficients (reference domain.py:52-106)."""

    nodes: FieldArray | None = None
    inf_index: int | None = None

    def is_natural(self) -> bool:
        """The plain naturals {0..m-1} with no infinity sample."""
        if self.nodes is None or self.inf_index is not None:
            return False
        nat = array(list(range(self.nodes.shape[0])), self.nodes.field)
        return bool((self.nodes == nat).all())

    @property
    def num_samples(self) -> int:
        if self.nodes is None:
            raise ValueError("a sampling domain needs an explicit node set")
        return self.nodes.shape[0] + (0 if self.inf_index is None else 1)

    def inf_position(self) -> int | None:
        """Where s(inf) sits among the samples (Python's modulo, as the
        reference's ``inf_index % (len(nodes) + 1)``)."""
        if self.inf_index is None:
            return None
        return self.inf_index % (self.nodes.shape[0] + 1)

    def coeff_matrix(self, field: Any = None) -> FieldArray:
        """Value -> coefficient matrix of the explicit finite nodes."""
        if self.nodes is None:
            raise ValueError("leading / naturals domain has no fixed matrix")
        f = as_field(field) if field is not None else self.nodes.field
        m = self.nodes.shape[0]
        cols = []
        for j in range(m):
            e = [array(0, f)] * m
            e[j] = array(1, f)
            cols.append(interp_coeffs(self.nodes, stack(e)))
        return stack(cols, axis=1)

    def to_coeffs(self, values: FieldArray) -> FieldArray:
        """Ascending coefficients from this domain's samples (domain.py:72-93):
        without infinity the interpolant through the nodes; with it, s(inf) is
        the top coefficient and the finite samples interpolate s - s(inf) x^d
        (d finite samples; the naturals {0..d-1} when nodes is None)."""
        if self.inf_index is None:
            if self.nodes is None:
                return interp_coeffs(naturals(values.shape[0], values.field), values)
            return interp_coeffs(self.nodes, values)
        n = values.shape[0]
        pos = self.inf_index % n
        v_inf = values[pos]
        finite = concatenate([values[:pos], values[pos + 1:]]) if n > 1 else values[:0]
        d = finite.shape[0]
        nodes = naturals(d, values.field) if self.nodes is None else self.nodes
        f = values.field
        if d:
            shifted = stack([finite[i] - v_inf * _pow(nodes[i].astype(f), d)
                             for i in range(d)])
            low = interp_coeffs(nodes, shifted)
            return concatenate([low, v_inf.reshape(1)])
        return v_inf.reshape(1)

    def sample(self, p0: FieldArray, p1: FieldArray) -> FieldArray:
        """The linear factor p0 + x (p1 - p0) at this domain's points, the
        leading axis indexing them; infinity gives the slope (domain.py:95-106).
        A host helper for small arrays."""
        if self.nodes is None:
            raise ValueError("a sampling domain needs an explicit node set")
        diff = p1 - p0
        pts = [p0 + diff * self.nodes[i].astype(p0.field) for i in range(self.nodes.shape[0])]
        pos = self.inf_position()
        if pos is not None:
            pts.insert(pos, diff)
        return stack(pts)

    def wire(self, chal_field: Any):
        """The C ABI wire form (a ``zs_wire`` and the buffers it points at)
        for messages sampled on this domain in the challenge field."""
        if self.nodes is None:
            raise ValueError("a sampling domain needs an explicit node set")
        f = as_field(chal_field)
        nodes = np.ascontiguousarray(self.nodes.astype(f).words) \
            if self.nodes.field is not f else np.ascontiguousarray(self.nodes.words)
        pos = self.inf_position()
        w = _lib.WireC(kind=_lib.ZS_WIRE_DOMAIN, num_nodes=self.nodes.shape[0],
                       has_inf=0 if pos is None else 1, inf_pos=0 if pos is None else pos,
                       nodes=nodes.ctypes.data if nodes.size else None)
        return w, nodes


def _pow(x: FieldArray, e: int) -> FieldArray:
    out = array(1, x.field)
    for _ in range(e):
        out = out * x
    return out


def natural_domain(degree: int, dtype: Any) -> EvalDomain:
    """{0, 1, ..., degree} in the base field of `dtype` (domain.py:131-137)."""
    return EvalDomain(naturals(degree + 1, dtype))


def uhat_domain(degree: int, dtype: Any) -> EvalDomain:
    """The compressed product round domain Û_degree = {inf, 0, 2, ..., degree-1}:
    infinity first, the node 1 dropped (domain.py:140-146)."""
    nat = naturals(degree, dtype)
    return EvalDomain(concatenate([nat[:1], nat[2:]]), inf_index=0)


def compressed_domain(node: int, dtype: Any) -> EvalDomain:
    """The two-point domain [s(node), s(inf)] with node 0 or 1 (domain.py:148-155)."""
    if node not in (0, 1):
        raise ValueError(f"compressed node must be 0 or 1, got {node}")
    return EvalDomain(naturals(node + 1, dtype)[node:], inf_index=-1)


def extend_to_round_domain(p0: FieldArray, p1: FieldArray, d: int, *,
                           skip_one: bool = False) -> FieldArray:
    """A linear pair lifted onto U_d = [inf, 0, 1, ..., d-1] (or Û_d without 1
    when skip_one; domain.py:109-127). A host helper for small arrays."""
    diff = p1 - p0
    if d == 1:
        return stack([diff, p0])
    base = [diff, p0] if skip_one else [diff, p0, p1]
    if d == 2:
        return stack(base)
    rest = [p0 + diff * array(u, p0.field) for u in range(2, d)]
    return stack(base + rest)


def split_halves(arr: FieldArray):
    """MSB split of the last axis into contiguous halves (domain.py:158-163)."""
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


def _device_message(stacked, wire_kind: int, domain: EvalDomain | None, chal: Any = None):
    """One round's message of the product summand over stacked tables, computed
    by the device prover (a scratch state and transcript; nothing advances)."""
    from zorch_amd.runtime import DeviceFoldState
    from zorch_amd.sha256_field_transcript import Sha256FieldTranscript
    from zorch_amd.fields import zeros

    f = as_field(chal if chal is not None else stacked.dtype)
    k = stacked.shape[0]
    st = DeviceFoldState(stacked, zeros((), f), f, _lib.ZS_MODE_PRODUCT, k,
                         wire_kind=wire_kind, domain=domain)
    msg, _, _, _ = st.round(Sha256FieldTranscript.new(b"", f.base))
    return msg


def product_round_poly(stacked) -> FieldArray:
    """s = sum_x prod_k f_k over Û_m for m stacked tables (domain.py:208-213),
    on the device."""
    m = stacked.shape[0]
    return _device_message(stacked, _lib.ZS_WIRE_DOMAIN, uhat_domain(m, stacked.dtype))


def product_round_coeffs(stacked) -> FieldArray:
    """Ascending coefficients of the degree-m product round polynomial of m
    stacked tables (domain.py:216-223), on the device."""
    return _device_message(stacked, _lib.ZS_WIRE_COEFFS, None)


# --- two-adic subgroup (univariate skip; reference domain.py:237-311) -------
# Host helpers for small arrays (the device round 0 has its own transforms).
# The subgroup of order 2^k is generated by g^((p-1)/2^k), g the field's
# multiplicative generator (BabyBear 31, Goldilocks 7); the reference takes
# its root from FRX's NTT, which is not visible here.

_GENERATOR = {"babybear_mont": 31, "goldilocks": 7}


def subgroup_root(dtype: Any, k: int) -> int:
    """A primitive 2^k-th root of unity of the prime field under `dtype`
    (canonical integer)."""
    f = as_field(dtype).base
    if (f.p - 1) % (1 << k):
        raise ValueError(f"{f.name} has no subgroup of order 2^{k}")
    return pow(_GENERATOR[f.name], (f.p - 1) >> k, f.p)


def _dft_last(values: FieldArray, inverse: bool) -> FieldArray:
    f = values.field
    if f.degree != 1:
        raise ValueError("subgroup transforms take base-field values")
    n = values.shape[-1]
    k = n.bit_length() - 1
    if n & (n - 1):
        raise ValueError(f"subgroup size {n} is not a power of two")
    w = subgroup_root(f, k) if n > 1 else 1
    if inverse:
        w = pow(w, f.p - 2, f.p)
    rows = [[int(x) for x in r] for r in values.canonical()[..., 0].reshape(-1, n).tolist()]
    ninv = pow(n, f.p - 2, f.p) if inverse else 1
    out = []
    for r in rows:
        row = []
        for j in range(n):
            wj, x, acc = pow(w, j, f.p), 1, 0
            for v in r:
                acc = (acc + v * x) % f.p
                x = x * wj % f.p
            row.append(acc * ninv % f.p)
        out.append(row)
    return array(out, f).reshape(values.shape)


def subgroup_to_coeffs(values: FieldArray) -> FieldArray:
    """Ascending coefficients of the degree < |D| univariate with `values` on
    the order-|D| subgroup (last axis; the iNTT, domain.py:245-252)."""
    return _dft_last(values, True)


def subgroup_evals(coeffs: FieldArray, size: int) -> FieldArray:
    """`coeffs` (last axis) evaluated on the order-`size` subgroup: zero-pad,
    then the NTT (domain.py:255-266)."""
    pad = coeffs.shape[-1]
    if size < pad:
        raise ValueError(f"subgroup_evals size {size} < coeff count {pad}")
    from zorch_amd.fields import zeros
    tail = zeros(coeffs.shape[:-1] + (size - pad,), coeffs.field)
    return _dft_last(concatenate([coeffs, tail], axis=-1), False)


def subgroup_sum(coeffs: FieldArray, skip_rounds: int) -> FieldArray:
    """sum over the order-2^skip_rounds subgroup D of the univariate `coeffs`:
    |D| times the sum of the coefficients at multiples of |D| (domain.py:
    269-278)."""
    size = 1 << skip_rounds
    f = coeffs.field
    acc = array(0, f)
    for i in range(0, coeffs.shape[0], size):
        acc = acc + coeffs[i]
    return acc * array(size, f)


@dataclass(frozen=True)
class UnivariateSkipDomain:
    """The order-2^skip_rounds subgroup D of the skip's round 0 (domain.py:
    281-311), with the value/coefficient maps carried by the transforms."""

    skip_rounds: int
    dtype: Any

    @property
    def size(self) -> int:
        return 1 << self.skip_rounds

    def nodes(self) -> FieldArray:
        """The |D| points in transform order: w^0, w^1, ..."""
        f = as_field(self.dtype).base
        w = subgroup_root(f, self.skip_rounds)
        return array([pow(w, j, f.p) for j in range(self.size)], f)

    def to_coeffs(self, values: FieldArray) -> FieldArray:
        return subgroup_to_coeffs(values)

    def sum_over_subgroup(self, coeffs: FieldArray) -> FieldArray:
        return subgroup_sum(coeffs, self.skip_rounds)


__all__ = [
    "EvalDomain", "natural_domain", "uhat_domain", "compressed_domain",
    "extend_to_round_domain", "split_halves", "fold", "interp_coeffs",
    "product_round_poly", "product_round_coeffs", "subgroup_root", "subgroup_to_coeffs",
    "subgroup_evals", "subgroup_sum", "UnivariateSkipDomain",
]
