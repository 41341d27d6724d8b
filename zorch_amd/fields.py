"""Field dtypes and host field arrays.

The reference computes in third-party field dtypes (``zk_dtypes``
``babybear_mont``, ``babybearx4_mont``, ``goldilocks``; call sites
zorch/sumcheck/domain.py:101-102, zorch/poly/univariate.py:55-65). Here a
field is a small descriptor and a :class:`FieldArray` is a numpy array of
*storage words* (the bits the reference's arrays hold):

* ``babybear_mont``    one u32 Montgomery word per element (a * 2^32 mod p)
* ``babybearx4_mont``  four such words c0..c3, value in F_p[X]/(X^4 - 11)
* ``goldilocks``       one canonical u64 per element

Storage is kept as ``words`` with a trailing limb axis (length 1 for prime
fields). Host arithmetic (used for claims, test helpers and the element-wise
operators) is vectorised numpy on canonical residues; the prover's arithmetic
runs on the GPU in ``libzorch_amd.so``.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Iterable, Sequence

import numpy as np

from zorch_amd import _lib

BB_P = (1 << 31) - (1 << 27) + 1
GL_P = (1 << 64) - (1 << 32) + 1
_R = 1 << 32
_BB_RINV = pow(_R, -1, BB_P)
_M32 = np.uint64(0xFFFFFFFF)


@dataclass(frozen=True)
class Field:
    name: str
    fid: int
    word_dtype: Any
    words: int          # storage words per element
    degree: int         # extension degree over the prime field
    p: int
    montgomery: bool

    @property
    def itemsize(self) -> int:
        return self.words * np.dtype(self.word_dtype).itemsize

    @property
    def base(self) -> "Field":
        return babybear_mont if self is babybearx4_mont else self

    def __repr__(self) -> str:
        return self.name


babybear_mont = Field("babybear_mont", _lib.ZS_F_BABYBEAR, np.uint32, 1, 1, BB_P, True)
babybearx4_mont = Field("babybearx4_mont", _lib.ZS_F_BABYBEAR4, np.uint32, 4, 4, BB_P, True)
goldilocks = Field("goldilocks", _lib.ZS_F_GOLDILOCKS, np.uint64, 1, 1, GL_P, False)
FIELDS = {f.name: f for f in (babybear_mont, babybearx4_mont, goldilocks)}


def as_field(dtype: Any) -> Field:
    if isinstance(dtype, Field):
        return dtype
    if isinstance(dtype, str) and dtype in FIELDS:
        return FIELDS[dtype]
    raise ValueError(f"unknown field dtype {dtype!r}")


def base_field(dtype: Any) -> Field:
    """The prime field under `dtype` (reference zorch/utils/field.py:15-22)."""
    return as_field(dtype).base


def challenge_limbs(dtype: Any) -> int:
    """Base-field squeezes per element (reference zorch/challenge.py:16-21)."""
    return as_field(dtype).degree


# ------------------------------------------------------------ codecs (numpy)

def _bb_decode(words: np.ndarray) -> np.ndarray:
    """Montgomery words (any u32, reduced or not) -> canonical residues (u64)."""
    w = words.astype(np.uint64) % np.uint64(BB_P)
    return (w * np.uint64(_BB_RINV)) % np.uint64(BB_P)


def _bb_encode(vals: np.ndarray) -> np.ndarray:
    v = vals.astype(np.uint64) % np.uint64(BB_P)
    return ((v << np.uint64(32)) % np.uint64(BB_P)).astype(np.uint32)


def decode(field: Field, words: np.ndarray) -> np.ndarray:
    """Storage words (..., words) -> canonical residues (..., words) as u64."""
    if field.montgomery:
        return _bb_decode(words)
    w = words.astype(np.uint64)
    return np.where(w >= np.uint64(GL_P), w - np.uint64(GL_P), w)


def encode(field: Field, vals: np.ndarray) -> np.ndarray:
    if field.montgomery:
        return _bb_encode(vals)
    return vals.astype(np.uint64)


# -------------------------------------------------- canonical arithmetic

def _gl_reduce(lo: np.ndarray, hi: np.ndarray) -> np.ndarray:
    eps = _M32
    hi_hi = hi >> np.uint64(32)
    hi_lo = hi & eps
    t0 = lo - hi_hi
    t0 = np.where(lo < hi_hi, t0 - eps, t0)
    t1 = hi_lo * eps
    t2 = t0 + t1
    t2 = np.where(t2 < t0, t2 + eps, t2)
    return np.where(t2 >= np.uint64(GL_P), t2 - np.uint64(GL_P), t2)


def _gl_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        a0, a1 = a & _M32, a >> np.uint64(32)
        b0, b1 = b & _M32, b >> np.uint64(32)
        lo = a0 * b0
        m1, m2 = a0 * b1, a1 * b0
        mid = m1 + m2
        mid_c = (mid < m1).astype(np.uint64)
        hi = a1 * b1
        m_lo = (mid & _M32) << np.uint64(32)
        x_lo = lo + m_lo
        c1 = (x_lo < lo).astype(np.uint64)
        x_hi = hi + (mid >> np.uint64(32)) + (mid_c << np.uint64(32)) + c1
        return _gl_reduce(x_lo, x_hi)


def c_add(field: Field, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    p = np.uint64(field.p)
    with np.errstate(over="ignore"):
        s = a + b
        if field is goldilocks:
            s = np.where(s < a, s + _M32, s)
        return np.where(s >= p, s - p, s)


def c_sub(field: Field, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    p = np.uint64(field.p)
    with np.errstate(over="ignore"):
        return np.where(a >= b, a - b, a + (p - b))


_BB4_SMALL = 1024
# _BB4_FOLD[4 i + j, k]: the weight of a_i b_j in limb k of a product in
# BB[X] / (X^4 - 11).
_BB4_FOLD = np.zeros((16, 4), dtype=np.uint64)
for _i in range(4):
    for _j in range(4):
        _BB4_FOLD[4 * _i + _j, (_i + _j) % 4] = 11 if _i + _j >= 4 else 1


def c_mul(field: Field, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Element-wise product of canonical arrays with trailing limb axes."""
    p = np.uint64(BB_P)
    if field is goldilocks:
        return _gl_mul(a, b)
    if field is babybear_mont:
        return (a * b) % p
    # X^4 = 11 quartic extension, schoolbook on canonical limbs.
    if a.shape == (4,) and b.shape == (4,):
        x, y, P = a.tolist(), b.tolist(), BB_P
        out = [0, 0, 0, 0]
        for i in range(4):
            for j in range(4):
                out[(i + j) & 3] += x[i] * y[j] * (11 if i + j >= 4 else 1)
        return np.array([v % P for v in out], dtype=np.uint64)
    a = np.broadcast_to(a, np.broadcast_shapes(a.shape, b.shape))
    b = np.broadcast_to(b, a.shape)
    if a.size <= _BB4_SMALL:
        # Few elements (challenges, claims, openings): the sixteen limb
        # products at once, each reduced below p, then collected into the
        # four output limbs by the weights of X^(i+j) mod X^4 - 11 (sums stay
        # below 2^38).
        prod = (a[..., :, None] * b[..., None, :]) % p
        return (prod.reshape(a.shape[:-1] + (16,)) @ _BB4_FOLD) % p
    out = np.zeros(a.shape, dtype=np.uint64)
    for k in range(4):
        acc = np.zeros(a.shape[:-1], dtype=np.uint64)
        for i in range(4):
            j = (k - i) % 4
            term = (a[..., i] * b[..., j]) % p
            if i + j >= 4:
                term = (term * np.uint64(11)) % p
            acc = (acc + term) % p
        out[..., k] = acc
    return out


def c_sum(field: Field, a: np.ndarray, axis: int = 0) -> np.ndarray:
    """Field sum over a leading axis of a canonical (..., words) array."""
    if field is goldilocks:
        lo = np.sum(a & _M32, axis=axis, dtype=np.uint64)
        hi = np.sum(a >> np.uint64(32), axis=axis, dtype=np.uint64)
        flat_lo, flat_hi = np.atleast_1d(lo).ravel(), np.atleast_1d(hi).ravel()
        res = np.array([((int(h) << 32) + int(l)) % GL_P for l, h in zip(flat_lo, flat_hi)],
                       dtype=np.uint64).reshape(np.shape(lo))
        return res
    # BabyBear limbs: residues < 2^31; chunk so partial sums stay < 2^64.
    a = np.moveaxis(a, axis, 0)
    acc = np.zeros(a.shape[1:], dtype=np.uint64)
    step = 1 << 30
    for i in range(0, a.shape[0], step):
        acc = (acc + np.sum(a[i:i + step], axis=0, dtype=np.uint64)) % np.uint64(BB_P)
    return acc


def c_inv(field: Field, a: np.ndarray) -> np.ndarray:
    flat = a.reshape(-1, field.words)
    out = np.empty_like(flat)
    for i, row in enumerate(flat):
        out[i] = _inv_int(field, [int(x) for x in row])
    return out.reshape(a.shape)


def _inv_int(field: Field, limbs: list[int]) -> list[int]:
    if field.degree == 1:
        if limbs[0] % field.p == 0:
            raise ZeroDivisionError("inverse of zero")
        return [pow(limbs[0], field.p - 2, field.p)]
    if not any(limbs):
        raise ZeroDivisionError("inverse of zero")
    e = field.p ** 4 - 2
    res = np.array([1, 0, 0, 0], dtype=np.uint64)
    base = np.array(limbs, dtype=np.uint64)
    while e:
        if e & 1:
            res = c_mul(field, res, base)
        base = c_mul(field, base, base)
        e >>= 1
    return [int(x) for x in res]


# ------------------------------------------------------------- FieldArray

class FieldArray:
    """A host array of field elements in storage encoding.

    ``words`` has shape ``(*shape, field.words)``. Immutable by convention:
    every operation returns a new array, as the reference's arrays do.
    """

    __slots__ = ("words", "field")

    def __init__(self, words: np.ndarray, field: Any):
        field = as_field(field)
        words = np.asarray(words, dtype=field.word_dtype)
        if words.ndim == 0 or words.shape[-1] != field.words:
            raise ValueError(f"storage for {field.name} needs a trailing axis of {field.words}")
        self.words = words
        self.field = field

    # --- shape ---
    @property
    def shape(self) -> tuple[int, ...]:
        return tuple(self.words.shape[:-1])

    @property
    def ndim(self) -> int:
        return self.words.ndim - 1

    @property
    def dtype(self) -> Field:
        return self.field

    def __len__(self) -> int:
        return self.shape[0]

    def __getitem__(self, idx) -> "FieldArray":
        if not isinstance(idx, tuple):
            idx = (idx,)
        return FieldArray(self.words[idx + (slice(None),)], self.field)

    def __iter__(self):
        for i in range(len(self)):
            yield self[i]

    def reshape(self, *shape) -> "FieldArray":
        if len(shape) == 1 and isinstance(shape[0], tuple):
            shape = shape[0]
        return FieldArray(self.words.reshape(*shape, self.field.words), self.field)

    # --- values ---
    def canonical(self) -> np.ndarray:
        """Canonical residues, shape (*shape, words), u64."""
        return decode(self.field, self.words)

    def normalized(self) -> "FieldArray":
        """The same values with every storage word reduced."""
        return FieldArray(encode(self.field, self.canonical()), self.field)

    def reduced(self) -> "FieldArray":
        """The same values with every storage word below p: this array when
        it already is (one comparison pass), else ``normalized()``."""
        if bool((self.words < np.asarray(self.field.p, dtype=self.words.dtype)).all()):
            return self
        return self.normalized()

    def tobytes(self) -> bytes:
        return np.ascontiguousarray(self.words).astype(
            np.dtype(self.field.word_dtype).newbyteorder("<")).tobytes()

    def to_ints(self):
        """Python values: ints (prime fields) or 4-tuples (extension)."""
        can = self.canonical()
        flat = can.reshape(-1, self.field.words)
        vals = [int(r[0]) if self.field.words == 1 else tuple(int(x) for x in r) for r in flat]
        if self.ndim == 0:
            return vals[0]
        return np.array(vals, dtype=object).reshape(self.shape).tolist() \
            if self.field.words == 1 else _nest(vals, self.shape)

    def astype(self, field: Any) -> "FieldArray":
        """Embed into an extension (the base field's limbs go to c0)."""
        field = as_field(field)
        if field is self.field:
            return self
        if self.field is babybear_mont and field is babybearx4_mont:
            out = np.zeros(self.shape + (4,), dtype=np.uint32)
            out[..., 0] = self.words[..., 0]
            return FieldArray(out, field)
        raise ValueError(f"cannot convert {self.field.name} to {field.name}")

    def view(self, field: Any) -> "FieldArray":
        """Reinterpret consecutive words as elements of another field over the
        same storage words (reference ``reinterpret_challenge``,
        zorch/transcript.py:97-119)."""
        field = as_field(field)
        flat = self.words.reshape(-1)
        if flat.dtype != np.dtype(field.word_dtype) or flat.size % field.words:
            raise ValueError(f"{self.shape} elements of {self.field.name} do not view as "
                             f"{field.name}")
        return FieldArray(flat.reshape(-1, field.words), field)

    # --- comparisons / arithmetic ---
    def _coerce(self, other) -> "FieldArray":
        if isinstance(other, FieldArray):
            if other.field is not self.field:
                if other.field is self.field.base:
                    return other.astype(self.field)
                raise ValueError("field mismatch")
            return other
        return array(np.asarray(other), self.field)

    def __eq__(self, other) -> np.ndarray:  # type: ignore[override]
        o = self._coerce(other)
        return np.all(self.canonical() == o.canonical(), axis=-1)

    def __ne__(self, other) -> np.ndarray:  # type: ignore[override]
        return ~self.__eq__(other)

    def __add__(self, other) -> "FieldArray":
        o = self._coerce(other)
        return _from_canon(self.field, c_add(self.field, self.canonical(), o.canonical()))

    def __sub__(self, other) -> "FieldArray":
        o = self._coerce(other)
        return _from_canon(self.field, c_sub(self.field, self.canonical(), o.canonical()))

    def __mul__(self, other) -> "FieldArray":
        o = self._coerce(other)
        return _from_canon(self.field, c_mul(self.field, self.canonical(), o.canonical()))

    __radd__ = __add__
    __rmul__ = __mul__

    def __rsub__(self, other) -> "FieldArray":
        return self._coerce(other) - self

    def __hash__(self):
        return id(self)

    def __repr__(self) -> str:
        return f"FieldArray({self.field.name}, shape={self.shape})"

    def at_set(self, idx, value) -> "FieldArray":
        """Functional update (the reference's ``x.at[idx].set(v)``)."""
        w = self.words.copy()
        v = self._coerce(value)
        if not isinstance(idx, tuple):
            idx = (idx,)
        w[idx + (slice(None),)] = v.words
        return FieldArray(w, self.field)


def _nest(vals, shape):
    if len(shape) == 1:
        return vals
    step = int(np.prod(shape[1:]))
    return [_nest(vals[i * step:(i + 1) * step], shape[1:]) for i in range(shape[0])]


def _from_canon(field: Field, can: np.ndarray) -> FieldArray:
    return FieldArray(encode(field, can), field)


def array(values, field: Any) -> FieldArray:
    """Field array from canonical integer values (encoded to storage).

    `values`: ints (prime field) with any shape, or for the extension either
    ints (embedded into c0) or a trailing axis of 4 limbs given as tuples.
    """
    field = as_field(field)
    if isinstance(values, FieldArray):
        return values.astype(field)
    if field.words == 1:
        arr = np.asarray(values, dtype=object)
        can = np.vectorize(lambda x: int(x) % field.p, otypes=[object])(arr) if arr.size else arr
        can = np.asarray(can, dtype=np.uint64)[..., None]
        return _from_canon(field, can)
    arr = np.asarray(values, dtype=object)
    if arr.ndim == 0 or arr.shape[-1] != 4:
        base = np.asarray(np.vectorize(lambda x: int(x) % field.p, otypes=[object])(arr)
                          if arr.size else arr, dtype=np.uint64)
        can = np.zeros(base.shape + (4,), dtype=np.uint64)
        can[..., 0] = base
    else:
        can = np.asarray(np.vectorize(lambda x: int(x) % field.p, otypes=[object])(arr),
                         dtype=np.uint64)
    return _from_canon(field, can)


def from_storage(words, field: Any) -> FieldArray:
    """Field array over given storage words; prime fields may omit the limb axis."""
    field = as_field(field)
    w = np.asarray(words, dtype=field.word_dtype)
    if field.words == 1 and (w.ndim == 0 or w.shape[-1] != 1):
        w = w[..., None]
    return FieldArray(w, field)


def from_bytes(data: bytes, field: Any, shape: Sequence[int] | None = None) -> FieldArray:
    field = as_field(field)
    w = np.frombuffer(data, dtype=np.dtype(field.word_dtype).newbyteorder("<")).astype(
        field.word_dtype)
    n = w.size // field.words
    shape = (n,) if shape is None else tuple(shape)
    return FieldArray(w.reshape(*shape, field.words), field)


def zeros(shape, field: Any) -> FieldArray:
    field = as_field(field)
    shape = (shape,) if isinstance(shape, int) else tuple(shape)
    return FieldArray(np.zeros(shape + (field.words,), dtype=field.word_dtype), field)


def ones(shape, field: Any) -> FieldArray:
    field = as_field(field)
    shape = (shape,) if isinstance(shape, int) else tuple(shape)
    can = np.zeros(shape + (field.words,), dtype=np.uint64)
    can[..., 0] = 1
    return _from_canon(field, can)


def stack(arrays: Iterable[FieldArray], axis: int = 0) -> FieldArray:
    arrays = list(arrays)
    if not arrays:
        raise ValueError("need at least one array to stack")
    field = arrays[0].field
    for a in arrays:
        if a.field is not field:
            raise ValueError("stack of mixed fields")
    if axis < 0:
        axis += arrays[0].ndim + 1
    return FieldArray(np.stack([a.words for a in arrays], axis=axis), field)


def fsum(a: FieldArray, axis: int = 0) -> FieldArray:
    return _from_canon(a.field, c_sum(a.field, a.canonical(), axis=axis))


def fprod(a: FieldArray, axis: int = 0) -> FieldArray:
    can = np.moveaxis(a.canonical(), axis, 0)
    out = can[0]
    for i in range(1, can.shape[0]):
        out = c_mul(a.field, out, can[i])
    return _from_canon(a.field, out)


def inv(a: FieldArray) -> FieldArray:
    return _from_canon(a.field, c_inv(a.field, a.canonical()))


def naturals(n: int, dtype: Any) -> FieldArray:
    """[0, 1, ..., n-1] in the base field of `dtype` (zorch/utils/field.py:25-34)."""
    return array(list(range(n)), base_field(dtype))


def concatenate(arrays: Iterable[FieldArray], axis: int = 0) -> FieldArray:
    arrays = list(arrays)
    field = arrays[0].field
    if axis < 0:
        axis += arrays[0].ndim
    return FieldArray(np.concatenate([a.words for a in arrays], axis=axis), field)
