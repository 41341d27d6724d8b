"""Known-answer tests of the device field arithmetic (zs_field_op) on edge
operands, against Python big-integer arithmetic.

The kernels' products reduce BabyBear through a Montgomery REDC and
Goldilocks through 2^64 = 2^32 - 1, 2^96 = -1 with a borrow, a carry and a
final subtraction that fire about once in 2^32 random products, so random
tables almost never reach them. Operands here are built to reach each of
them (the branches are named in the comments of the Goldilocks selection);
non-canonical BabyBear words check the reduction of squeezed words.
"""

from __future__ import annotations

import ctypes
import random

import numpy as np
import pytest

from zorch_amd import _lib

BB_P = (1 << 31) - (1 << 27) + 1
GL_P = (1 << 64) - (1 << 32) + 1
R = 1 << 32
M64 = (1 << 64) - 1


def _gl_branches(a: int, b: int) -> set:
    """Which conditional steps of the 128 -> 64 bit Goldilocks reduction a
    product takes (lo - hi_hi borrows; the 2^64 (hi_lo) term carries; the
    result needs the final subtraction of p) — used only to pick operands."""
    x = a * b
    lo, hi = x & M64, x >> 64
    hi_hi, hi_lo = hi >> 32, hi & 0xFFFFFFFF
    out = set()
    t0 = (lo - hi_hi) & M64
    if lo < hi_hi:
        out.add("borrow")
        t0 = (t0 - 0xFFFFFFFF) & M64
    t1 = hi_lo * 0xFFFFFFFF
    t2 = (t0 + t1) & M64
    if t2 < t0:
        out.add("carry")
        t2 = (t2 + 0xFFFFFFFF) & M64
    if t2 >= GL_P:
        out.add("final")
    return out


def _gl_operands():
    edge = [0, 1, 2, 3, GL_P - 1, GL_P - 2, (1 << 32) - 1, 1 << 32, (1 << 32) + 1,
            1 << 63, (1 << 63) - 1, (1 << 63) + 1, 1 << 48, (1 << 64) - (1 << 33) + 1,
            0xFFFFFFFF00000000, 0x00000000FFFFFFFF, 0x8000000080000000]
    pairs = [(a, b) for a in edge for b in edge]
    rng = random.Random(1)
    pairs += [(rng.randrange(GL_P), rng.randrange(GL_P)) for _ in range(2000)]
    # Products that land just above a multiple of 2^64 near p: the final
    # subtraction and the carry; searched around b = k p / a.
    for a in range(2, 600):
        for k in range(1, a):
            b = (k * (1 << 64) + GL_P - k * 0xFFFFFFFF + a - 1) // a
            for d in range(-2, 3):
                if 0 <= b + d < GL_P:
                    pairs.append((a, b + d))
    hit = set()
    for a, b in pairs:
        hit |= _gl_branches(a, b)
    assert {"borrow", "carry", "final"} <= hit, hit
    return pairs


def _run(field_id: int, op: int, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    from zorch_amd.runtime import default_context
    ctx = default_context()
    out = np.zeros_like(a)
    _lib.check(_lib.lib().zs_field_op(ctx.handle, field_id, op, a.ctypes.data, b.ctypes.data,
                                      out.ctypes.data, a.shape[0]))
    return out


@pytest.mark.gpu
def test_goldilocks_add_sub_mul_edge_operands():
    pairs = _gl_operands()
    a = np.array([p[0] for p in pairs], dtype=np.uint64)
    b = np.array([p[1] for p in pairs], dtype=np.uint64)
    for op, f in [(_lib.ZS_OP_ADD, lambda x, y: (x + y) % GL_P),
                  (_lib.ZS_OP_SUB, lambda x, y: (x - y) % GL_P),
                  (_lib.ZS_OP_MUL, lambda x, y: (x * y) % GL_P)]:
        got = _run(_lib.ZS_F_GOLDILOCKS, op, a, b)
        want = [f(int(x), int(y)) for x, y in pairs]
        assert got.tolist() == want, op


def _bb_words():
    # Montgomery storage words: canonical edges and non-canonical words (the
    # kernels reduce a squeezed word before use).
    edge = [0, 1, 2, BB_P - 1, BB_P - 2, (BB_P - 1) // 2, (R % BB_P), (R * R) % BB_P,
            BB_P, BB_P + 1, (1 << 31), (1 << 32) - 1, 2 * BB_P - 1, 0x0FFFFFFE]
    rng = random.Random(2)
    return edge + [rng.randrange(R) for _ in range(200)]


def _bb_val(w: int) -> int:
    return (w % BB_P) * pow(R, -1, BB_P) % BB_P


def _bb_store(v: int) -> int:
    return v * R % BB_P


@pytest.mark.gpu
def test_babybear_montgomery_edge_operands():
    ws = _bb_words()
    pairs = [(x, y) for x in ws for y in ws]
    a = np.array([p[0] for p in pairs], dtype=np.uint32)
    b = np.array([p[1] for p in pairs], dtype=np.uint32)
    for op, f in [(_lib.ZS_OP_ADD, lambda x, y: x + y), (_lib.ZS_OP_SUB, lambda x, y: x - y),
                  (_lib.ZS_OP_MUL, lambda x, y: x * y)]:
        got = _run(_lib.ZS_F_BABYBEAR, op, a, b)
        want = [_bb_store(f(_bb_val(x), _bb_val(y)) % BB_P) for x, y in pairs]
        assert got.tolist() == want, op


def _bb4_mul(x, y):
    """(sum x_i X^i)(sum y_j X^j) mod (X^4 - 11) on canonical values."""
    out = [0] * 7
    for i in range(4):
        for j in range(4):
            out[i + j] += x[i] * y[j]
    return [(out[k] + 11 * (out[k + 4] if k + 4 < 7 else 0)) % BB_P for k in range(4)]


@pytest.mark.gpu
def test_babybear4_edge_operands():
    rng = random.Random(3)
    limbs = [0, 1, BB_P - 1, 2, 11, BB_P - 11]
    elems = [[a, b, c, d] for a in (0, 1, BB_P - 1) for b in (0, BB_P - 1)
             for c in (0, 1) for d in (0, 1, BB_P - 1)]
    # X, X^2, X^3 (the wrap X^4 = 11) and random elements
    elems += [[0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]]
    elems += [[rng.choice(limbs) for _ in range(4)] for _ in range(40)]
    elems += [[rng.randrange(BB_P) for _ in range(4)] for _ in range(40)]
    pairs = [(x, y) for x in elems for y in elems]
    a = np.array([[_bb_store(v) for v in p[0]] for p in pairs], dtype=np.uint32)
    b = np.array([[_bb_store(v) for v in p[1]] for p in pairs], dtype=np.uint32)
    for op, f in [(_lib.ZS_OP_ADD, lambda x, y: [(u + v) % BB_P for u, v in zip(x, y)]),
                  (_lib.ZS_OP_SUB, lambda x, y: [(u - v) % BB_P for u, v in zip(x, y)]),
                  (_lib.ZS_OP_MUL, _bb4_mul)]:
        got = _run(_lib.ZS_F_BABYBEAR4, op, a, b)
        want = [[_bb_store(v) for v in f(x, y)] for x, y in pairs]
        assert got.tolist() == want, op
    # X^3 * X = 11 (the reduction polynomial X^4 - 11)
    x3 = np.array([[0, 0, 0, _bb_store(1)]], dtype=np.uint32)
    x1 = np.array([[0, _bb_store(1), 0, 0]], dtype=np.uint32)
    assert _run(_lib.ZS_F_BABYBEAR4, _lib.ZS_OP_MUL, x3, x1).tolist() == \
        [[_bb_store(11), 0, 0, 0]]


def test_goldilocks_operand_selection_reaches_every_branch():
    """The operand search itself (host only): every conditional step of the
    reduction is reached."""
    _gl_operands()
