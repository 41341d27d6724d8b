"""Pure-Python SHA-256 compression (TEST INFRASTRUCTURE ONLY).

``hashlib`` gives digests but not the streaming midstate, and the transcript
state the reference threads is exactly that midstate plus the pending bytes
and the total length (hash-frx ``Sha256State`` = ``h``, ``pending``,
``counts``; reference zorch/sha256_field_transcript.py:45-50,108-114). This
module computes it from the FIPS 180-4 definition; ``tests/`` pins it to
``hashlib`` by finalizing the midstate and comparing digests.
"""

from __future__ import annotations

import struct

H0 = (
    0x6A09E667, 0xBB67AE85, 0x3C6EF372, 0xA54FF53A,
    0x510E527F, 0x9B05688C, 0x1F83D9AB, 0x5BE0CD19,
)

# FIPS 180-4 section 4.2.2: first 32 bits of the fractional parts of the cube
# roots of the first 64 primes, computed here rather than tabulated.


def _first_primes(n: int) -> list[int]:
    out, c = [], 2
    while len(out) < n:
        if all(c % q for q in out if q * q <= c):
            out.append(c)
        c += 1
    return out


def _icbrt(x: int) -> int:
    lo, hi = 0, 1 << ((x.bit_length() + 2) // 3 + 1)
    while lo < hi:
        mid = (lo + hi + 1) // 2
        if mid ** 3 <= x:
            lo = mid
        else:
            hi = mid - 1
    return lo


K = tuple(_icbrt(p << 96) & 0xFFFFFFFF for p in _first_primes(64))
_M = 0xFFFFFFFF


def _rotr(x: int, n: int) -> int:
    return ((x >> n) | (x << (32 - n))) & _M


def compress(h: tuple[int, ...], block: bytes) -> tuple[int, ...]:
    assert len(block) == 64
    w = list(struct.unpack(">16I", block))
    for t in range(16, 64):
        s0 = _rotr(w[t - 15], 7) ^ _rotr(w[t - 15], 18) ^ (w[t - 15] >> 3)
        s1 = _rotr(w[t - 2], 17) ^ _rotr(w[t - 2], 19) ^ (w[t - 2] >> 10)
        w.append((w[t - 16] + s0 + w[t - 7] + s1) & _M)
    a, b, c, d, e, f, g, hh = h
    for t in range(64):
        s1 = _rotr(e, 6) ^ _rotr(e, 11) ^ _rotr(e, 25)
        ch = (e & f) ^ (~e & g)
        t1 = (hh + s1 + ch + K[t] + w[t]) & _M
        s0 = _rotr(a, 2) ^ _rotr(a, 13) ^ _rotr(a, 22)
        mj = (a & b) ^ (a & c) ^ (b & c)
        t2 = (s0 + mj) & _M
        hh, g, f, e, d, c, b, a = g, f, e, (d + t1) & _M, c, b, a, (t1 + t2) & _M
    return tuple((x + y) & _M for x, y in zip(h, (a, b, c, d, e, f, g, hh)))


def midstate(stream: bytes) -> tuple[tuple[int, ...], bytes, int]:
    """(h after every full block, pending tail bytes, total length)."""
    h = H0
    full = len(stream) - len(stream) % 64
    for i in range(0, full, 64):
        h = compress(h, stream[i : i + 64])
    return h, stream[full:], len(stream)


def finalize(h: tuple[int, ...], pending: bytes, total: int) -> bytes:
    """Digest of the stream whose midstate is (h, pending, total)."""
    tail = pending + b"\x80"
    tail += b"\x00" * ((56 - len(tail)) % 64)
    tail += struct.pack(">Q", total * 8)
    for i in range(0, len(tail), 64):
        h = compress(h, tail[i : i + 64])
    return struct.pack(">8I", *h)
