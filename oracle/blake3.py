"""Pure-Python BLAKE3 (TEST INFRASTRUCTURE ONLY).

The reference's BLAKE3 field transcript (zorch/blake3_field_transcript.py:
206-215, 281-409) runs on hash-frx's streaming ``Blake3Stream`` (absent here;
hash-frx==0.2.0.dev20260911005315, requirements.in:39). This module restates
the published BLAKE3 specification (O'Connor, Aumasson, Neves, Wilcox-O'Hearn,
"BLAKE3: one function, fast everywhere", 2020, sections 2.1-2.6) in hash mode
twice, so that each checks the other:

* ``Blake3Stream`` — incremental: a chunk state plus a stack of subtree
  chaining values merged as chunks complete (the shape a resumable transcript
  state needs). A full block is compressed only once more input arrives,
  because the last block of a chunk carries CHUNK_END (and ROOT).
* ``blake3_tree`` — one-shot recursion over the binary tree: the left subtree
  of n > 1 chunks holds the largest power of two of chunks below n.

``tests/test_oracle.py`` pins them to the published vectors the reference's
own tests quote (zorch/testkit/testing/byte_hash_test.py:24-29) and to each
other on many lengths and random absorb splits.
"""

from __future__ import annotations

import struct
from dataclasses import dataclass, field

IV = (0x6A09E667, 0xBB67AE85, 0x3C6EF372, 0xA54FF53A,
      0x510E527F, 0x9B05688C, 0x1F83D9AB, 0x5BE0CD19)
MSG_PERMUTATION = (2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8)
CHUNK_START, CHUNK_END, PARENT, ROOT = 1, 2, 4, 8
BLOCK_LEN, CHUNK_LEN = 64, 1024
MASK = 0xFFFFFFFF


def _rotr(x: int, n: int) -> int:
    return ((x >> n) | (x << (32 - n))) & MASK


def _g(v: list, a: int, b: int, c: int, d: int, mx: int, my: int) -> None:
    v[a] = (v[a] + v[b] + mx) & MASK
    v[d] = _rotr(v[d] ^ v[a], 16)
    v[c] = (v[c] + v[d]) & MASK
    v[b] = _rotr(v[b] ^ v[c], 12)
    v[a] = (v[a] + v[b] + my) & MASK
    v[d] = _rotr(v[d] ^ v[a], 8)
    v[c] = (v[c] + v[d]) & MASK
    v[b] = _rotr(v[b] ^ v[c], 7)


def compress(cv, block: bytes, counter: int, block_len: int, flags: int) -> tuple:
    """The compression function: 16 output words (the first 8 are the new
    chaining value; all 16 are extended output at the root)."""
    m = list(struct.unpack("<16I", block.ljust(BLOCK_LEN, b"\0")))
    v = list(cv) + list(IV[:4]) + [counter & MASK, (counter >> 32) & MASK, block_len, flags]
    for rnd in range(7):
        _g(v, 0, 4, 8, 12, m[0], m[1])
        _g(v, 1, 5, 9, 13, m[2], m[3])
        _g(v, 2, 6, 10, 14, m[4], m[5])
        _g(v, 3, 7, 11, 15, m[6], m[7])
        _g(v, 0, 5, 10, 15, m[8], m[9])
        _g(v, 1, 6, 11, 12, m[10], m[11])
        _g(v, 2, 7, 8, 13, m[12], m[13])
        _g(v, 3, 4, 9, 14, m[14], m[15])
        if rnd < 6:
            m = [m[i] for i in MSG_PERMUTATION]
    for i in range(8):
        v[i] ^= v[i + 8]
        v[i + 8] ^= cv[i]
    return tuple(v)


@dataclass(frozen=True)
class Output:
    """A node awaiting finalisation: its chaining value, or the root XOF."""

    cv: tuple
    block: bytes
    counter: int
    block_len: int
    flags: int

    def chaining_value(self) -> tuple:
        return compress(self.cv, self.block, self.counter, self.block_len, self.flags)[:8]

    def root_bytes(self, nbytes: int) -> bytes:
        out, ctr = b"", 0
        while len(out) < nbytes:
            words = compress(self.cv, self.block, ctr, self.block_len, self.flags | ROOT)
            out += struct.pack("<16I", *words)
            ctr += 1
        return out[:nbytes]


def _parent(left: tuple, right: tuple) -> Output:
    return Output(IV, struct.pack("<16I", *left, *right), 0, BLOCK_LEN, PARENT)


@dataclass
class Blake3Stream:
    """Incremental hash-mode BLAKE3; ``copy`` forks it (transcripts are values)."""

    cv: tuple = IV
    chunk_counter: int = 0
    block: bytes = b""
    blocks_done: int = 0
    stack: list = field(default_factory=list)

    def copy(self) -> "Blake3Stream":
        return Blake3Stream(self.cv, self.chunk_counter, self.block, self.blocks_done,
                            list(self.stack))

    def _chunk_len(self) -> int:
        return BLOCK_LEN * self.blocks_done + len(self.block)

    def _chunk_output(self) -> Output:
        flags = CHUNK_END | (CHUNK_START if self.blocks_done == 0 else 0)
        return Output(self.cv, self.block, self.chunk_counter, len(self.block), flags)

    def absorb(self, data: bytes) -> "Blake3Stream":
        data = bytes(data)
        while data:
            if self._chunk_len() == CHUNK_LEN:
                # The chunk is complete and more input follows: its chaining
                # value joins the stack, merging every completed pair.
                new_cv = self._chunk_output().chaining_value()
                total = self.chunk_counter + 1
                while total & 1 == 0:
                    new_cv = _parent(self.stack.pop(), new_cv).chaining_value()
                    total >>= 1
                self.stack.append(new_cv)
                self.cv, self.chunk_counter = IV, self.chunk_counter + 1
                self.block, self.blocks_done = b"", 0
            if len(self.block) == BLOCK_LEN:
                flags = CHUNK_START if self.blocks_done == 0 else 0
                self.cv = compress(self.cv, self.block, self.chunk_counter, BLOCK_LEN, flags)[:8]
                self.blocks_done += 1
                self.block = b""
            take = min(BLOCK_LEN - len(self.block), len(data))
            self.block += data[:take]
            data = data[take:]
        return self

    def finalize(self, nbytes: int) -> bytes:
        """XOF output of everything absorbed; the stream is unchanged."""
        out = self._chunk_output()
        for cv in reversed(self.stack):
            out = _parent(cv, out.chaining_value())
        return out.root_bytes(nbytes)


def _tree_output(data: bytes, chunk0: int) -> Output:
    nchunks = max(1, -(-len(data) // CHUNK_LEN))
    if nchunks == 1:
        cv, done = IV, 0
        blocks = [data[i:i + BLOCK_LEN] for i in range(0, len(data), BLOCK_LEN)] or [b""]
        for blk in blocks[:-1]:
            cv = compress(cv, blk, chunk0, BLOCK_LEN, CHUNK_START if done == 0 else 0)[:8]
            done += 1
        last = blocks[-1]
        return Output(cv, last, chunk0, len(last), CHUNK_END | (CHUNK_START if done == 0 else 0))
    left = 1 << ((nchunks - 1).bit_length() - 1)  # largest power of two below nchunks
    lcv = _tree_output(data[:left * CHUNK_LEN], chunk0).chaining_value()
    rcv = _tree_output(data[left * CHUNK_LEN:], chunk0 + left).chaining_value()
    return _parent(lcv, rcv)


def blake3_tree(data: bytes, nbytes: int = 32) -> bytes:
    """One-shot BLAKE3 of `data` by recursion over the chunk tree."""
    return _tree_output(bytes(data), 0).root_bytes(nbytes)


def blake3(data: bytes, nbytes: int = 32) -> bytes:
    return Blake3Stream().absorb(data).finalize(nbytes)
