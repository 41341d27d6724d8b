"""A second, independently written BLAKE3 (hash mode and XOF) for checking
``oracle/blake3.py`` beyond the one-block inputs the published vectors cover
(TEST INFRASTRUCTURE ONLY).

It shares no code and no constant tables with the oracle:

* the compression keeps the state as four rows of four words and does the
  diagonal step by rotating rows 1-3 (the SIMD formulation of the BLAKE3
  paper, section 2.2) instead of naming the sixteen G calls; the message
  words of all seven rounds come from one precomputed schedule table of word
  indices rather than re-permuting the block each round;
* the tree is built bottom-up: chunk chaining values are paired left to
  right, an odd last node is carried up unchanged, and the final pair is
  compressed with ROOT (section 2.5 describes the same left-complete tree
  top-down: the left subtree holds the largest power of two of chunks that
  leaves at least one for the right);
* chunks are split by index arithmetic over the whole input, with no
  streaming state.

Mistakes the two could still share are those of reading the specification
the same wrong way; what this one removes is the shared code.
"""

from __future__ import annotations

_MASK = (1 << 32) - 1

# The eight words of the initialisation vector (those of SHA-256).
_IV = [0x6A09E667, 0xBB67AE85, 0x3C6EF372, 0xA54FF53A,
       0x510E527F, 0x9B05688C, 0x1F83D9AB, 0x5BE0CD19]

_F_CHUNK_START = 1 << 0
_F_CHUNK_END = 1 << 1
_F_PARENT = 1 << 2
_F_ROOT = 1 << 3


def _schedule() -> list[list[int]]:
    """Word indices of the block that round r consumes, in consumption order
    (round 0 takes them in order; each later round applies the fixed
    permutation of the specification, table 2 of the paper, to the previous
    round's order)."""
    perm = {0: 2, 1: 6, 2: 3, 3: 10, 4: 7, 5: 0, 6: 4, 7: 13,
            8: 1, 9: 11, 10: 12, 11: 5, 12: 9, 13: 14, 14: 15, 15: 8}
    rounds = [list(range(16))]
    for _ in range(6):
        prev = rounds[-1]
        rounds.append([prev[perm[i]] for i in range(16)])
    return rounds


_SCHEDULE = _schedule()


def _rot(x: int, n: int) -> int:
    return ((x >> n) | (x << (32 - n))) & _MASK


def _quarter(rows: list[list[int]], col: int, x: int, y: int) -> None:
    a, b, c, d = rows[0][col], rows[1][col], rows[2][col], rows[3][col]
    a = (a + b + x) & _MASK
    d = _rot(d ^ a, 16)
    c = (c + d) & _MASK
    b = _rot(b ^ c, 12)
    a = (a + b + y) & _MASK
    d = _rot(d ^ a, 8)
    c = (c + d) & _MASK
    b = _rot(b ^ c, 7)
    rows[0][col], rows[1][col], rows[2][col], rows[3][col] = a, b, c, d


def _shift_rows(rows: list[list[int]], back: bool) -> None:
    """Rotate row i left by i words (diagonals become columns); back undoes it."""
    for i in (1, 2, 3):
        k = (4 - i) % 4 if back else i
        rows[i] = rows[i][k:] + rows[i][:k]


def compress_words(cv: list[int], block: bytes, counter: int, block_len: int,
                   flags: int) -> list[int]:
    """All sixteen output words of one compression."""
    padded = bytes(block) + bytes(64 - len(block))
    m = [int.from_bytes(padded[4 * i:4 * i + 4], "little") for i in range(16)]
    rows = [list(cv[0:4]), list(cv[4:8]), list(_IV[0:4]),
            [counter & _MASK, counter >> 32, block_len, flags]]
    for order in _SCHEDULE:
        w = [m[i] for i in order]
        for col in range(4):
            _quarter(rows, col, w[2 * col], w[2 * col + 1])
        _shift_rows(rows, back=False)
        for col in range(4):
            _quarter(rows, col, w[8 + 2 * col], w[8 + 2 * col + 1])
        _shift_rows(rows, back=True)
    v = rows[0] + rows[1] + rows[2] + rows[3]
    return [v[i] ^ v[i + 8] for i in range(8)] + [v[i + 8] ^ cv[i] for i in range(8)]


def _le_bytes(words: list[int]) -> bytes:
    return b"".join(w.to_bytes(4, "little") for w in words)


def _chunk_node(chunk: bytes, index: int):
    """(chaining value input, last block, counter, length, flags) of a chunk:
    every block but the last is compressed here; the last is left to the
    caller, which may add ROOT."""
    blocks = [chunk[i:i + 64] for i in range(0, len(chunk), 64)] or [b""]
    cv = list(_IV)
    for j, blk in enumerate(blocks[:-1]):
        cv = compress_words(cv, blk, index, 64, _F_CHUNK_START if j == 0 else 0)[:8]
    start = _F_CHUNK_START if len(blocks) == 1 else 0
    return cv, blocks[-1], index, len(blocks[-1]), start | _F_CHUNK_END


def blake3_xof(data: bytes, nbytes: int = 32) -> bytes:
    data = bytes(data)
    nchunks = max(1, (len(data) + 1023) // 1024)
    nodes = [_chunk_node(data[1024 * i:1024 * (i + 1)], i) for i in range(nchunks)]
    while len(nodes) > 1:
        cvs = [compress_words(*n)[:8] if not isinstance(n, list) else n for n in nodes]
        if len(cvs) == 2:
            nodes = [(list(_IV), _le_bytes(cvs[0] + cvs[1]), 0, 64, _F_PARENT)]
            break
        nxt = []
        for i in range(0, len(cvs) - 1, 2):
            nxt.append(compress_words(list(_IV), _le_bytes(cvs[i] + cvs[i + 1]), 0, 64,
                                      _F_PARENT)[:8])
        if len(cvs) % 2:
            nxt.append(cvs[-1])
        nodes = nxt
    node = nodes[0]
    if isinstance(node, list):  # cannot happen: a lone node is a chunk or the root pair
        raise AssertionError("root node lost its inputs")
    cv, block, counter, block_len, flags = node
    out = b""
    t = 0
    while len(out) < nbytes:
        out += _le_bytes(compress_words(cv, block, t, block_len, flags | _F_ROOT))
        t += 1
    return out[:nbytes]
