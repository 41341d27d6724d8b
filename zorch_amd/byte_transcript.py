"""Byte-transcript wire vocabulary (reference zorch/byte_transcript.py:37-54).

The Merlin-style framing every field transcript puts on the stream: an op
tag, a kind tag, and lengths / counts as 8-byte little-endian integers.
"""

from __future__ import annotations

OP_DOMAIN = 0x01
OP_LABEL = 0x02
OP_OBSERVE = 0x03
OP_SQUEEZE = 0x04
OP_BYTES = 0x05
KIND_SCALAR = 0x01
KIND_SLICE = 0x02


def _len8(n: int) -> bytes:
    """A length / count as 8 little-endian bytes."""
    return int(n).to_bytes(8, "little")
