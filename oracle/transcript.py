"""Oracle SHA-256 field transcript (TEST INFRASTRUCTURE ONLY).

Restates ``Sha256FieldTranscript`` (reference
zorch/sha256_field_transcript.py:156-378) on an explicit byte stream with
``hashlib`` as the digest, i.e. the host ``ByteHashTranscript`` construction
(zorch/byte_transcript.py:85-171) it is byte-identical to:

* ``new(domain)``   S = [OP_DOMAIN] || le64(len) || domain          (:178-181)
* ``observe(V)``    S ||= [OP_OBSERVE, KIND_SLICE] || le64(|V|/s_T) || bytes(V)
                                                                     (:189-196)
* ``sample(n)``     S ||= [OP_SQUEEZE, KIND_SLICE] || le64(n);
                    out = SHA256(S || le64(0)) || SHA256(S || le64(1)) || ...
                    truncated to n*s_T; S ||= out                   (:251-259,
                                                                     :100-114)
Wire constants: zorch/byte_transcript.py:37-54.
"""

from __future__ import annotations

import hashlib
from dataclasses import dataclass

from oracle import blake3 as _b3
from oracle import sha256 as _sha

OP_DOMAIN, OP_LABEL, OP_OBSERVE, OP_SQUEEZE, OP_BYTES = 1, 2, 3, 4, 5
KIND_SCALAR, KIND_SLICE = 1, 2


def le64(n: int) -> bytes:
    return int(n).to_bytes(8, "little")


@dataclass(frozen=True)
class OracleTranscript:
    stream: bytes
    item_bytes: int  # storage bytes of one transcript-field word (s_T)

    @classmethod
    def new(cls, domain: bytes, item_bytes: int) -> "OracleTranscript":
        return cls(bytes([OP_DOMAIN]) + le64(len(domain)) + bytes(domain), item_bytes)

    def absorb(self, payload: bytes) -> "OracleTranscript":
        return type(self)(self.stream + payload, self.item_bytes)

    def observe(self, value_bytes: bytes) -> "OracleTranscript":
        count = len(value_bytes) // self.item_bytes
        return self.absorb(bytes([OP_OBSERVE, KIND_SLICE]) + le64(count) + value_bytes)

    def observe_scalar(self, value_bytes: bytes) -> "OracleTranscript":
        out = b""
        for i in range(0, len(value_bytes), self.item_bytes):
            out += bytes([OP_OBSERVE, KIND_SCALAR]) + value_bytes[i : i + self.item_bytes]
        return self.absorb(out)

    def observe_label(self, label: bytes) -> "OracleTranscript":
        return self.absorb(bytes([OP_LABEL]) + le64(len(label)) + bytes(label))

    def observe_bytes(self, data: bytes) -> "OracleTranscript":
        return self.absorb(bytes([OP_BYTES]) + le64(len(data)) + bytes(data))

    def _squeeze(self, nbytes: int) -> bytes:
        out, ctr = b"", 0
        while len(out) < nbytes:
            out += hashlib.sha256(self.stream + le64(ctr)).digest()
            ctr += 1
        return out[:nbytes]

    def sample(self, n: int) -> tuple["OracleTranscript", bytes]:
        t = self.absorb(bytes([OP_SQUEEZE, KIND_SLICE]) + le64(n))
        out = t._squeeze(n * self.item_bytes)
        return t.absorb(out), out

    def sample_scalar(self) -> tuple["OracleTranscript", bytes]:
        t = self.absorb(bytes([OP_SQUEEZE, KIND_SCALAR]))
        out = t._squeeze(self.item_bytes)
        return t.absorb(out), out

    def observe_and_sample(self, value_bytes: bytes, n: int):
        return self.observe(value_bytes).sample(n)

    def state(self) -> tuple[tuple[int, ...], bytes, int]:
        """Streaming state (midstate h, pending bytes, total length)."""
        return _sha.midstate(self.stream)


@dataclass(frozen=True)
class OracleBlake3Transcript(OracleTranscript):
    """``Blake3FieldTranscript`` (reference zorch/blake3_field_transcript.py:
    247-409): the same framing; a squeeze is the first n * s_T bytes of the
    BLAKE3 extendable output of S (:206-215), re-absorbed."""

    def _squeeze(self, nbytes: int) -> bytes:
        return _b3.blake3(self.stream, nbytes)

    def state(self) -> tuple:
        """Streaming state (chunk chaining value, subtree stack, current block,
        chunk index, blocks compressed in the chunk)."""
        s = _b3.Blake3Stream().absorb(self.stream)
        return (tuple(s.cv), tuple(tuple(c) for c in s.stack), bytes(s.block),
                s.chunk_counter, s.blocks_done)
