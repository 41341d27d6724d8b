"""BLAKE3 field transcript (drop-in for reference
zorch/blake3_field_transcript.py:247-409).

The squeeze reads the BLAKE3 extendable output of the whole stream, then
re-absorbs it (zorch/blake3_field_transcript.py:206-215, 320-330); there is
no counter chain. The state is the streaming BLAKE3 state (current chunk,
subtree stack); the shared surface is
``zorch_amd.field_transcript.FieldTranscript``. Proof-of-work grinding is not
part of the sumcheck path and is not provided.
"""

from __future__ import annotations

from dataclasses import dataclass

from zorch_amd import _lib
from zorch_amd.field_transcript import FieldTranscript

# The unpadded proof-of-work pre-image: state digest || le64 nonce
# (zorch/blake3_field_transcript.py:63-66).
_POW_PREIMAGE_BYTES = 40


@dataclass(frozen=True, eq=False)
class Blake3FieldTranscript(FieldTranscript):
    KIND = _lib.ZS_TR_BLAKE3

    @classmethod
    def new(cls, domain: bytes, dtype, *, pow_preimage_bytes: int = _POW_PREIMAGE_BYTES):
        if pow_preimage_bytes < _POW_PREIMAGE_BYTES:
            raise ValueError(
                f"pow_preimage_bytes must be at least {_POW_PREIMAGE_BYTES} — the nonce does "
                f"not fit below it — got {pow_preimage_bytes}")
        return super().new(domain, dtype)

    @property
    def has_dedicated_fusion(self) -> bool:
        return False

    def stream_state(self) -> tuple:
        """(chunk chaining value, subtree stack, current block bytes, chunk
        index, blocks of the chunk already compressed)."""
        s = self.state.blake3
        stack = tuple(tuple(int(x) for x in s.stack[i]) for i in range(s.stack_len))
        return (tuple(int(x) for x in s.cv), stack, bytes(s.block[: s.block_len]),
                int(s.chunk_counter), int(s.blocks_done))
