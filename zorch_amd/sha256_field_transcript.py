"""SHA-256 field transcript (drop-in for reference
zorch/sha256_field_transcript.py:156-378).

The squeeze is the counter chain SHA256(S || le64(ctr)), truncated and
re-absorbed (zorch/sha256_field_transcript.py:100-114). The state is the
streaming SHA-256 midstate, pending tail and byte count; the shared surface
is ``zorch_amd.field_transcript.FieldTranscript``.
"""

from __future__ import annotations

from dataclasses import dataclass

from zorch_amd import _lib
from zorch_amd.field_transcript import FieldTranscript


@dataclass(frozen=True, eq=False)
class Sha256FieldTranscript(FieldTranscript):
    KIND = _lib.ZS_TR_SHA256

    @property
    def has_dedicated_fusion(self) -> bool:
        return True

    def stream_state(self) -> tuple[tuple[int, ...], bytes, int]:
        """(SHA-256 midstate words, pending tail bytes, total bytes absorbed)."""
        s = self.state.sha
        return (tuple(int(x) for x in s.h), bytes(s.pending[: s.pending_len]),
                int(s.total_len))
