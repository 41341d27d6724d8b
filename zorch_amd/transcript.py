"""The ``Transcript`` seam and challenge reinterpretation.

Mirrors reference zorch/transcript.py:66-137: a transcript squeezes words of
its own field; a challenge in an extension is packed from consecutive words.
"""

from __future__ import annotations

from typing import Any, Protocol, TypeVar

from zorch_amd.fields import FieldArray, as_field


class Transcript(Protocol):
    @property
    def field(self) -> Any: ...

    @property
    def has_dedicated_fusion(self) -> bool: ...

    def observe(self, values: FieldArray) -> "Transcript": ...

    def sample(self, n: int = 1) -> tuple["Transcript", FieldArray]: ...

    def observe_and_sample(self, values: FieldArray, n: int = 1) -> tuple["Transcript", FieldArray]: ...


TranscriptT = TypeVar("TranscriptT", bound=Transcript)


def reinterpret_challenge(raw: FieldArray, dtype: Any) -> FieldArray:
    """The `dtype` element whose storage words are the squeezes `raw`
    (reference zorch/transcript.py:97-119). Raises if they do not make exactly
    one element."""
    field = as_field(dtype)
    viewed = raw.view(field)
    if viewed.shape != (1,):
        raise ValueError(
            f"{raw.shape[0]} squeezes reinterpret to {viewed.shape} elements of "
            f"{field.name}; a challenge needs exactly one")
    return viewed[0]


def sample_challenge(transcript, dtype: Any, limbs: int = 1):
    """Squeeze one `dtype` challenge as `limbs` transcript words
    (reference zorch/transcript.py:122-137)."""
    if limbs < 1:
        raise ValueError(f"limbs must be >= 1, got {limbs}")
    transcript, raw = transcript.sample(limbs)
    return transcript, reinterpret_challenge(raw, dtype)
