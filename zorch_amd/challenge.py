"""Fiat-Shamir challenge-field policy (drop-in for reference zorch/challenge.py).

``limbs_over`` = deg(C) / deg(T) transcript words per challenge (raises
``ValueError`` when C does not tile T); one squeeze call per challenge set; the
words are regrouped into elements of C by reinterpretation (limb i is
coefficient i).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any

from zorch_amd.fields import FieldArray, as_field, challenge_limbs, stack
from zorch_amd.transcript import reinterpret_challenge


def _limbs_of(dtype: Any) -> int:
    try:
        return challenge_limbs(dtype)
    except ValueError:
        return 1  # a raw word dtype is its own field


@dataclass(frozen=True)
class ChallengePolicy:
    """The field Fiat-Shamir challenges are drawn in (zorch/challenge.py:24-106)."""

    dtype: Any

    def __post_init__(self) -> None:
        object.__setattr__(self, "dtype", as_field(self.dtype))

    @property
    def base_limbs(self) -> int:
        return challenge_limbs(self.dtype)

    def limbs_over(self, transcript_field: Any) -> int:
        words = challenge_limbs(self.dtype)
        per_word = _limbs_of(transcript_field)
        if words % per_word:
            raise ValueError(
                f"{self.dtype} does not tile {transcript_field}: "
                f"{words} coefficients over words of {per_word}")
        return words // per_word

    def _regroup(self, raw: FieldArray, count: int, limbs: int) -> FieldArray:
        return stack([reinterpret_challenge(raw[i * limbs:(i + 1) * limbs], self.dtype)
                      for i in range(count)])

    def sample(self, transcript):
        transcript, challenges = self.sample_many(transcript, 1)
        return transcript, challenges[0]

    def sample_many(self, transcript, count: int):
        limbs = self.limbs_over(transcript.field)
        transcript, raw = transcript.sample(count * limbs)
        return transcript, self._regroup(raw, count, limbs)

    def observe_and_sample_many(self, transcript, values, count: int):
        limbs = self.limbs_over(transcript.field)
        transcript, raw = transcript.observe_and_sample(values, count * limbs)
        return transcript, self._regroup(raw, count, limbs)

    def observe_and_sample(self, transcript, values):
        transcript, challenges = self.observe_and_sample_many(transcript, values, 1)
        return transcript, challenges[0]
