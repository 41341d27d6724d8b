"""Sumcheck verifier rounds (drop-in for reference zorch/sumcheck/verifier.py:45-134).

``SumcheckRound(degree, challenges)`` checks ``claim == s(0) + s(1)`` on the
natural-domain values and reduces the claim to ``s(r)`` with the challenge
drawn exactly as the prover drew it (the library's host code,
``zs_verify_round``). ``CoeffsSumcheckRound`` is the same check for ascending
coefficients (s(0) = c_0, s(1) = sum c), and ``CompressedCoeffsSumcheckRound``
the compressed [c_0, c_2] wire, whose c_1 is rebuilt from the claim (its ``ok``
is constant true; binding rests on the terminal check). Those two reduce on
host field arrays (``sumcheck.reduce``) over the library's host transcript.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from zorch_amd import _lib
from zorch_amd.challenge import ChallengePolicy
from zorch_amd.fields import FieldArray, as_field
from zorch_amd.round import RunningClaim
from zorch_amd.sumcheck.reduce import (reduce_coeffs, reduce_compressed, reduce_evals,
                                       require_width)


@dataclass(frozen=True)
class SumcheckRound:
    degree: int
    challenges: ChallengePolicy

    def __post_init__(self) -> None:
        if self.degree < 1:
            raise ValueError("degree must be >= 1")

    def check_reduce(self, claim: FieldArray, msg: FieldArray, r: FieldArray):
        """The round identity and reduction for an externally drawn challenge."""
        return reduce_evals(claim, msg, r, self.degree)

    def __call__(self, claim: RunningClaim, transcript, msg: FieldArray):
        f = as_field(self.challenges.dtype)
        limbs = self.challenges.limbs_over(transcript.field)
        if msg.field is not f:
            raise ValueError("round messages must be in the challenge field "
                             f"{f.name}, got {msg.field.name}")
        value = claim.value.astype(f)
        if msg.ndim != 1:
            raise ValueError("a round message is one row of the proof")
        st = transcript.state.copy()
        if int(st.item_bytes) * limbs != f.itemsize:
            raise ValueError("challenge field does not tile the transcript field")
        mw = np.ascontiguousarray(msg.words)
        cw = np.ascontiguousarray(value.words)
        chal = np.zeros((f.words,), dtype=f.word_dtype)
        red = np.zeros((f.words,), dtype=f.word_dtype)
        ok = ctypes.c_int(0)
        _lib.check(_lib.lib().zs_verify_round(
            f.fid, ctypes.byref(st), cw.ctypes.data, mw.ctypes.data, msg.shape[0], self.degree,
            chal.ctypes.data, red.ctypes.data, ctypes.byref(ok)))
        new_t = type(transcript)(st, transcript.dtype)
        return claim.bind(FieldArray(red, f), FieldArray(chal, f)), new_t, bool(ok.value)


def _message_in(challenges: ChallengePolicy, msg: FieldArray, value: FieldArray):
    f = as_field(challenges.dtype)
    if msg.field is not f:
        raise ValueError("round messages must be in the challenge field "
                         f"{f.name}, got {msg.field.name}")
    if msg.ndim != 1:
        raise ValueError("a round message is one row of the proof")
    return value.astype(f)


@dataclass(frozen=True)
class CoeffsSumcheckRound:
    """Verifier of a coefficient-form round (verifier.py:76-100)."""

    degree: int
    challenges: ChallengePolicy

    def __post_init__(self) -> None:
        if self.degree < 1:
            raise ValueError("degree must be >= 1")

    def __call__(self, claim: RunningClaim, transcript, msg: FieldArray):
        # Structural rejection precedes any read of the claim.
        require_width(msg, self.degree + 1, "coefficients")
        value = _message_in(self.challenges, msg, claim.value)
        transcript, r = self.challenges.observe_and_sample(transcript, msg)
        reduced, ok = reduce_coeffs(value, msg, r, self.degree)
        return claim.bind(reduced, r), transcript, ok


@dataclass(frozen=True)
class CompressedCoeffsSumcheckRound:
    """Verifier of the compressed [c_0, c_2] wire (verifier.py:104-134)."""

    challenges: ChallengePolicy

    def check_reduce(self, claim: FieldArray, msg: FieldArray, r: FieldArray):
        """Rebuild c_1 from the claim and reduce, for an external challenge."""
        return reduce_compressed(claim, msg, r)

    def __call__(self, claim: RunningClaim, transcript, msg: FieldArray):
        value = _message_in(self.challenges, msg, claim.value)
        transcript, r = self.challenges.observe_and_sample(transcript, msg)
        reduced, ok = self.check_reduce(value, msg, r)
        return claim.bind(reduced, r), transcript, ok


@dataclass(frozen=True)
class UnivariateSkipRound:
    """Verifier of the skip's round 0 (reference verifier.py, UnivariateSkipRound;
    univariate_skip.py:193-245): the base-field coefficient message is
    observed, the challenge drawn in the challenge field, the subgroup-sum
    identity checked and the claim reduced to s_0(r_0)."""

    skip_rounds: int
    degree: int
    challenges: ChallengePolicy

    def __call__(self, claim: RunningClaim, transcript, msg: FieldArray):
        require_width(msg, self.degree * ((1 << self.skip_rounds) - 1) + 1, "coefficients")
        if msg.ndim != 1:
            raise ValueError("a round message is one row of the proof")
        from zorch_amd.sumcheck.reduce import reduce_subgroup
        transcript, r = self.challenges.observe_and_sample(transcript, msg)
        reduced, ok = reduce_subgroup(claim.value, msg, r, self.skip_rounds, self.degree)
        return claim.bind(reduced, r), transcript, ok
