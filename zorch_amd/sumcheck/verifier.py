"""Sumcheck verifier round (drop-in for reference zorch/sumcheck/verifier.py:44-71).

``SumcheckRound(degree, challenges)`` checks ``claim == s(0) + s(1)`` and
reduces the claim to ``s(r)`` with the challenge drawn exactly as the prover
drew it. The step runs in the library's host code (``zs_verify_round``).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from zorch_amd import _lib
from zorch_amd.challenge import ChallengePolicy
from zorch_amd.fields import FieldArray, as_field
from zorch_amd.round import RunningClaim
from zorch_amd.sumcheck.reduce import reduce_evals


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
