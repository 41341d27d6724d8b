"""Sumcheck replay driver (reference zorch/verify.py:26-48)."""

from __future__ import annotations

from zorch_amd.fields import FieldArray, zeros
from zorch_amd.round import RunningClaim


def verify(verifier, claim: FieldArray, proof: FieldArray, transcript):
    """Replay `proof` against `claim` -> (point, final_claim, transcript, ok).

    `ok` ANDs every round's check; structural problems raise ValueError.
    """
    if proof.ndim != 2 or proof.shape[0] == 0:
        raise ValueError("proof must be a non-empty 2-D array (one row per round)")
    state = RunningClaim(claim, zeros((proof.shape[0],), claim.dtype), 0)
    ok = True
    for i in range(proof.shape[0]):
        state, transcript, ok_i = verifier(state, transcript, proof[i])
        ok = ok and bool(ok_i)
    return state.point, state.value, transcript, ok
