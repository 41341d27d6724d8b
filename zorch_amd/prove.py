"""Scheme-agnostic multi-round proving driver (reference zorch/prove.py:26-37)."""

from __future__ import annotations

from typing import Any


def fold_rounds(rnd, state: Any, transcript: Any, rounds: int):
    """Run `rnd` exactly `rounds` times; return (state, transcript, list[msg])."""
    msgs: list[Any] = []
    for _ in range(rounds):
        state, transcript, msg = rnd(state, transcript)
        msgs.append(msg)
    return state, transcript, msgs
