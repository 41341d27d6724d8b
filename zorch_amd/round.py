"""Round protocols and the running claim (drop-in for reference zorch/round.py)."""

from __future__ import annotations

from collections.abc import Iterable, Sequence
from dataclasses import dataclass
from typing import Any, Protocol

from zorch_amd.fields import FieldArray


@dataclass(frozen=True)
class RunningClaim:
    """A partially built evaluation claim (zorch/round.py:28-49): the value after
    the rounds bound so far, the point being built and the write cursor."""

    value: FieldArray
    point: FieldArray
    index: int

    def bind(self, value: FieldArray, challenge: FieldArray) -> "RunningClaim":
        """Advance to the reduced claim, recording this round's challenge."""
        return RunningClaim(value, self.point.at_set(self.index, challenge), self.index + 1)


class ProverRound(Protocol):
    """A prover step under fold_rounds: (carry, transcript) -> (carry,
    transcript, message) (reference zorch/round.py:52-60)."""

    def __call__(self, carry: Any, transcript: Any, /) -> tuple[Any, Any, Any]: ...


class VerifierRound(Protocol):
    """A verifier step: (carry, transcript, message) -> (carry, transcript,
    ok); whatever the step derives travels in the carry (zorch/round.py:63-78)."""

    def __call__(self, carry: Any, transcript: Any, message: Any, /) -> tuple[Any, Any, Any]: ...


def prove_rounds(rounds: Iterable[ProverRound], carry: Any, transcript: Any):
    """Run prover rounds, collecting each step's message (zorch/round.py:81-91)."""
    msgs = []
    for rnd in rounds:
        carry, transcript, msg = rnd(carry, transcript)
        msgs.append(msg)
    return carry, transcript, msgs


def verify_rounds(rounds: Iterable[VerifierRound], carry: Any, msgs: Sequence[Any],
                  transcript: Any):
    """Replay verifier rounds over a message list, ANDing verdicts
    (zorch/round.py:94-115)."""
    materialized = list(rounds)
    if len(msgs) != len(materialized):
        raise ValueError(
            f"need one message per round: {len(materialized)} rounds, got {len(msgs)} messages")
    ok: Any = True
    for rnd, msg in zip(materialized, msgs):
        carry, transcript, ok_round = rnd(carry, transcript, msg)
        ok = ok and bool(ok_round)
    return carry, transcript, ok
