"""Stage roles (drop-in for reference zorch/stage.py:48-95)."""

from __future__ import annotations

from abc import abstractmethod
from dataclasses import dataclass
from typing import Any, Generic, Protocol, TypeVar

R = TypeVar("R")
P = TypeVar("P")
T = TypeVar("T")


@dataclass(frozen=True)
class TrivialClaim:
    """The claim that holds by construction."""


@dataclass(frozen=True)
class ProveResult(Generic[R, P, T]):
    """A reduced claim, its conditional reduction proof, and the transcript."""

    reduced_claim: Any
    reduction_proof: Any
    transcript: Any


@dataclass(frozen=True)
class VerifyResult(Generic[R, T]):
    """The verifier-derived reduced claim, advanced transcript, and verdict."""

    reduced_claim: Any
    transcript: Any
    ok: Any


class ProverStage(Protocol):
    @abstractmethod
    def prove(self, claim: Any, witness: Any, transcript: Any) -> ProveResult: ...


class VerifierStage(Protocol):
    @abstractmethod
    def verify(self, claim: Any, reduction_proof: Any, transcript: Any) -> VerifyResult: ...
