"""Dense LogUp-GKR prover (drop-in for reference zorch/logup_gkr/prover.py).

``logup_combine`` is the summand eq (lam (n0 d1 + n1 d0) + d0 d1) over the
factors [eq, n0, d1, n1, d0] (prover.py:56-68), ``LogupSummand`` its degree
and arity (:71-117). ``LogupSumcheckRound`` is one per-variable round on the
GPU (:232-300): its carry is the stacked five-table state (host field arrays
or device tables before the first round, then a device fold state) and each
call folds, evaluates and hashes one round (``zs_state_round`` of a
``ZS_MODE_LOGUP`` stage). ``GkrLayerRound`` proves one layer (:344-378): lam,
the eq table of the carry's point, the layer's whole sumcheck in one device
call (``zs_prove``), the final openings observed for the child selector and
``fold_carry`` (:159-172). ``bind_output`` is the chain's head (:307-341).

Layers are proved where they live: a ``DeviceLayer`` is one device buffer of
five rows laid out as the stage's tables [eq, n0, d1, n1, d0];
``device_pyramid`` uploads the first layer once and builds the others with
``zs_logup_transition``, and each layer's row 0 receives the eq table of its
point from ``zs_logup_eq`` before the stage runs on the buffer in place.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Any, Sequence

import numpy as np

from zorch_amd import _lib
from zorch_amd.challenge import ChallengePolicy
from zorch_amd.fields import FieldArray, as_field, concatenate, fsum, ones, stack, zeros
from zorch_amd.logup_gkr.circuit import GkrLayer, LogUpGkrOutput, build_pyramid, extract_outputs
from zorch_amd.round import prove_rounds
from zorch_amd.runtime import (DeviceFoldState, DeviceTables, default_context, el_buffer,
                               log2_strict)
from zorch_amd.sumcheck._device import prove_stage

DEGREE = 3
NUM_FACTORS = 5


def logup_combine(lam, eq, n0, d1, n1, d0):
    """The LogUp summand eq (lam (n0 d1 + n1 d0) + d0 d1) (host arrays)."""
    return eq * (lam * (n0 * d1 + n1 * d0) + d0 * d1)


class LogupSummand:
    """The LogUp combine, its scalar lam and its degree (prover.py:71-117)."""

    DEGREE: int = DEGREE
    NUM_FACTORS: int = NUM_FACTORS

    def __init__(self, lam: FieldArray) -> None:
        self.lam = lam

    @property
    def degree(self) -> int:
        return self.DEGREE

    def combine_scalars(self) -> tuple:
        return (self.lam,)

    def combine(self, scalars: Sequence[Any], *factors: FieldArray) -> FieldArray:
        if len(factors) != self.NUM_FACTORS:
            raise ValueError(f"LogUp combine needs {self.NUM_FACTORS} factors "
                             f"[eq, n0, d1, n1, d0], got {len(factors)}")
        (lam,) = scalars
        return logup_combine(lam, *factors)

    def _combine(self, *factors: FieldArray) -> FieldArray:
        return self.combine(self.combine_scalars(), *factors)


@dataclass(frozen=True)
class RoundMsg:
    """One round's message and its challenge (reference zorch/sumcheck/prover.py)."""

    round_poly: FieldArray
    challenge: FieldArray


def _logup_state(tables, lam: FieldArray, chal: Any) -> DeviceFoldState:
    if isinstance(tables, FieldArray):
        if tables.ndim != 2 or tables.shape[0] != NUM_FACTORS:
            raise ValueError(f"LogUp combine needs {NUM_FACTORS} factors [eq, n0, d1, n1, d0], "
                             f"got a state of shape {tables.shape}")
        tables = tables.astype(chal)
    elif not isinstance(tables, DeviceTables):
        raise TypeError("state must be stacked tables (FieldArray or DeviceTables)")
    return DeviceFoldState(tables, zeros((), chal), chal, _lib.ZS_MODE_LOGUP, DEGREE, lam=lam)


@dataclass(frozen=True)
class LogupSumcheckRound:
    """Per-variable LogUp sumcheck round (prover.py:232-300): emits a RoundMsg."""

    lam: FieldArray
    challenges: ChallengePolicy

    @property
    def _summand(self) -> LogupSummand:
        return LogupSummand(self.lam)

    @property
    def degree(self) -> int:
        return DEGREE

    def combine_scalars(self) -> tuple:
        return (self.lam,)

    def combine(self, scalars: Sequence[Any], *factors: FieldArray) -> FieldArray:
        return self._summand.combine(scalars, *factors)

    def _combine(self, *factors: FieldArray) -> FieldArray:
        return self.combine(self.combine_scalars(), *factors)

    def _round_poly(self, folded) -> FieldArray:
        """The round polynomial at t = 0..3 for a stacked state, without
        advancing any caller state (a scratch state and transcript)."""
        from zorch_amd.sha256_field_transcript import Sha256FieldTranscript

        chal = as_field(self.challenges.dtype)
        st = _logup_state(folded, self.lam, chal)
        msg, _, _, _ = st.round(Sha256FieldTranscript.new(b"", chal.base))
        return msg

    def __call__(self, folded, transcript):
        state = folded
        if not isinstance(state, DeviceFoldState):
            state = _logup_state(folded, self.lam, as_field(self.challenges.dtype))
        msg, chal, _, transcript = state.round(transcript)
        return state, transcript, RoundMsg(msg, chal.normalized())


@dataclass(frozen=True)
class LayerProof:
    """One layer's sumcheck transcript: round polynomials, the bound point
    and the final openings (prover.py:192-212)."""

    round_polys: FieldArray  # (num_variables, 4)
    point: FieldArray
    numerator_0: FieldArray
    numerator_1: FieldArray
    denominator_0: FieldArray
    denominator_1: FieldArray


def fold_carry(n0, n1, d0, d1, point: FieldArray, r: FieldArray):
    """Bind the openings at the child selector r, appended as the low
    coordinate of the MSB-first point (prover.py:159-172)."""
    num_eval = n0 + (n1 - n0) * r
    den_eval = d0 + (d1 - d0) * r
    return num_eval, den_eval, concatenate([point, r.reshape(1)])


def eval_mle(values: FieldArray, point: FieldArray) -> FieldArray:
    """MLE evaluation binding the high variable first (reference
    zorch/poly/multilinear.py eval_mle)."""
    cur = values
    for i in range(point.shape[0]):
        h = cur.shape[0] // 2
        cur = cur[:h] + point[i] * (cur[h:] - cur[:h])
    return cur[0]


def expand_eq_to_hypercube(point: FieldArray, scale: FieldArray) -> FieldArray:
    """scale * eq(point, x) over {0,1}^n, point[0] on the index MSB
    (reference zorch/poly/eq.py expand_eq_to_hypercube)."""
    f = point.field
    cur = scale.reshape(1).astype(f)
    for i in range(point.shape[0] - 1, -1, -1):
        hi = cur * point[i]
        cur = concatenate([cur - hi, hi])
    return cur


class DeviceLayer:
    """A layer in HBM: rows [eq, n0, d1, n1, d0] of `width` challenge-field
    elements (row 0 is scratch for the eq table of the point it is proved
    at)."""

    def __init__(self, tables: DeviceTables, num_batch_variables: int) -> None:
        self.tables = tables
        self.num_batch_variables = num_batch_variables

    @property
    def width(self) -> int:
        return self.tables.shape[1]

    @property
    def num_variables(self) -> int:
        return log2_strict(self.width)

    @property
    def num_row_variables(self) -> int:
        return self.num_variables - self.num_batch_variables

    @classmethod
    def upload(cls, layer: GkrLayer, field, ctx=None) -> "DeviceLayer":
        f = as_field(field)
        ctx = ctx or default_context()
        w = layer.numerator_0.shape[0]
        t = DeviceTables.empty((NUM_FACTORS, w), f, ctx)
        row_bytes = w * f.itemsize
        for q, plane in enumerate((layer.numerator_0, layer.denominator_1, layer.numerator_1,
                                   layer.denominator_0), start=1):
            words = np.ascontiguousarray(plane.astype(f).reduced().words)
            _lib.check(_lib.lib().zs_memcpy_htod(
                ctx.handle, ctypes.c_void_p(t.data_ptr() + q * row_bytes), words.ctypes.data,
                row_bytes))
        return cls(t, layer.num_batch_variables)

    def transition(self) -> "DeviceLayer":
        if self.num_row_variables < 1:
            raise ValueError("no row variable left to fold")
        f = self.tables.field
        nxt = DeviceTables.empty((NUM_FACTORS, self.width // 2), f, self.tables.ctx)
        _lib.check(_lib.lib().zs_logup_transition(self.tables.ctx.handle, f.fid,
                                                  ctypes.c_void_p(self.tables.data_ptr()),
                                                  self.width, ctypes.c_void_p(nxt.data_ptr())))
        return DeviceLayer(nxt, self.num_batch_variables)

    def set_eq(self, point: FieldArray) -> None:
        f = self.tables.field
        if point.shape[0] != self.num_variables:
            raise ValueError(f"eq point mismatch: {point.shape[0]} coords for a layer of "
                             f"{self.num_variables} variables")
        buf = el_buffer(point, f)
        _lib.check(_lib.lib().zs_logup_eq(self.tables.ctx.handle, f.fid, buf.ctypes.data,
                                          point.shape[0], ctypes.c_void_p(self.tables.data_ptr())))

    def to_host(self) -> GkrLayer:
        rows = self.tables.to_host()
        return GkrLayer(rows[1], rows[3], rows[4], rows[2], self.num_batch_variables)

    def free(self) -> None:
        self.tables.free()


def device_pyramid(first: GkrLayer, field, ctx=None) -> list[DeviceLayer]:
    """The circuit on the device: the first layer uploaded once, every
    transition a launch (zs_logup_transition)."""
    layers = [DeviceLayer.upload(first, field, ctx)]
    while layers[-1].num_row_variables > 0:
        layers.append(layers[-1].transition())
    return layers


def bind_output(output: LogUpGkrOutput, transcript, challenges: ChallengePolicy):
    """Observe the output MLEs and draw the initial claim point in one hop
    (prover.py:307-341)."""
    num_vars = log2_strict(output.numerator.shape[0])
    f = as_field(challenges.dtype)
    transcript, point = challenges.observe_and_sample_many(
        transcript, concatenate([output.numerator, output.denominator]), num_vars)
    point = point.normalized()
    return (eval_mle(output.numerator.astype(f), point),
            eval_mle(output.denominator.astype(f), point), point), transcript


class GkrLayerRound:
    """Prove one GKR layer (prover.py:344-378); the chain of these, floor
    outward, is the GKR prover."""

    def __init__(self, layer, challenges: ChallengePolicy) -> None:
        self.layer = layer  # a GkrLayer (uploaded when proved) or a DeviceLayer
        self.challenges = challenges

    def __call__(self, claim, transcript):
        num_eval, den_eval, eval_point = claim
        f = as_field(self.challenges.dtype)
        transcript, lam = self.challenges.sample(transcript)
        lam = lam.normalized()
        lay = self.layer
        owned = not isinstance(lay, DeviceLayer)
        if owned:
            lay = DeviceLayer.upload(lay, f)
        elif lay.tables.field is not f:
            raise ValueError(f"device layer in {lay.tables.field}, challenges in {f}")
        try:
            lay.set_eq(eval_point.astype(f))
            # The stage's claim input feeds only the reduced-claim output, which
            # this round does not read (the verifier recomputes it).
            proof = prove_stage(lay.tables, zeros((), f), f, _lib.ZS_MODE_LOGUP, DEGREE,
                                transcript, lam=lam)
        finally:
            if owned:
                lay.free()
        transcript = proof.transcript
        point = proof.point.normalized()
        _, n0, d1, n1, d0 = (proof.final[i] for i in range(NUM_FACTORS))
        transcript, r = self.challenges.observe_and_sample(transcript, stack([n0, n1, d0, d1]))
        carry = fold_carry(n0, n1, d0, d1, point, r.normalized())
        return carry, transcript, LayerProof(proof.msgs, point, n0, n1, d0, d1)


def prove_gkr(first: GkrLayer, transcript, challenges: ChallengePolicy):
    """The GKR prover chain over `first`'s pyramid (reference test kit
    zorch/logup_gkr/testing/__init__.py prove_gkr_with_transcript), the
    pyramid built and proved on the device. Returns (layers, output,
    layer_proofs, final carry, transcript); the layers are DeviceLayers
    (``to_host()`` reads one back)."""
    f = as_field(challenges.dtype)
    layers = device_pyramid(first, f)
    output = extract_outputs(layers[-1].to_host())
    carry, transcript = bind_output(output, transcript, challenges)
    final, transcript, proofs = prove_rounds(
        [GkrLayerRound(layer, challenges) for layer in reversed(layers[:-1])], carry, transcript)
    return layers, output, proofs, final, transcript


__all__ = ["logup_combine", "LogupSummand", "RoundMsg", "LogupSumcheckRound", "LayerProof",
           "fold_carry", "eval_mle", "expand_eq_to_hypercube", "bind_output", "GkrLayerRound",
           "prove_gkr", "DEGREE", "DeviceLayer", "device_pyramid"]
