"""Jagged LogUp-GKR prover (drop-in for reference zorch/logup_gkr/jagged_prover.py).

A jagged layer materialises only sum(row_counts) of its virtual
2^(niv + nrv) positions; the others hold the neutral fraction (n = 0,
d = 1), whose LogUp summand is its eq weight alone. The layer's sumcheck
binds the row variables LSB first, then the batch variables, and sends each
round's cubic as four ascending coefficients, interpolated through 0, 1,
1/2 and the root of the round's eq factor with the virtual rows' mass added
in closed form (jagged_prover.py:1-33). The whole layer sumcheck is one
device call (``zs_prove_jagged_layer``: one fold-and-sum launch per row
round with the finisher block, one block for the batch rounds);
``JaggedGkrLayerRound`` adds the carry around it as the dense
``GkrLayerRound`` does (jagged_prover.py:453-548, 671-712). A
``JaggedDeviceLayer`` keeps a layer's four planes in HBM;
``jagged_device_pyramid`` uploads the first layer once and runs every
transition on the device (``zs_jagged_transition``), and the layer rounds
prove those planes in place.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from zorch_amd import _lib
from zorch_amd.challenge import ChallengePolicy
from zorch_amd.fields import FieldArray, as_field, stack
from zorch_amd.logup_gkr.circuit import JaggedGkrLayer
from zorch_amd.logup_gkr.prover import fold_carry
from zorch_amd.runtime import DeviceTables, default_context, el_buffer
from zorch_amd.sumcheck.jagged.types import RoundWidthCaps

DEGREE = 3


@dataclass(frozen=True)
class JaggedLayerProof:
    """One jagged layer's sumcheck transcript: the batching challenge and
    the claim the layer entered with, the coefficient-form round polynomials,
    the bound point (most significant first: the challenges reversed) and
    the four openings (jagged_prover.py:89-121)."""

    lam: FieldArray
    claim: FieldArray
    round_polys: FieldArray  # (niv + nrv, 4) ascending coefficients
    point: FieldArray
    numerator_0: FieldArray
    numerator_1: FieldArray
    denominator_0: FieldArray
    denominator_1: FieldArray


def _device_prove(layer: JaggedGkrLayer, lam: FieldArray, claim: FieldArray,
                  eval_point: FieldArray, transcript, chal, ctx=None, resident=None):
    """`resident`: the four live planes already in HBM (DeviceTables of
    shape (4, live) in the challenge field) instead of the layer's host
    arrays."""
    ctx = ctx or default_context()
    niv = layer.num_batch_variables
    nrv = eval_point.shape[0] - niv
    live = layer.live
    if resident is not None:
        if tuple(resident.shape) != (4, live) or resident.dtype != chal:
            raise ValueError(f"resident planes must be (4, {live}) in {chal}")
        words, on_dev, ptr = None, 1, resident.data_ptr()
    else:
        planes = stack([a[:live].astype(chal) for a in (layer.numerator_0, layer.numerator_1,
                                                        layer.denominator_0, layer.denominator_1)])
        words = np.ascontiguousarray(planes.reduced().words)
        on_dev, ptr = 0, words.ctypes.data
    counts = np.ascontiguousarray(np.asarray(layer.row_counts, dtype=np.uint32))
    pt = el_buffer(eval_point, chal)
    lam_b = el_buffer(lam, chal)
    claim_b = el_buffer(claim, chal)
    R = niv + nrv
    polys = np.zeros((R, 4, chal.words), dtype=chal.word_dtype)
    chals = np.zeros((R, chal.words), dtype=chal.word_dtype)
    opens = np.zeros((4, chal.words), dtype=chal.word_dtype)
    value = np.zeros((chal.words,), dtype=chal.word_dtype)
    args = _lib.JaggedArgsC(field=chal.fid, num_segments=layer.num_batches, row_variables=nrv,
                            planes_on_device=on_dev, row_counts=counts.ctypes.data,
                            planes=ptr, eval_point=pt.ctypes.data,
                            lambda_=lam_b.ctypes.data, claim=claim_b.ctypes.data)
    out = _lib.JaggedProofC(round_polys=polys.ctypes.data, challenges=chals.ctypes.data,
                            openings=opens.ctypes.data, value=value.ctypes.data)
    st = transcript.state.copy()
    _lib.check(_lib.lib().zs_prove_jagged_layer(ctx.handle, ctypes.byref(args), ctypes.byref(st),
                                                ctypes.byref(out)))
    return (FieldArray(polys, chal), FieldArray(chals, chal).normalized(), FieldArray(opens, chal),
            type(transcript)(st, transcript.dtype))


class JaggedDeviceLayer:
    """A jagged layer in HBM: the planes n0, n1, d0, d1 (one device buffer,
    row stride = the live rows) in the challenge field, and the row counts."""

    def __init__(self, planes: DeviceTables, row_counts: tuple) -> None:
        self.planes = planes
        self.row_counts = tuple(int(c) for c in row_counts)

    @property
    def num_batches(self) -> int:
        return len(self.row_counts)

    @property
    def num_batch_variables(self) -> int:
        from zorch_amd.runtime import log2_strict
        return log2_strict(self.num_batches)

    @property
    def live(self) -> int:
        return sum(self.row_counts)

    width = live

    @classmethod
    def upload(cls, layer: JaggedGkrLayer, field, ctx=None) -> "JaggedDeviceLayer":
        f = as_field(field)
        ctx = ctx or default_context()
        live = layer.live
        t = DeviceTables.empty((4, live), f, ctx)
        row_bytes = live * f.itemsize
        for q, plane in enumerate((layer.numerator_0, layer.numerator_1, layer.denominator_0,
                                   layer.denominator_1)):
            words = np.ascontiguousarray(plane[:live].astype(f).reduced().words)
            _lib.check(_lib.lib().zs_memcpy_htod(
                ctx.handle, ctypes.c_void_p(t.data_ptr() + q * row_bytes), words.ctypes.data,
                row_bytes))
        return cls(t, layer.row_counts)

    def transition(self, out_row_counts) -> "JaggedDeviceLayer":
        out = tuple(int(c) for c in out_row_counts)
        if len(out) != self.num_batches:
            raise ValueError(f"schedule must cover all {self.num_batches} batches, got "
                             f"{len(out)} entries")
        f = self.planes.field
        nxt = DeviceTables.empty((4, sum(out)), f, self.planes.ctx)
        rows = np.ascontiguousarray(np.asarray(self.row_counts, dtype=np.uint32))
        orows = np.ascontiguousarray(np.asarray(out, dtype=np.uint32))
        _lib.check(_lib.lib().zs_jagged_transition(
            self.planes.ctx.handle, f.fid, self.num_batches, rows.ctypes.data, orows.ctypes.data,
            ctypes.c_void_p(self.planes.data_ptr()), self.live, ctypes.c_void_p(nxt.data_ptr()),
            sum(out)))
        return JaggedDeviceLayer(nxt, out)

    def to_host(self) -> JaggedGkrLayer:
        p = self.planes.to_host()
        return JaggedGkrLayer(p[0], p[1], p[2], p[3], self.row_counts)

    def free(self) -> None:
        self.planes.free()


def jagged_device_pyramid(first: JaggedGkrLayer, schedules, field, ctx=None):
    """[first, ..., floor] in HBM: the first layer uploaded once, one
    zs_jagged_transition per schedule entry (out row counts)."""
    layers = [JaggedDeviceLayer.upload(first, field, ctx)]
    for sched in schedules:
        layers.append(layers[-1].transition(sched))
    return layers


def _check_layer(layer: JaggedGkrLayer, eval_point: FieldArray, caps) -> int:
    if caps is None:
        raise ValueError("a jagged layer proves under caps; pass RoundWidthCaps")
    niv = layer.num_batch_variables
    nrv = eval_point.shape[0] - niv
    if nrv < 1:
        raise ValueError(f"eval_point must carry at least one row variable: got "
                         f"{eval_point.shape[0]} coordinates for {niv} batch variables")
    if max(layer.row_counts) > (1 << nrv):
        raise ValueError(f"row count {max(layer.row_counts)} exceeds the virtual row space "
                         f"2^{nrv}; the row-eq lookup would run out of bounds")
    if min(layer.row_counts) < 1:
        raise ValueError("every segment needs at least one row")
    caps.check(layer.width, nrv, niv)
    return nrv


def prove_jagged_layer(layer: JaggedGkrLayer, lam: FieldArray, claim: FieldArray,
                       eval_point: FieldArray, transcript, *, challenges: ChallengePolicy,
                       caps: RoundWidthCaps | None = None):
    """One jagged layer's sumcheck from an explicit lam and claim (the
    standalone seam, jagged_prover.py:124-183). `eval_point` is most
    significant first over (batch || row) variables. Returns (bound point,
    transcript, proof)."""
    _check_layer(layer, eval_point, caps)
    chal = as_field(challenges.dtype)
    lam = lam.astype(chal).normalized()
    claim = claim.astype(chal).normalized()
    eval_point = eval_point.astype(chal)
    resident = None
    if isinstance(layer, JaggedDeviceLayer):
        if layer.planes.field is not chal:
            raise ValueError(f"device layer in {layer.planes.field}, challenges in {chal}")
        resident = layer.planes
    polys, chals, opens, transcript = _device_prove(layer, lam, claim, eval_point, transcript,
                                                    chal, resident=resident)
    point = chals[::-1]
    proof = JaggedLayerProof(lam, claim, polys, point, opens[0], opens[1], opens[2], opens[3])
    return point, transcript, proof


class JaggedGkrLayerRound:
    """Prove one jagged layer with the dense chain's carry: sample lam, form
    lam num_eval + den_eval, prove, observe the openings, sample the child
    selector and fold the carry (jagged_prover.py:453-548, 671-712)."""

    def __init__(self, layer: JaggedGkrLayer, challenges: ChallengePolicy, *,
                 caps: RoundWidthCaps | None = None, layer_bufs=None) -> None:
        self.layer = layer
        self.challenges = challenges
        self.caps = caps

    def __call__(self, claim, transcript):
        num_eval, den_eval, eval_point = claim
        if self.caps is None:
            raise ValueError("a jagged layer proves under caps; pass RoundWidthCaps")
        transcript, lam = self.challenges.sample(transcript)
        lam = lam.normalized()
        batched = lam * num_eval + den_eval
        point, transcript, proof = prove_jagged_layer(
            self.layer, lam, batched, eval_point, transcript, challenges=self.challenges,
            caps=self.caps)
        n0, n1 = proof.numerator_0, proof.numerator_1
        d0, d1 = proof.denominator_0, proof.denominator_1
        transcript, r = self.challenges.observe_and_sample(transcript, stack([n0, n1, d0, d1]))
        return fold_carry(n0, n1, d0, d1, point, r.normalized()), transcript, proof


__all__ = ["JaggedLayerProof", "prove_jagged_layer", "JaggedGkrLayerRound", "RoundWidthCaps"]
