"""Whole-stage device proving through ``zs_prove`` / ``zs_prove_batch``.

One C call runs every round of a stage on the GPU (fused fold+evaluate pass,
device Fiat-Shamir hop, claim reduction) and returns the stacked messages,
the raw challenge words (the point), the reduced value, the advanced
transcript and the final folded values.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Any, Sequence

import numpy as np

from zorch_amd import _lib
from zorch_amd.fields import Field, FieldArray, as_field
from zorch_amd.runtime import (Context, default_context, el_buffer, log2_strict, message_width,
                               tables_pointer, wire_struct)


@dataclass
class DeviceProof:
    msgs: FieldArray        # (rounds, npts)
    point: FieldArray       # (rounds,) raw squeezed words
    value: FieldArray       # ()
    transcript: Any
    final: FieldArray       # (k,) folded values (+ eq mass in eq mode)


@dataclass
class StageJob:
    """One instance's arguments, kept alive for the duration of a C call."""

    args: _lib.ProveArgsC
    proof: _lib.ProofC
    keep: list
    chal: Field
    npts: int
    k: int
    rounds: int
    mode: int
    msgs: np.ndarray
    point: np.ndarray
    value: np.ndarray
    final: np.ndarray


def make_job(tables, claim: FieldArray, chal_field: Any, mode: int, degree: int,
             eq_point: FieldArray | None = None, *, wire_kind: int = _lib.ZS_WIRE_VALUES,
             domain: Any = None, refold_rounds: int = 0,
             lam: FieldArray | None = None) -> StageJob:
    """Arguments and output buffers of one stage; `wire_kind` / `domain` pick
    the message wire form (natural values by default); `refold_rounds` first
    rounds refold the untouched input (the sqrt-space engine); `lam` is a
    LogUp stage's batching challenge."""
    chal = as_field(chal_field)
    ptr, on_dev, keep, shape = tables_pointer(tables)
    k, width = shape
    n = log2_strict(width)
    table_field = tables.dtype if not isinstance(tables, FieldArray) else tables.field
    claim_buf = el_buffer(claim, chal)
    eq_buf = el_buffer(eq_point, chal) if eq_point is not None else None
    if mode == _lib.ZS_MODE_EQ and (eq_point is None or eq_point.shape[0] != n):
        raise ValueError(f"equality point needs {n} coordinates")
    wire, wire_keep = wire_struct(wire_kind, domain, chal)
    wire_ptr = ctypes.pointer(wire) if wire is not None else None
    lam_buf = el_buffer(lam, chal) if lam is not None else None
    args = _lib.ProveArgsC(
        table_field=as_field(table_field).fid, chal_field=chal.fid, mode=mode, degree=degree,
        num_tables=k, rounds=n, tables=ptr, tables_on_device=1 if on_dev else 0,
        refold_rounds=refold_rounds,
        eq_point=eq_buf.ctypes.data if eq_buf is not None else None, claim=claim_buf.ctypes.data,
        wire=wire_ptr, lambda_=lam_buf.ctypes.data if lam_buf is not None else None)
    npts = message_width(args)
    msgs = np.zeros((n, npts, chal.words), dtype=chal.word_dtype)
    point = np.zeros((n, chal.words), dtype=chal.word_dtype)
    value = np.zeros((chal.words,), dtype=chal.word_dtype)
    final = np.zeros((k + 1, chal.words), dtype=chal.word_dtype)
    proof = _lib.ProofC(msgs=msgs.ctypes.data, point=point.ctypes.data, value=value.ctypes.data,
                        final_tables=final.ctypes.data)
    return StageJob(args, proof, [keep, claim_buf, eq_buf, wire, wire_ptr, wire_keep, lam_buf], chal,
                    npts, k, n, mode, msgs, point, value, final)


def _result(job: StageJob, st: _lib.TranscriptC, transcript) -> DeviceProof:
    c = job.chal
    nfin = job.k + (1 if job.mode == _lib.ZS_MODE_EQ else 0)
    return DeviceProof(FieldArray(job.msgs, c), FieldArray(job.point, c), FieldArray(job.value, c),
                       type(transcript)(st, transcript.dtype), FieldArray(job.final[:nfin], c))


def prove_stage(tables, claim: FieldArray, chal_field: Any, mode: int, degree: int, transcript,
                eq_point: FieldArray | None = None, ctx: Context | None = None, *,
                wire_kind: int = _lib.ZS_WIRE_VALUES, domain: Any = None,
                refold_rounds: int = 0, lam: FieldArray | None = None) -> DeviceProof:
    ctx = ctx or default_context()
    job = make_job(tables, claim, chal_field, mode, degree, eq_point, wire_kind=wire_kind,
                   domain=domain, refold_rounds=refold_rounds, lam=lam)
    st = transcript.state.copy()
    _lib.check(_lib.lib().zs_prove(ctx.handle, ctypes.byref(job.args), ctypes.byref(st),
                                   ctypes.byref(job.proof)))
    return _result(job, st, transcript)


def prove_stage_batch(jobs: Sequence[StageJob], transcripts: Sequence[Any],
                      ctx: Context | None = None,
                      tune_key: str | None = "batch") -> list[DeviceProof]:
    """Every job in one ``zs_prove_batch``; the launch shapes cached by the
    autotuner under `tune_key` apply for the call (results never depend on
    them)."""
    from zorch_amd import autotune

    ctx = ctx or default_context()
    n = len(jobs)
    args = (_lib.ProveArgsC * n)(*[j.args for j in jobs])
    sts = (_lib.TranscriptC * n)(*[t.state.copy() for t in transcripts])
    proofs = (_lib.ProofC * n)(*[j.proof for j in jobs])
    if tune_key is None:
        _lib.check(_lib.lib().zs_prove_batch(ctx.handle, n, args, sts, proofs))
    else:
        with autotune.applied(ctx, tune_key):
            _lib.check(_lib.lib().zs_prove_batch(ctx.handle, n, args, sts, proofs))
    return [_result(j, sts[i].copy(), transcripts[i]) for i, j in enumerate(jobs)]
