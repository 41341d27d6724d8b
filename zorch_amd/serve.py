"""Continuous-batching proof server (``zs_server``).

Requests are sumcheck stages (the arguments of ``prove_stage``). The server's
scheduler thread admits them against an HBM budget and advances every
admitted request by one round per step, one batched launch per request kind,
so requests that arrive while others are mid-proof join at the next step.
Results equal ``prove_stage`` on the same arguments and transcript.

The reference has no serving loop (it proves one stage per call); this is the
serving form of the same prover.
"""

from __future__ import annotations

import ctypes
from typing import Any

from zorch_amd import _lib
from zorch_amd.fields import FieldArray
from zorch_amd.runtime import Context
from zorch_amd.sumcheck._device import DeviceProof, _result, make_job


class ProofServer:
    """A proof server on a context of its own (``ctx=None``, the default) or on
    `ctx`, which it takes over: the server's scheduler thread owns the
    context's stream and launch shapes until ``close``. Inputs resident on
    another context of the same device may be submitted."""

    def __init__(self, ctx: Context | None = None, hbm_budget_bytes: int = 0,
                 tune_key: str | None = "serve"):
        from zorch_amd import autotune

        self._own_ctx = ctx is None
        self.ctx = ctx if ctx is not None else Context(0, hbm_budget_bytes)
        # The launch shapes the autotuner cached for served proving, if any
        # (the server owns the context's shapes; results never depend on them).
        if tune_key is not None:
            autotune.apply(self.ctx, tune_key)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().zs_server_create(self.ctx.handle, hbm_budget_bytes,
                                               ctypes.byref(h)))
        self._h = h
        self._jobs: dict[int, tuple] = {}

    @property
    def handle(self) -> ctypes.c_void_p:
        if self._h is None:
            raise RuntimeError("server is closed")
        return self._h

    def submit(self, tables, claim: FieldArray, chal_field: Any, mode: int, degree: int,
               transcript, eq_point: FieldArray | None = None) -> int:
        """Queue one stage; `tables` must stay alive until it completes (kept here)."""
        return self.submit_job(make_job(tables, claim, chal_field, mode, degree, eq_point),
                               transcript)

    def submit_job(self, job, transcript) -> int:
        """Queue one stage prepared by ``sumcheck._device.make_job`` (its output
        buffers receive the proof; one job per request)."""
        st = transcript.state.copy()
        rid = ctypes.c_uint64()
        _lib.check(_lib.lib().zs_server_submit(self.handle, ctypes.byref(job.args),
                                               ctypes.byref(st), ctypes.byref(rid)))
        self._jobs[rid.value] = (job, transcript, st)
        return rid.value

    def _take(self, rid: int, wait: float | None) -> tuple[DeviceProof, float] | None:
        if rid not in self._jobs:
            raise ValueError(f"unknown request id {rid}")
        job, transcript, st = self._jobs[rid]
        done = ctypes.c_int(0)
        lat = ctypes.c_double(0.0)
        lib = _lib.lib()
        try:
            if wait is None:
                rc = lib.zs_server_poll(self.handle, rid, ctypes.byref(job.proof), ctypes.byref(st),
                                        ctypes.byref(lat), ctypes.byref(done))
            else:
                rc = lib.zs_server_wait(self.handle, rid, ctypes.byref(job.proof),
                                        ctypes.byref(st), ctypes.byref(lat), wait,
                                        ctypes.byref(done))
            _lib.check(rc)
        finally:
            if done.value:
                del self._jobs[rid]
        if not done.value:
            return None
        return _result(job, st, transcript), lat.value

    def poll(self, rid: int) -> tuple[DeviceProof, float] | None:
        """(proof, latency in s) once finished, else None."""
        return self._take(rid, None)

    def wait(self, rid: int, timeout: float = 120.0) -> tuple[DeviceProof, float]:
        res = self._take(rid, timeout)
        if res is None:
            raise TimeoutError(f"request {rid} still running after {timeout} s")
        return res

    def completed(self, cap: int = 4096) -> list[int]:
        """Ids finished since the last call."""
        buf = (ctypes.c_uint64 * cap)()
        n = ctypes.c_size_t(0)
        _lib.check(_lib.lib().zs_server_completed(self.handle, buf, cap, ctypes.byref(n)))
        return list(buf[: n.value])

    def stats(self) -> dict:
        s = _lib.ServerStatsC()
        _lib.check(_lib.lib().zs_server_stats_read(self.handle, ctypes.byref(s)))
        return {name: getattr(s, name) for name, _ in s._fields_}

    def close(self) -> None:
        if self._h is not None:
            _lib.lib().zs_server_destroy(self._h)
            self._h = None
            self._jobs.clear()
        if self._own_ctx and self.ctx is not None:
            self.ctx.close()
            self.ctx = None

    def __enter__(self) -> "ProofServer":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
