"""Device context and device-resident folding state.

``Context`` wraps a ``zs_ctx`` (one HIP stream and a budgeted device memory
pool on one GPU). ``DeviceFoldState`` is the folding state a sumcheck round
threads through ``fold_rounds``: the stacked tables live on the device and are
folded there; each round advances it once.

There is no host fallback: creating a context without a GPU raises
``RuntimeError``.
"""

from __future__ import annotations

import ctypes
import threading
from typing import Any

import numpy as np

from zorch_amd import _lib
from zorch_amd.fields import Field, FieldArray, as_field, from_bytes


class Context:
    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().zs_ctx_create(device, hbm_budget_bytes, ctypes.byref(h)))
        self._h = h
        self.device = device

    @property
    def handle(self) -> ctypes.c_void_p:
        if self._h is None:
            raise RuntimeError("context is closed")
        return self._h

    def reserved_bytes(self) -> int:
        out = ctypes.c_uint64()
        _lib.check(_lib.lib().zs_ctx_reserved_bytes(self.handle, ctypes.byref(out)))
        return int(out.value)

    def get_tune(self) -> dict:
        """Launch shapes of later proofs: tail_pairs, max_grid, pairs_per_thread,
        chain_pairs, chain_grid, step_rounds, lookahead."""
        t = _lib.TuneC()
        _lib.check(_lib.lib().zs_ctx_get_tune(self.handle, ctypes.byref(t)))
        return {"tail_pairs": t.tail_pairs, "max_grid": t.max_grid,
                "pairs_per_thread": t.pairs_per_thread, "chain_pairs": t.chain_pairs,
                "chain_grid": t.chain_grid, "step_rounds": t.step_rounds,
                "lookahead": t.lookahead}

    def set_tune(self, **kw) -> None:
        """Change launch shapes (results never depend on them)."""
        cur = self.get_tune()
        unknown = set(kw) - set(cur)
        if unknown:
            raise ValueError(f"unknown tuning keys {sorted(unknown)}")
        cur.update(kw)
        t = _lib.TuneC(cur["tail_pairs"], cur["max_grid"], cur["pairs_per_thread"],
                       cur["chain_pairs"], cur["chain_grid"], cur["step_rounds"],
                       cur["lookahead"])
        _lib.check(_lib.lib().zs_ctx_set_tune(self.handle, ctypes.byref(t)))

    def close(self) -> None:
        if self._h is not None:
            _lib.lib().zs_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default: Context | None = None
_default_lock = threading.Lock()


def default_context() -> Context:
    """The process-wide context on device 0 (created on first use)."""
    global _default
    if _default is None:
        with _default_lock:
            if _default is None:
                _default = Context(0)
    return _default


def device_count() -> int:
    n = ctypes.c_int(0)
    rc = _lib.lib().zs_device_count(ctypes.byref(n))
    return int(n.value) if rc == 0 else 0


class DeviceTables:
    """Stacked factor tables resident in HBM (a device buffer of the context's
    pool), shape (k, 2^n), in storage encoding. Hands the prover inputs that
    are already on the GPU, so proving starts with no host-to-device copy.
    """

    def __init__(self, ptr: int, shape: tuple[int, int], field: Any, ctx: Context,
                 owned: bool = True):
        self._ptr = ptr
        self.shape = tuple(shape)
        self.field = as_field(field)
        self.ctx = ctx
        self._owned = owned
        if len(self.shape) != 2:
            raise ValueError("device tables must be (k, 2^n)")

    @classmethod
    def from_host(cls, tables: FieldArray, ctx: Context | None = None) -> "DeviceTables":
        ctx = ctx or default_context()
        if tables.ndim != 2:
            raise ValueError("device tables must be (k, 2^n)")
        words = np.ascontiguousarray(tables.words)
        p = ctypes.c_void_p()
        _lib.check(_lib.lib().zs_device_alloc(ctx.handle, words.nbytes, ctypes.byref(p)))
        _lib.check(_lib.lib().zs_memcpy_htod(ctx.handle, p, words.ctypes.data, words.nbytes))
        return cls(int(p.value), tables.shape, tables.field, ctx)

    @classmethod
    def empty(cls, shape: tuple[int, int], field: Any, ctx: Context | None = None) -> "DeviceTables":
        """Uninitialised device tables of `shape` in `field`."""
        ctx = ctx or default_context()
        f = as_field(field)
        nbytes = int(shape[0]) * int(shape[1]) * f.itemsize
        p = ctypes.c_void_p()
        _lib.check(_lib.lib().zs_device_alloc(ctx.handle, nbytes, ctypes.byref(p)))
        return cls(int(p.value), shape, f, ctx)

    def to_host(self) -> FieldArray:
        f = self.field
        out = np.zeros(self.shape + (f.words,), dtype=f.word_dtype)
        _lib.check(_lib.lib().zs_memcpy_dtoh(self.ctx.handle, out.ctypes.data,
                                             ctypes.c_void_p(self._ptr), out.nbytes))
        return FieldArray(out, f)

    @property
    def dtype(self) -> Field:
        return self.field

    def data_ptr(self) -> int:
        if self._ptr is None:
            raise ValueError("device tables were freed")
        return self._ptr

    def free(self) -> None:
        if self._ptr is not None and self._owned:
            _lib.check(_lib.lib().zs_device_free(self.ctx.handle, ctypes.c_void_p(self._ptr)))
        self._ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def tables_pointer(state) -> tuple[int, bool, Any, tuple[int, ...]]:
    """(pointer, on_device, keepalive, shape) of a stacked table argument."""
    if isinstance(state, DeviceTables):
        return state.data_ptr(), True, state, state.shape
    if isinstance(state, FieldArray):
        words = np.ascontiguousarray(state.words)
        return words.ctypes.data, False, words, state.shape
    raise TypeError("tables must be a FieldArray or DeviceTables")


def log2_strict(n: int) -> int:
    if n <= 0 or n & (n - 1):
        raise ValueError(f"{n} is not a power of two")
    return n.bit_length() - 1


def el_buffer(value: FieldArray, field: Field):
    """Contiguous storage words of `value` lifted into `field` (kept alive by the caller)."""
    return np.ascontiguousarray(value.astype(field).words)


def wire_struct(kind: int, domain: Any, chal_field: Any):
    """(``zs_wire`` or None, keepalive) for a message wire form: ZS_WIRE_VALUES
    (None), ZS_WIRE_DOMAIN with an EvalDomain, ZS_WIRE_COEFFS or
    ZS_WIRE_COMPRESSED."""
    if kind == _lib.ZS_WIRE_VALUES:
        return None, None
    if kind == _lib.ZS_WIRE_DOMAIN:
        return domain.wire(chal_field)
    return _lib.WireC(kind=kind), None


def message_width(args: "_lib.ProveArgsC", transcript=None) -> int:
    """Elements per round message of a request (``zs_prove_check``); raises
    ValueError on a malformed request."""
    out = ctypes.c_uint32(0)
    st = ctypes.byref(transcript) if transcript is not None else None
    _lib.check(_lib.lib().zs_prove_check(ctypes.byref(args), st, ctypes.byref(out)))
    return int(out.value)


class DeviceFoldState:
    """Device-resident folding state of one sumcheck (a ``zs_state``).

    Rounds advance it in place; ``bound`` counts the rounds done. A carry that
    refers to an earlier ``bound`` has been consumed and cannot be replayed.
    With ``refold_rounds`` the first rounds leave the input unfolded and
    refold it over the bound prefix (the sqrt-space engine).
    """

    def __init__(self, tables, claim: FieldArray, chal_field: Any, mode: int, degree: int,
                 eq_point: FieldArray | None = None, ctx: Context | None = None, *,
                 wire_kind: int = _lib.ZS_WIRE_VALUES, domain: Any = None,
                 eq_mass: FieldArray | None = None, refold_rounds: int = 0,
                 lam: FieldArray | None = None):
        self.ctx = ctx or default_context()
        ptr, on_dev, keep, shape = tables_pointer(tables)
        self.k, width = shape
        self.rounds = log2_strict(width)
        self.table_field = tables.dtype if isinstance(tables, DeviceTables) else tables.field
        self.chal_field = as_field(chal_field)
        self.mode = mode
        self.degree = degree
        self.npts = degree + 1 + (1 if mode == _lib.ZS_MODE_EQ else 0)
        claim_buf = el_buffer(claim, self.chal_field)
        lam_buf = el_buffer(lam, self.chal_field) if lam is not None else None
        eq_buf = el_buffer(eq_point, self.chal_field) if eq_point is not None else None
        wire, wire_keep = wire_struct(wire_kind, domain, self.chal_field)
        mass_buf = el_buffer(eq_mass, self.chal_field) if eq_mass is not None else None
        args = _lib.ProveArgsC(
            table_field=self.table_field.fid, chal_field=self.chal_field.fid, mode=mode,
            degree=degree, num_tables=self.k, rounds=self.rounds, tables=ptr,
            tables_on_device=1 if on_dev else 0, refold_rounds=refold_rounds,
            eq_point=eq_buf.ctypes.data if eq_buf is not None else None,
            claim=claim_buf.ctypes.data,
            wire=ctypes.pointer(wire) if wire is not None else None,
            eq_mass=mass_buf.ctypes.data if mass_buf is not None else None,
            lambda_=lam_buf.ctypes.data if lam_buf is not None else None)
        self.mlen = message_width(args)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().zs_state_create(self.ctx.handle, ctypes.byref(args), ctypes.byref(h)))
        self._h = h
        self._keep = (keep, lam_buf)
        self.bound = 0
        self.start_index = 0  # point index of this state's first round in its claim
        self.eq_point = eq_point

    def round(self, transcript):
        """One round: returns (msg, raw challenge, reduced claim, new transcript)."""
        f = self.chal_field
        msg = np.zeros((self.mlen, f.words), dtype=f.word_dtype)
        chal = np.zeros((f.words,), dtype=f.word_dtype)
        claim = np.zeros((f.words,), dtype=f.word_dtype)
        st = transcript.state.copy()
        _lib.check(_lib.lib().zs_state_round(self._h, ctypes.byref(st), msg.ctypes.data,
                                             chal.ctypes.data, claim.ctypes.data))
        self.bound += 1
        new_t = type(transcript)(st, transcript.dtype)
        return FieldArray(msg, f), FieldArray(chal, f), FieldArray(claim, f), new_t

    def read_tables(self) -> FieldArray:
        """Current folded tables (k, 2^(n - bound)) in the challenge field."""
        f = self.chal_field
        length = 1 << (self.rounds - self.bound)
        out = np.zeros((self.k, length, f.words), dtype=f.word_dtype)
        _lib.check(_lib.lib().zs_state_read_tables(self._h, out.ctypes.data, out.nbytes))
        return FieldArray(out, f)

    @property
    def shape(self) -> tuple[int, int]:
        return (self.k, 1 << (self.rounds - self.bound))

    def __del__(self):
        h = getattr(self, "_h", None)
        ctx = getattr(self, "ctx", None)
        # A state whose context is already closed (finalisers of a cycle run in
        # any order at shutdown) went with the context's memory.
        if h is not None and ctx is not None and ctx._h is not None:
            try:
                _lib.lib().zs_state_destroy(h)
            except Exception:
                pass
        self._h = None


def profile(ctx: Context, enable: bool) -> None:
    """Turn per-kernel event timing on or off (clears previous records)."""
    _lib.check(_lib.lib().zs_ctx_profile(ctx.handle, 1 if enable else 0))


def profile_read(ctx: Context, cap: int = 1 << 16) -> list[dict]:
    """Kernel records since the last read: round, kind, grid, algorithmic bytes, ms."""
    recs = (_lib.KernelRecordC * cap)()
    n = ctypes.c_uint32(0)
    _lib.check(_lib.lib().zs_ctx_profile_read(ctx.handle, recs, cap, ctypes.byref(n)))
    kinds = {1: "pass", 2: "finish", 3: "tail", 4: "batch", 5: "chain", 6: "pass_la",
             7: "chain_la", 8: "tail_la", 9: "fused01", 10: "fold2", 11: "refold", 12: "engine",
             13: "logup", 14: "jagged", 15: "jagged_batch"}
    return [{"round": r.round, "kind": kinds.get(r.kind, str(r.kind)), "grid": r.grid,
             "bytes": int(r.bytes), "ms": float(r.ms)} for r in recs[: n.value]]


def trace(ctx: Context, enable: bool) -> None:
    """Turn the in-kernel phase trace on or off (clears it)."""
    _lib.check(_lib.lib().zs_ctx_trace(ctx.handle, 1 if enable else 0))


def trace_read(ctx: Context, cap: int = 2047) -> list[tuple[int, int, float]]:
    """Phase events since the last read as (round, phase, microseconds)."""
    buf = np.zeros(2 * cap, dtype=np.uint64)
    n = ctypes.c_size_t(0)
    khz = ctypes.c_uint64(0)
    _lib.check(_lib.lib().zs_ctx_trace_read(ctx.handle, buf.ctypes.data, cap, ctypes.byref(n),
                                            ctypes.byref(khz)))
    rate = max(int(khz.value), 1) / 1e3  # ticks per microsecond
    ev = buf[: 2 * n.value].reshape(-1, 2)
    return [(int(k) >> 8, int(k) & 0xFF, float(t) / rate) for k, t in ev]
