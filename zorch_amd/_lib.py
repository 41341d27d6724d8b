"""ctypes binding of ``libzorch_amd.so`` (the C ABI in include/zorch_amd.h).

The library is built in-tree by ``python -m zorch_amd.build``. There is no
fallback: if the library is missing, or no GPU is present when a device call
is made, the call raises.
"""

from __future__ import annotations

import ctypes
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libzorch_amd.so"

ZS_OK, ZS_EINVAL, ZS_ENOMEM, ZS_EHIP, ZS_EBUDGET, ZS_ENODEV = 0, 1, 2, 3, 4, 5
ZS_F_BABYBEAR, ZS_F_BABYBEAR4, ZS_F_GOLDILOCKS = 1, 2, 3
ZS_TR_SHA256, ZS_TR_BLAKE3 = 1, 2
ZS_MODE_PRODUCT, ZS_MODE_EQ, ZS_MODE_LOGUP = 1, 2, 3
ZS_MAX_ROUNDS = 31
ZS_NOALLOC_SCRATCH_BYTES = 64 << 20
ZS_WIRE_VALUES, ZS_WIRE_DOMAIN, ZS_WIRE_COEFFS, ZS_WIRE_COMPRESSED = 0, 1, 2, 3
ZS_OP_ADD, ZS_OP_SUB, ZS_OP_MUL = 1, 2, 3


class Sha256StateC(ctypes.Structure):
    _fields_ = [
        ("h", ctypes.c_uint32 * 8),
        ("pending", ctypes.c_uint8 * 64),
        ("pending_len", ctypes.c_uint64),
        ("total_len", ctypes.c_uint64),
    ]


class Blake3StateC(ctypes.Structure):
    _fields_ = [
        ("cv", ctypes.c_uint32 * 8),
        ("stack", (ctypes.c_uint32 * 8) * 16),
        ("block", ctypes.c_uint8 * 64),
        ("chunk_counter", ctypes.c_uint64),
        ("block_len", ctypes.c_uint32),
        ("blocks_done", ctypes.c_uint32),
        ("stack_len", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class _StreamC(ctypes.Union):
    _fields_ = [("sha", Sha256StateC), ("blake3", Blake3StateC)]


class TranscriptC(ctypes.Structure):
    _anonymous_ = ("stream",)
    _fields_ = [
        ("kind", ctypes.c_uint32),
        ("field", ctypes.c_uint32),
        ("item_bytes", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("stream", _StreamC),
    ]

    def copy(self) -> "TranscriptC":
        out = TranscriptC()
        ctypes.memmove(ctypes.byref(out), ctypes.byref(self), ctypes.sizeof(TranscriptC))
        return out


class WireC(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_uint32),
        ("num_nodes", ctypes.c_uint32),
        ("has_inf", ctypes.c_uint32),
        ("inf_pos", ctypes.c_uint32),
        ("nodes", ctypes.c_void_p),
    ]


class ProveArgsC(ctypes.Structure):
    _fields_ = [
        ("table_field", ctypes.c_uint32),
        ("chal_field", ctypes.c_uint32),
        ("mode", ctypes.c_uint32),
        ("degree", ctypes.c_uint32),
        ("num_tables", ctypes.c_uint32),
        ("rounds", ctypes.c_uint32),
        ("tables", ctypes.c_void_p),
        ("tables_on_device", ctypes.c_uint32),
        ("refold_rounds", ctypes.c_uint32),
        ("eq_point", ctypes.c_void_p),
        ("claim", ctypes.c_void_p),
        ("wire", ctypes.POINTER(WireC)),
        ("eq_mass", ctypes.c_void_p),
        ("lambda_", ctypes.c_void_p),
    ]


class ProofC(ctypes.Structure):
    _fields_ = [
        ("msgs", ctypes.c_void_p),
        ("point", ctypes.c_void_p),
        ("value", ctypes.c_void_p),
        ("final_tables", ctypes.c_void_p),
    ]


class JaggedArgsC(ctypes.Structure):
    _fields_ = [
        ("field", ctypes.c_uint32),
        ("num_segments", ctypes.c_uint32),
        ("row_variables", ctypes.c_uint32),
        ("planes_on_device", ctypes.c_uint32),
        ("row_counts", ctypes.c_void_p),
        ("planes", ctypes.c_void_p),
        ("eval_point", ctypes.c_void_p),
        ("lambda_", ctypes.c_void_p),
        ("claim", ctypes.c_void_p),
    ]


class JaggedProofC(ctypes.Structure):
    _fields_ = [
        ("round_polys", ctypes.c_void_p),
        ("challenges", ctypes.c_void_p),
        ("openings", ctypes.c_void_p),
        ("value", ctypes.c_void_p),
    ]


class KernelRecordC(ctypes.Structure):
    _fields_ = [
        ("round", ctypes.c_uint32),
        ("kind", ctypes.c_uint32),
        ("grid", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("bytes", ctypes.c_uint64),
        ("ms", ctypes.c_double),
    ]


class TuneC(ctypes.Structure):
    _fields_ = [
        ("tail_pairs", ctypes.c_uint32),
        ("max_grid", ctypes.c_uint32),
        ("pairs_per_thread", ctypes.c_uint32),
        ("chain_pairs", ctypes.c_uint32),
        ("chain_grid", ctypes.c_uint32),
        ("step_rounds", ctypes.c_uint32),
        ("lookahead", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32 * 1),
    ]


class ServerStatsC(ctypes.Structure):
    _fields_ = [(name, ctypes.c_uint64) for name in (
        "submitted", "admitted", "completed", "failed", "steps", "queued", "active",
        "resident_bytes", "peak_resident_bytes", "peak_active", "budget_bytes")] + [
        ("busy_s", ctypes.c_double)]


assert ctypes.sizeof(Sha256StateC) == 112
assert ctypes.sizeof(Blake3StateC) == 632
assert ctypes.sizeof(TranscriptC) == 648
assert ctypes.sizeof(ProveArgsC) == 80
assert ctypes.sizeof(JaggedArgsC) == 56

_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
_U64 = ctypes.c_uint64
_SZ = ctypes.c_size_t
_INT = ctypes.c_int

# name -> (restype, argtypes); every symbol include/zorch_amd.h declares.
SIGNATURES = {
    "zs_last_error": (ctypes.c_char_p, []),
    "zs_version": (ctypes.c_char_p, []),
    "zs_device_count": (_INT, [ctypes.POINTER(_INT)]),
    "zs_ctx_create": (_INT, [_INT, _U64, ctypes.POINTER(_P)]),
    "zs_ctx_destroy": (None, [_P]),
    "zs_ctx_reserved_bytes": (_INT, [_P, ctypes.POINTER(_U64)]),
    "zs_device_alloc": (_INT, [_P, _SZ, ctypes.POINTER(_P)]),
    "zs_device_free": (_INT, [_P, _P]),
    "zs_memcpy_htod": (_INT, [_P, _P, _P, _SZ]),
    "zs_memcpy_dtoh": (_INT, [_P, _P, _P, _SZ]),
    "zs_memcpy_dtod": (_INT, [_P, _P, _P, _SZ]),
    "zs_ctx_trace": (_INT, [_P, _INT]),
    "zs_ctx_trace_read": (_INT, [_P, _P, _SZ, ctypes.POINTER(_SZ), ctypes.POINTER(_U64)]),
    "zs_ctx_get_tune": (_INT, [_P, ctypes.POINTER(TuneC)]),
    "zs_ctx_set_tune": (_INT, [_P, ctypes.POINTER(TuneC)]),
    "zs_ctx_profile": (_INT, [_P, _INT]),
    "zs_ctx_profile_read": (_INT, [_P, ctypes.POINTER(KernelRecordC), _U32,
                                   ctypes.POINTER(_U32)]),
    "zs_transcript_new": (_INT, [_U32, _U32, _P, _SZ, ctypes.POINTER(TranscriptC)]),
    "zs_transcript_absorb": (_INT, [ctypes.POINTER(TranscriptC), _P, _SZ]),
    "zs_transcript_squeeze": (_INT, [ctypes.POINTER(TranscriptC), _SZ, _P]),
    "zs_transcript_digest": (_INT, [ctypes.POINTER(TranscriptC), _P]),
    "zs_prove_check": (_INT, [ctypes.POINTER(ProveArgsC), ctypes.POINTER(TranscriptC),
                              ctypes.POINTER(_U32)]),
    "zs_field_op": (_INT, [_P, _U32, _U32, _P, _P, _P, _SZ]),
    "zs_hbm_bandwidth": (_INT, [_P, _SZ, _U32, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_double)]),
    "zs_prove": (_INT, [_P, ctypes.POINTER(ProveArgsC), ctypes.POINTER(TranscriptC),
                        ctypes.POINTER(ProofC)]),
    "zs_prove_noalloc": (_INT, [_P, ctypes.POINTER(ProveArgsC), ctypes.POINTER(TranscriptC),
                                ctypes.POINTER(ProofC)]),
    "zs_prove_batch": (_INT, [_P, _SZ, ctypes.POINTER(ProveArgsC), ctypes.POINTER(TranscriptC),
                              ctypes.POINTER(ProofC)]),
    "zs_prove_univariate_skip": (_INT, [_P, ctypes.POINTER(ProveArgsC), _U32,
                                        ctypes.POINTER(TranscriptC), _P, ctypes.POINTER(ProofC)]),
    "zs_univariate_skip_round0": (_INT, [_P, ctypes.POINTER(ProveArgsC), _U32,
                                         ctypes.POINTER(TranscriptC), _P, _P, _P, _P]),
    "zs_prove_small_value": (_INT, [_P, ctypes.POINTER(ProveArgsC), _U32,
                                    ctypes.POINTER(TranscriptC), ctypes.POINTER(ProofC)]),
    "zs_logup_transition": (_INT, [_P, _U32, _P, _SZ, _P]),
    "zs_logup_eq": (_INT, [_P, _U32, _P, _U32, _P]),
    "zs_jagged_transition": (_INT, [_P, _U32, _U32, _P, _P, _P, _SZ, _P, _SZ]),
    "zs_prove_jagged_layer": (_INT, [_P, ctypes.POINTER(JaggedArgsC),
                                     ctypes.POINTER(TranscriptC), ctypes.POINTER(JaggedProofC)]),
    "zs_state_create": (_INT, [_P, ctypes.POINTER(ProveArgsC), ctypes.POINTER(_P)]),
    "zs_state_round": (_INT, [_P, ctypes.POINTER(TranscriptC), _P, _P, _P]),
    "zs_state_bound": (_INT, [_P, ctypes.POINTER(_U32)]),
    "zs_state_read_tables": (_INT, [_P, _P, _SZ]),
    "zs_state_destroy": (None, [_P]),
    "zs_verify_round": (_INT, [_U32, ctypes.POINTER(TranscriptC), _P, _P, _U32, _U32, _P, _P,
                               ctypes.POINTER(_INT)]),
    "zs_server_create": (_INT, [_P, _U64, ctypes.POINTER(_P)]),
    "zs_server_submit": (_INT, [_P, ctypes.POINTER(ProveArgsC), ctypes.POINTER(TranscriptC),
                                ctypes.POINTER(_U64)]),
    "zs_server_poll": (_INT, [_P, _U64, ctypes.POINTER(ProofC), ctypes.POINTER(TranscriptC),
                              ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_INT)]),
    "zs_server_wait": (_INT, [_P, _U64, ctypes.POINTER(ProofC), ctypes.POINTER(TranscriptC),
                              ctypes.POINTER(ctypes.c_double), ctypes.c_double,
                              ctypes.POINTER(_INT)]),
    "zs_server_completed": (_INT, [_P, ctypes.POINTER(_U64), _SZ, ctypes.POINTER(_SZ)]),
    "zs_server_stats_read": (_INT, [_P, ctypes.POINTER(ServerStatsC)]),
    "zs_server_destroy": (None, [_P]),
}

_lib = None
_lock = threading.Lock()


class LibraryMissing(RuntimeError):
    pass


def lib() -> ctypes.CDLL:
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not LIB_PATH.exists():
                    raise LibraryMissing(
                        f"{LIB_PATH} is not built; run `python -m zorch_amd.build`")
                handle = ctypes.CDLL(str(LIB_PATH))
                for name, (res, args) in SIGNATURES.items():
                    fn = getattr(handle, name)
                    fn.restype = res
                    fn.argtypes = args
                _lib = handle
    return _lib


def last_error() -> str:
    msg = lib().zs_last_error()
    return msg.decode() if msg else ""


def check(rc: int) -> None:
    """Map a C ABI return code to the reference's exception types."""
    if rc == ZS_OK:
        return
    msg = last_error()
    if rc == ZS_EINVAL:
        raise ValueError(msg)
    if rc in (ZS_ENOMEM, ZS_EBUDGET):
        raise MemoryError(msg)
    raise RuntimeError(msg or f"zorch_amd error {rc}")
