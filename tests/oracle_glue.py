"""Test-side glue: run the oracles on a workload instance and compare.

Only tests, the smoke check and bench.py's CPU-baseline leg use this module.
"""

from __future__ import annotations

import ctypes
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from oracle import fields as OF
from oracle import sumcheck as OS
from oracle import transcript as OT

ROOT = Path(__file__).resolve().parents[1]
OREF = ROOT / "oracle" / "_build" / "liboref.so"

FIELD_IDS = {"babybear_mont": 1, "babybearx4_mont": 2, "goldilocks": 3}
OFIELD = {"babybear_mont": OF.BABYBEAR, "babybearx4_mont": OF.BABYBEAR4,
          "goldilocks": OF.GOLDILOCKS}


class OrefSha(ctypes.Structure):
    _fields_ = [("h", ctypes.c_uint32 * 8), ("pending", ctypes.c_uint8 * 64),
                ("pending_len", ctypes.c_uint64), ("total_len", ctypes.c_uint64)]


_oref = None


def oref() -> ctypes.CDLL:
    global _oref
    if _oref is None:
        if not OREF.exists():
            subprocess.run(["make", "-C", str(ROOT / "oracle")], check=True,
                           capture_output=True)
        lib = ctypes.CDLL(str(OREF))
        lib.oref_prove.argtypes = [ctypes.c_int] * 7 + [ctypes.c_void_p] * 8 + [ctypes.c_int]
        lib.oref_prove.restype = ctypes.c_int
        lib.oref_claim.argtypes = [ctypes.c_int] * 5 + [ctypes.c_void_p] * 3 + [ctypes.c_int]
        lib.oref_claim.restype = ctypes.c_int
        _oref = lib
    return _oref


@dataclass
class OracleResult:
    msgs: np.ndarray     # storage words (n, npts, words)
    point: np.ndarray    # raw words (n, words)
    value: np.ndarray    # (words,)
    final: np.ndarray    # (k [+1], words)
    h: tuple
    pending: bytes
    total: int
    state: tuple = ()    # the transcript's stream_state() form (either hash)


def sha_from_stream(stream: bytes) -> OrefSha:
    from oracle import sha256 as S
    h, pend, tot = S.midstate(stream)
    s = OrefSha()
    for i in range(8):
        s.h[i] = h[i]
    for i, b in enumerate(pend):
        s.pending[i] = b
    s.pending_len = len(pend)
    s.total_len = tot
    return s


def initial_stream(label: bytes) -> bytes:
    return OT.OracleTranscript.new(label, 4).stream


def run_c(inst, threads: int = 0, stream: bytes | None = None) -> OracleResult:
    """Run the C oracle on a workloads.Instance (SHA-256 transcript)."""
    if inst.transcript_kind != "sha256":
        raise ValueError("the C oracle restates the SHA-256 transcript only")
    lib = oref()
    tf, cf, trf = inst.tables.field, inst.chal, inst.transcript_field
    eq = 1 if inst.mode == "zerocheck" else 0
    k, N = inst.tables.shape
    n = inst.n
    npts = inst.degree + 1 + eq
    tabs = np.ascontiguousarray(inst.tables.words)
    claim = np.ascontiguousarray(inst.claim.words)
    z = np.ascontiguousarray(inst.z.words) if inst.z is not None else np.zeros(1, np.uint32)
    msgs = np.zeros((n, npts, cf.words), dtype=cf.word_dtype)
    point = np.zeros((n, cf.words), dtype=cf.word_dtype)
    value = np.zeros((cf.words,), dtype=cf.word_dtype)
    final = np.zeros((k + 1, cf.words), dtype=cf.word_dtype)
    st = sha_from_stream(stream if stream is not None
                         else OT.OracleTranscript.new(inst.label, trf.itemsize).stream)
    rc = lib.oref_prove(FIELD_IDS[tf.name], FIELD_IDS[cf.name], FIELD_IDS[trf.name], eq,
                        inst.degree, k, n, tabs.ctypes.data, z.ctypes.data, claim.ctypes.data,
                        ctypes.addressof(st), msgs.ctypes.data, point.ctypes.data,
                        value.ctypes.data, final.ctypes.data, threads)
    assert rc == 0, rc
    h, pend, tot = tuple(st.h), bytes(st.pending[:st.pending_len]), int(st.total_len)
    return OracleResult(msgs, point, value, final[:k + eq], h, pend, tot, (h, pend, tot))


def run_py(inst) -> tuple[OracleResult, OS.OracleProof]:
    """Run the Python oracle on a (small) workloads.Instance."""
    tf, cf, trf = (OFIELD[inst.tables.field.name], OFIELD[inst.chal.name],
                   OFIELD[inst.transcript_field.name])
    tabs = [[tf.from_words([int(w) for w in inst.tables.words[q, x]])
             for x in range(inst.tables.shape[1])] for q in range(inst.tables.shape[0])]
    claim = cf.from_words([int(w) for w in inst.claim.words.reshape(-1)])
    tcls = OT.OracleBlake3Transcript if inst.transcript_kind == "blake3" else OT.OracleTranscript
    tr = tcls.new(inst.label, trf.item_bytes)
    if inst.mode == "zerocheck":
        z = [cf.from_words([int(w) for w in inst.z.words[i]]) for i in range(inst.n)]
        pf = OS.prove_eq(cf, trf, tabs, z, claim, tr, inst.degree)
    else:
        pf = OS.prove_product(cf, trf, tabs, claim, tr, inst.degree)
    wd = np.uint64 if cf is OF.GOLDILOCKS else np.uint32
    msgs = np.array([[cf.to_words(m) for m in row] for row in pf.msgs], dtype=wd)
    point = np.array(pf.point_words, dtype=wd)
    value = np.array(cf.to_words(pf.value), dtype=wd)
    fin = [cf.to_words(t[0]) for t in pf.final_tables]
    if inst.mode == "zerocheck":
        fin.append(cf.to_words(pf.eq_prev))
    state = pf.transcript.state()
    h, pend, tot = state if inst.transcript_kind == "sha256" else ((), b"", 0)
    return OracleResult(msgs, point, value, np.array(fin, dtype=wd), h, pend, tot, state), pf


def oracle_inputs(inst):
    """(table field, challenge field, transcript field, tables, claim, fresh
    transcript) of an instance in the oracle's representation."""
    tf, cf, trf = (OFIELD[inst.tables.field.name], OFIELD[inst.chal.name],
                   OFIELD[inst.transcript_field.name])
    tabs = [[tf.from_words([int(w) for w in inst.tables.words[q, x]])
             for x in range(inst.tables.shape[1])] for q in range(inst.tables.shape[0])]
    claim = cf.from_words([int(w) for w in inst.claim.words.reshape(-1)])
    tcls = OT.OracleBlake3Transcript if inst.transcript_kind == "blake3" else OT.OracleTranscript
    return tf, cf, trf, tabs, claim, tcls.new(inst.label, trf.item_bytes)


def proof_result(inst, cf, pf: OS.OracleProof, eq: bool = False) -> OracleResult:
    """An oracle proof as the storage words the device returns."""
    wd = np.uint64 if cf is OF.GOLDILOCKS else np.uint32
    msgs = np.array([[cf.to_words(m) for m in row] for row in pf.msgs], dtype=wd)
    point = np.array(pf.point_words, dtype=wd)
    value = np.array(cf.to_words(pf.value), dtype=wd)
    fin = [cf.to_words(t[0]) for t in pf.final_tables]
    if eq:
        fin.append(cf.to_words(pf.eq_prev))
    state = pf.transcript.state()
    h, pend, tot = state if inst.transcript_kind == "sha256" else ((), b"", 0)
    return OracleResult(msgs, point, value, np.array(fin, dtype=wd), h, pend, tot, state)


def oracle_wire(inst, wire_py) -> tuple:
    """(F, wire) of the oracle for a package wire description: ("values",),
    ("domain", EvalDomain), ("coeffs",) or ("compressed",)."""
    cf = OFIELD[inst.chal.name]
    if wire_py[0] != "domain":
        return cf, wire_py
    dom = wire_py[1]
    nodes = [cf.embed(OFIELD[dom.nodes.field.name].from_words(
        [int(w) for w in dom.nodes.words[i]])) for i in range(dom.nodes.shape[0])]
    return cf, ("domain", nodes, dom.inf_index)


def run_py_wire(inst, wire_py, eq_mass=None) -> tuple[OracleResult, OS.OracleProof]:
    """The Python oracle's stage with `wire_py` messages (small instances)."""
    tf, cf, trf = (OFIELD[inst.tables.field.name], OFIELD[inst.chal.name],
                   OFIELD[inst.transcript_field.name])
    tabs = [[tf.from_words([int(w) for w in inst.tables.words[q, x]])
             for x in range(inst.tables.shape[1])] for q in range(inst.tables.shape[0])]
    claim = cf.from_words([int(w) for w in inst.claim.words.reshape(-1)])
    tcls = OT.OracleBlake3Transcript if inst.transcript_kind == "blake3" else OT.OracleTranscript
    tr = tcls.new(inst.label, trf.item_bytes)
    _, wire = oracle_wire(inst, wire_py)
    mass = None if eq_mass is None else cf.from_words([int(w) for w in eq_mass.words.reshape(-1)])
    if inst.mode == "zerocheck":
        z = [cf.from_words([int(w) for w in inst.z.words[i]]) for i in range(inst.n)]
        pf = OS.prove_eq_wire(cf, trf, tabs, z, claim, tr, inst.degree, wire, mass)
    else:
        pf = OS.prove_product_wire(cf, trf, tabs, claim, tr, inst.degree, wire)
    wd = np.uint64 if cf is OF.GOLDILOCKS else np.uint32
    msgs = np.array([[cf.to_words(m) for m in row] for row in pf.msgs], dtype=wd)
    point = np.array(pf.point_words, dtype=wd)
    value = np.array(cf.to_words(pf.value), dtype=wd)
    fin = [cf.to_words(t[0]) for t in pf.final_tables]
    if inst.mode == "zerocheck":
        fin.append(cf.to_words(pf.eq_prev))
    state = pf.transcript.state()
    h, pend, tot = state if inst.transcript_kind == "sha256" else ((), b"", 0)
    return OracleResult(msgs, point, value, np.array(fin, dtype=wd), h, pend, tot, state), pf
