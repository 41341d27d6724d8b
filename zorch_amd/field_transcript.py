"""Field transcripts over the library's streaming hashes (shared surface).

``FieldTranscript`` is the reference's field-transcript surface
(zorch/sha256_field_transcript.py:156-378, zorch/blake3_field_transcript.py:
247-409): a frozen value holding one streaming hash state in the C ABI's
``zs_transcript`` layout, so the same bytes can be handed to the device
prover, which runs the transcript's hop on the GPU. Framing is the
reference's Merlin wire (zorch/byte_transcript.py:37-54):

* ``new(domain)``   [OP_DOMAIN] || le64(len) || domain
* ``observe(V)``    [OP_OBSERVE, KIND_SLICE] || le64(|bytes(V)|/itemsize) || bytes(V)
* ``sample(n)``     [OP_SQUEEZE, KIND_SLICE] || le64(n), then the hash's
                    squeeze of n * itemsize bytes, re-absorbed.

Subclasses fix the hash (``KIND``): SHA-256's counter squeeze or BLAKE3's
extendable output. The hashing is the library's host code (``zs_transcript_*``).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Any

import numpy as np

from zorch_amd import _lib
from zorch_amd.byte_transcript import (
    KIND_SCALAR,
    KIND_SLICE,
    OP_BYTES,
    OP_LABEL,
    OP_OBSERVE,
    OP_SQUEEZE,
    _len8,
)
from zorch_amd.fields import Field, FieldArray, as_field, from_bytes

_RAW_IDS = {4: _lib.ZS_F_BABYBEAR, 8: _lib.ZS_F_GOLDILOCKS}


def _field_id(dtype: Any) -> tuple[int, int]:
    """(C field id, item bytes) of a transcript dtype: a field, or a raw
    unsigned word dtype of 4 or 8 bytes."""
    if isinstance(dtype, Field) or isinstance(dtype, str):
        f = as_field(dtype)
        return f.fid, f.itemsize
    item = np.dtype(dtype).itemsize
    if item not in _RAW_IDS:
        raise ValueError(f"unsupported transcript word dtype {dtype!r}")
    return _RAW_IDS[item], item


@dataclass(frozen=True)
class FieldTranscript:
    state: _lib.TranscriptC
    dtype: Any

    KIND = 0  # ZS_TR_* of the subclass

    @property
    def field(self) -> Any:
        return self.dtype

    @property
    def item_bytes(self) -> int:
        return int(self.state.item_bytes)

    # ---- construction / absorb / squeeze --------------------------------
    @classmethod
    def new(cls, domain: bytes, dtype: Any):
        fid, _ = _field_id(dtype)
        st = _lib.TranscriptC()
        domain = bytes(domain)
        buf = ctypes.create_string_buffer(domain, len(domain)) if domain else None
        _lib.check(_lib.lib().zs_transcript_new(cls.KIND, fid, buf, len(domain),
                                                ctypes.byref(st)))
        if not isinstance(dtype, (Field, str)):
            st.item_bytes = np.dtype(dtype).itemsize
        return cls(st, dtype if not isinstance(dtype, str) else as_field(dtype))

    def _absorb(self, payload: bytes):
        st = self.state.copy()
        if payload:
            buf = ctypes.create_string_buffer(payload, len(payload))
            _lib.check(_lib.lib().zs_transcript_absorb(ctypes.byref(st), buf, len(payload)))
        return type(self)(st, self.dtype)

    def _squeeze_after(self, framing: bytes, nbytes: int):
        t = self._absorb(framing)
        st = t.state
        out = ctypes.create_string_buffer(max(nbytes, 1))
        _lib.check(_lib.lib().zs_transcript_squeeze(ctypes.byref(st), nbytes, out))
        return t, out.raw[:nbytes]

    def _words(self, data: bytes, n: int) -> FieldArray | np.ndarray:
        if isinstance(self.dtype, Field):
            return from_bytes(data, self.dtype, (n,))
        return np.frombuffer(data, dtype=np.dtype(self.dtype).newbyteorder("<")).astype(
            self.dtype)[:n]

    @staticmethod
    def _elem_bytes(values) -> bytes:
        if isinstance(values, FieldArray):
            return values.tobytes()
        arr = np.asarray(values)
        return arr.astype(arr.dtype.newbyteorder("<")).tobytes()

    # ---- the Transcript surface -----------------------------------------
    def observe(self, values):
        vb = self._elem_bytes(values)
        count = len(vb) // self.item_bytes
        return self._absorb(bytes([OP_OBSERVE, KIND_SLICE]) + _len8(count) + vb)

    def observe_scalar(self, value):
        vb = self._elem_bytes(value)
        item = self.item_bytes
        payload = b"".join(bytes([OP_OBSERVE, KIND_SCALAR]) + vb[i:i + item]
                           for i in range(0, len(vb), item))
        return self._absorb(payload)

    def observe_label(self, label: bytes):
        return self._absorb(bytes([OP_LABEL]) + _len8(len(label)) + bytes(label))

    def observe_bytes(self, data):
        data = bytes(np.asarray(data, dtype=np.uint8).reshape(-1)) \
            if not isinstance(data, (bytes, bytearray)) else bytes(data)
        return self._absorb(bytes([OP_BYTES]) + _len8(len(data)) + data)

    def sample(self, n: int = 1):
        t, raw = self._squeeze_after(bytes([OP_SQUEEZE, KIND_SLICE]) + _len8(n),
                                     n * self.item_bytes)
        return t, self._words(raw, n)

    def sample_scalar(self):
        t, raw = self._squeeze_after(bytes([OP_SQUEEZE, KIND_SCALAR]), self.item_bytes)
        return t, self._words(raw, 1)[0]

    def observe_and_sample(self, values, n: int = 1):
        return self.observe(values).sample(n)

    def observe_scalar_and_sample(self, value):
        return self.observe_scalar(value).sample_scalar()

    def digest(self) -> bytes:
        """The 32-byte hash of the stream so far (the proof-of-work state digest)."""
        out = ctypes.create_string_buffer(32)
        st = self.state.copy()
        _lib.check(_lib.lib().zs_transcript_digest(ctypes.byref(st), out))
        return out.raw

    # ---- state inspection ------------------------------------------------
    def stream_state(self) -> tuple:
        raise NotImplementedError

    def __eq__(self, other) -> bool:
        return (type(other) is type(self)
                and self.stream_state() == other.stream_state()
                and self.item_bytes == other.item_bytes)

    def __hash__(self) -> int:
        return hash(self.stream_state())
