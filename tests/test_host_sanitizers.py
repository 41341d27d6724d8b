"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (no GPU).

`tests/sanitize/host_check.cpp` is built host-only with ASan and UBSan
(recoverable checks turned into failures) together with the C oracle
(`oracle/sumcheck_ref.c`), and driven over stdin:

* the host field arithmetic of `zorch_amd/csrc/field.h` (BabyBear,
  BabyBear^4, Goldilocks) against 128-bit integer references, with edge
  operands that reach the Goldilocks reduction's borrow, carry and final
  subtraction;
* the SHA-256 stream of `csrc/sha256.h` against hashlib, lengths 0..130 and
  1000;
* the BLAKE3 stream of `csrc/blake3.h` against the oracle's BLAKE3 over chunk
  and subtree boundaries, absorbed in two parts;
* one Fiat-Shamir hop of either field transcript (the host path of
  `zs_transcript_*` and `zs_verify_round`) against the oracle transcript;
* a whole zerocheck proof of the C oracle against the same oracle built
  without sanitizers.

Any sanitizer report ends the harness with a non-zero status.
"""

from __future__ import annotations

import hashlib
import os
import random
import shutil
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import blake3 as OB3
from oracle.transcript import OracleBlake3Transcript, OracleTranscript

ROOT = Path(__file__).resolve().parents[1]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CLANG = "/opt/rocm/llvm/bin/clang"
CLANGXX = "/opt/rocm/llvm/bin/clang++"
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
       "-O1", "-g"]

pytestmark = pytest.mark.skipif(not (Path(HIPCC).exists() and Path(CLANGXX).exists()),
                                reason="ROCm toolchain not present")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    d = tmp_path_factory.mktemp("san")
    oref_o, host_o, exe = d / "oref.o", d / "host.o", d / "host_check"
    subprocess.run([CLANG, *SAN, "-c", str(ROOT / "oracle/sumcheck_ref.c"), "-o", str(oref_o)],
                   check=True)
    # Host pass only; each sanitizer flag applies to the host compilation.
    subprocess.run([HIPCC, "--cuda-host-only", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
                    "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                    "-Xarch_host", "-fno-sanitize-recover=all",
                    "-x", "hip", "-c", str(ROOT / "tests/sanitize/host_check.cpp"), "-o", str(host_o)],
                   check=True)
    subprocess.run([CLANGXX, "-fsanitize=address,undefined", str(host_o), str(oref_o),
                    "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib", "-o", str(exe)],
                   check=True)
    return exe


def _run(exe, lines: list[str]) -> list[str]:
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    res = subprocess.run([str(exe)], input="\n".join(lines) + "\n", capture_output=True,
                         text=True, env=env, timeout=600)
    assert res.returncode == 0, res.stderr[-4000:]
    assert "runtime error" not in res.stderr and "AddressSanitizer" not in res.stderr, \
        res.stderr[-4000:]
    return res.stdout.split("\n")


def _hx(b: bytes) -> str:
    return b.hex() if b else "-"


def test_field_arithmetic(harness):
    assert _run(harness, ["field"])[0] == "ok"


def test_sha256_stream(harness):
    rng = random.Random(1)
    msgs = [bytes(rng.randrange(256) for _ in range(n)) for n in list(range(131)) + [1000]]
    out = _run(harness, [f"sha {_hx(m)}" for m in msgs])
    for m, got in zip(msgs, out):
        assert got == hashlib.sha256(m).hexdigest()


def test_blake3_stream(harness):
    rng = random.Random(2)
    cases = []
    for n in (0, 1, 63, 64, 65, 1023, 1024, 1025, 2048, 3073, 5000):
        m = bytes(rng.randrange(256) for _ in range(n))
        cases.append((m, rng.randrange(n + 1), rng.choice((32, 64, 100, 250))))
    out = _run(harness, [f"b3 {_hx(m)} {split} {nb}" for m, split, nb in cases])
    for (m, _, nb), got in zip(cases, out):
        assert got == OB3.blake3(m, nb).hex()


def test_transcript_hops(harness):
    rng = random.Random(3)
    lines, want = [], []
    for kind, cls in ((0, OracleTranscript), (1, OracleBlake3Transcript)):
        for item, limbs, mlen in ((4, 4, 64), (4, 1, 12), (8, 1, 32), (16, 1, 96), (4, 4, 500)):
            t = cls.new(b"sanitize-%d" % rng.randrange(1000), item)
            t = t.absorb(bytes(rng.randrange(256) for _ in range(rng.randrange(200))))
            msg = bytes(rng.randrange(256) for _ in range(mlen))
            lines.append(f"hop {kind} {item} {limbs} {_hx(t.stream)} {_hx(msg)}")
            want.append(t.observe_and_sample(msg, limbs)[1].hex())
    assert _run(harness, lines)[:len(want)] == want


def test_oracle_proof_under_sanitizers(harness, tmp_path):
    from tests.oracle_glue import run_c
    from zorch_amd import workloads as W

    inst = W.zerocheck_bb4(8, (61, 62), 63, 64)
    ref = run_c(inst, threads=1)
    stream = OracleTranscript.new(inst.label, 4).stream
    req = tmp_path / "req.bin"
    req.write_bytes(struct.pack("<I", inst.n) + np.ascontiguousarray(inst.tables.words,
                                                                      dtype="<u4").tobytes()
                    + np.ascontiguousarray(inst.z.words, dtype="<u4").tobytes() + stream)
    msgs, point, value = _run(harness, [f"oref {req}"])[:3]
    assert msgs == np.ascontiguousarray(ref.msgs, dtype="<u4").tobytes().hex()
    assert point == np.ascontiguousarray(ref.point, dtype="<u4").tobytes().hex()
    assert value == np.ascontiguousarray(ref.value, dtype="<u4").tobytes().hex()
