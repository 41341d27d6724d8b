"""GPU parity of the continuous-batching proof server.

Requests submitted at different times (so that later ones join while earlier
ones are mid-proof) must each equal the proof of the same request made alone,
bit for bit, whether they run in the batched steps or, when large, whole on
the server's own streams; admission respects the HBM budget; a request larger than the
whole budget fails with MemoryError instead of blocking the queue.
"""

from __future__ import annotations

import time

import numpy as np
import pytest

from tests.oracle_glue import run_c, run_py
from zorch_amd import _lib
from zorch_amd import workloads as W
from zorch_amd.runtime import Context, DeviceTables
from zorch_amd.serve import ProofServer
from zorch_amd.sha256_field_transcript import Sha256FieldTranscript
from zorch_amd.sumcheck._device import prove_stage

pytestmark = pytest.mark.gpu


def _mode(inst):
    return _lib.ZS_MODE_EQ if inst.mode == "zerocheck" else _lib.ZS_MODE_PRODUCT


def _t0(inst):
    return W.new_transcript(inst)


def _submit(srv, inst, tables=None):
    return srv.submit(inst.tables if tables is None else tables, inst.claim, inst.chal,
                      _mode(inst), inst.degree, _t0(inst), inst.z)


def _alone(inst, ctx):
    return prove_stage(inst.tables, inst.claim, inst.chal, _mode(inst), inst.degree, _t0(inst),
                       inst.z, ctx=ctx)


def _same(a, b):
    assert a.msgs.words.tolist() == b.msgs.words.tolist()
    assert a.point.words.tolist() == b.point.words.tolist()
    assert a.value.words.tolist() == b.value.words.tolist()
    assert a.final.words.tolist() == b.final.words.tolist()
    assert a.transcript.stream_state() == b.transcript.stream_state()


def _requests():
    insts = []
    for j, n in enumerate([16, 3, 12, 14, 1, 13, 9, 15, 11, 18]):
        insts.append(W.zerocheck_bb4(n, (1000 + j, 1100 + j), 1200 + j, 1300 + j))
    insts.append(W.product_bb_ext(13, (1401, 1402), 1403))
    insts.append(W.product_bb(12, (1404, 1405, 1406), 1407))
    insts.append(W.product_gl(14, (1411, 1412), 1413))
    b3 = W.zerocheck_bb4(12, (1421, 1422), 1423, 1424)
    b3.transcript_kind = "blake3"
    insts.append(b3)
    # Large requests, proved whole on the server's own streams.
    insts.append(W.zerocheck_bb4(20, (1431, 1432), 1433, 1434))
    insts.append(W.product_gl(20, (1441, 1442), 1443))
    b3 = W.zerocheck_bb4(21, (1451, 1452), 1453, 1454)
    b3.transcript_kind = "blake3"
    insts.append(b3)
    return insts


@pytest.mark.parametrize("step_rounds", [0, 2])
def test_served_results_equal_alone(step_rounds):
    ctx = Context(0)
    insts = _requests()
    ref = [_alone(i, ctx) for i in insts]
    # step_rounds 2: a request's tail rounds spread over several steps.
    ctx.set_tune(step_rounds=step_rounds, tail_pairs=4096)
    with ProofServer(ctx, tune_key=None) as srv:
        ids = []
        for k, inst in enumerate(insts):
            ids.append(_submit(srv, inst))
            if k % 3 == 2:
                time.sleep(0.002)  # later requests join a running batch
        got = [srv.wait(i) for i in ids]
        st = srv.stats()
    for (res, lat), r in zip(got, ref):
        _same(res, r)
        assert lat > 0
    # Anchored to the oracle: the C oracle for the SHA-256 requests (up to
    # n = 21, batched steps and the server's own streams alike), the Python
    # oracle for the small BLAKE3 one (the n = 21 BLAKE3 request is checked
    # against its one-call proof above only: no oracle restates BLAKE3 at
    # that size in reasonable time).
    if step_rounds == 0:
        for (res, _), inst in zip(got, insts):
            if inst.transcript_kind == "sha256":
                ref_o = run_c(inst, threads=16)
            elif inst.n <= 12:
                ref_o, _ = run_py(inst)
            else:
                continue
            assert res.msgs.words.tolist() == ref_o.msgs.tolist(), inst.name
            assert res.point.words.tolist() == ref_o.point.tolist(), inst.name
            assert res.value.words.reshape(-1).tolist() == ref_o.value.tolist(), inst.name
            assert res.transcript.stream_state() == tuple(ref_o.state), inst.name
    assert st["completed"] == len(insts) and st["failed"] == 0
    assert st["steps"] >= 1 and st["resident_bytes"] == 0
    ctx.close()


def test_budget_admission_and_oversized_request():
    ctx = Context(0)
    insts = [W.zerocheck_bb4(14, (1500 + j, 1600 + j), 1700 + j, 1800 + j) for j in range(6)]
    dev = [DeviceTables.from_host(i.tables, ctx) for i in insts]
    ref = [_alone(i, ctx) for i in insts]
    # Room for two n=14 requests at a time (device-resident inputs: about
    # 280 KB each, mostly the folded tables).
    budget = 600_000
    with ProofServer(ctx, budget, tune_key=None) as srv:
        ids = [_submit(srv, i, d) for i, d in zip(insts, dev)]
        big = W.zerocheck_bb4(17, (1901, 1902), 1903, 1904)
        big_id = _submit(srv, big)
        got = [srv.wait(i) for i in ids]
        with pytest.raises(MemoryError):
            srv.wait(big_id)
        st = srv.stats()
    for (res, _), r in zip(got, ref):
        _same(res, r)
    assert st["peak_resident_bytes"] <= budget
    assert st["completed"] == len(insts) and st["failed"] == 1
    for d in dev:
        d.free()
    ctx.close()


def test_poll_and_completed_ids():
    ctx = Context(0)
    insts = _requests()[:5]
    with ProofServer(ctx, tune_key=None) as srv:
        ids = [_submit(srv, i) for i in insts]
        seen = set()
        deadline = time.time() + 60
        while len(seen) < len(ids) and time.time() < deadline:
            for rid in srv.completed():
                assert srv.poll(rid) is not None
                seen.add(rid)
            time.sleep(0.001)
        assert seen == set(ids)
        with pytest.raises(ValueError):
            srv.poll(ids[0])  # already taken
    ctx.close()


def _digest(msgs, point, value, state) -> bytes:
    import hashlib
    h = hashlib.sha256()
    for a in (msgs, point, value):
        h.update(np.ascontiguousarray(a).tobytes())
    h.update(repr(state).encode())
    return h.digest()


def test_poisson_trace_at_full_size():
    """The served configuration as stated: Poisson arrivals at 2000 requests/s
    over 60 s (seed 5), degree-3 zerocheck requests over BabyBear^4 with n
    uniform in {14..24}, a 200 GB HBM budget, requests submitted open loop on
    the trace's clock. Each request uses one of two distinct device-resident
    inputs of its size and a transcript label of its own. Every served proof
    equals the same request proved alone, and a seeded sample (its n = 24
    members included) equals the C oracle."""
    import ctypes

    from zorch_amd import _lib
    from zorch_amd.sumcheck._device import make_job

    spec = W.serve_spec()
    pool = W.serve_inputs(spec, per_size=2)
    ctx = Context(0)
    dev = {n: [DeviceTables.from_host(i.tables, ctx) for i in insts] for n, insts in pool.items()}
    N = len(spec.arrivals)

    def inst_of(i):
        n = spec.sizes[i]
        return pool[n][i % 2], dev[n][i % 2]

    def t0(i):
        inst, _ = inst_of(i)
        return Sha256FieldTranscript.new(W.serve_label(i), inst.transcript_field)

    served = {}
    with ProofServer(hbm_budget_bytes=200 * 10**9) as srv:
        ids = {}
        i = done = 0
        start = time.perf_counter()
        while done < N:
            now = time.perf_counter() - start
            busy = False
            while i < N and spec.arrivals[i] <= now:
                inst, d = inst_of(i)
                ids[srv.submit(d, inst.claim, inst.chal, _lib.ZS_MODE_EQ, inst.degree, t0(i),
                               inst.z)] = i
                i += 1
                busy = True
            for rid in srv.completed():
                res, _ = srv.poll(rid)
                k = ids.pop(rid)
                served[k] = _digest(res.msgs.words, res.point.words, res.value.words,
                                    res.transcript.stream_state())
                done += 1
                busy = True
            if not busy:
                time.sleep(0.0002)
        st = srv.stats()
    assert st["completed"] == N and st["failed"] == 0 and st["resident_bytes"] == 0
    # Alone: the same requests one per call.
    lib = _lib.lib()
    jobs = {(n, j): make_job(dev[n][j], pool[n][j].claim, pool[n][j].chal, _lib.ZS_MODE_EQ,
                             pool[n][j].degree, pool[n][j].z)
            for n in pool for j in range(2)}
    for k in range(N):
        inst, _ = inst_of(k)
        job = jobs[(inst.n, k % 2)]
        tr = t0(k)
        stc = tr.state.copy()
        _lib.check(lib.zs_prove(ctx.handle, ctypes.byref(job.args), ctypes.byref(stc),
                                ctypes.byref(job.proof)))
        alone = type(tr)(stc, tr.dtype)
        assert served[k] == _digest(job.msgs, job.point, job.value, alone.stream_state()), k
    # The C oracle on a seeded sample, n = 24 requests among it.
    rng = np.random.default_rng(5)
    big = [k for k in range(N) if spec.sizes[k] == 24]
    sample = set(rng.choice(big, size=2, replace=False).tolist())
    sample |= set(rng.choice(N, size=4, replace=False).tolist())
    for k in sorted(sample):
        inst, _ = inst_of(k)
        inst = W.Instance(inst.name, inst.mode, inst.n, inst.tables, inst.claim, inst.chal,
                          inst.transcript_field, W.serve_label(k), inst.degree, inst.z)
        ref = run_c(inst, threads=16)
        assert served[k] == _digest(ref.msgs, ref.point, ref.value, tuple(ref.state)), k
    for ds in dev.values():
        for d in ds:
            d.free()
    ctx.close()
