"""zs_prove_noalloc (the entry point that allocates nothing) against zs_prove.

Two calls in a row on one device, the second with a smaller input, each
compared bit for bit with the allocating entry point on the same arguments;
a stage too large for the scratch array is rejected with ValueError.
"""

from __future__ import annotations

import ctypes

import pytest

from zorch_amd import _lib
from zorch_amd import workloads as W
from zorch_amd.runtime import DeviceTables, default_context
from zorch_amd.sumcheck._device import _result, make_job, prove_stage

pytestmark = pytest.mark.gpu


def _mode(inst):
    return _lib.ZS_MODE_EQ if inst.mode == "zerocheck" else _lib.ZS_MODE_PRODUCT


def _noalloc(inst, tables, ctx):
    job = make_job(tables, inst.claim, inst.chal, _mode(inst), inst.degree, inst.z)
    t0 = W.new_transcript(inst)
    st = t0.state.copy()
    _lib.check(_lib.lib().zs_prove_noalloc(ctx.handle, ctypes.byref(job.args), ctypes.byref(st),
                                           ctypes.byref(job.proof)))
    return _result(job, st, t0)


def _same(a, b):
    assert a.msgs.words.tolist() == b.msgs.words.tolist()
    assert a.point.words.tolist() == b.point.words.tolist()
    assert a.value.words.tolist() == b.value.words.tolist()
    assert a.final.words.tolist() == b.final.words.tolist()
    assert a.transcript.stream_state() == b.transcript.stream_state()


@pytest.mark.parametrize("pair", [
    ("zerocheck", 18, 12), ("product_gl", 16, 10), ("product_bb", 20, 3)])
def test_noalloc_equals_prove_two_calls(pair):
    kind, n_big, n_small = pair
    ctx = default_context()
    make = {"zerocheck": lambda n: W.zerocheck_bb4(n, (n, n + 1), n + 2, n + 3),
            "product_gl": lambda n: W.product_gl(n, (n, n + 1), n + 2),
            "product_bb": lambda n: W.product_bb(n, (n, n + 1), n + 2)}[kind]
    big, small = make(n_big), make(n_small)
    # First call: device-resident tables; second: smaller, host tables.
    dev = DeviceTables.from_host(big.tables, ctx)
    got_big = _noalloc(big, dev, ctx)
    got_small = _noalloc(small, small.tables, ctx)
    ref_big = prove_stage(dev, big.claim, big.chal, _mode(big), big.degree,
                          W.new_transcript(big), big.z, ctx=ctx)
    ref_small = prove_stage(small.tables, small.claim, small.chal, _mode(small), small.degree,
                            W.new_transcript(small), small.z, ctx=ctx)
    _same(got_big, ref_big)
    _same(got_small, ref_small)
    dev.free()


def test_noalloc_rejects_a_stage_beyond_the_scratch_array():
    ctx = default_context()
    inst = W.zerocheck_bb4(22, (5, 6), 7, 8)
    dev = DeviceTables.from_host(inst.tables, ctx)
    with pytest.raises(ValueError, match="bytes of device memory"):
        _noalloc(inst, dev, ctx)
    dev.free()
