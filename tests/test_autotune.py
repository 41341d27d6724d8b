"""Host logic of the launch-shape autotuner (no GPU: a stand-in context)."""

from __future__ import annotations

import pytest

from zorch_amd import autotune


class FakeCtx:
    def __init__(self):
        self.tune = {"max_grid": 2048, "pairs_per_thread": 1, "tail_pairs": 2048}
        self.log = []

    def get_tune(self):
        return dict(self.tune)

    def set_tune(self, **kw):
        self.tune.update(kw)
        self.log.append(dict(self.tune))


def test_tune_picks_fastest_and_caches(tmp_path, monkeypatch):
    monkeypatch.setenv("ZS_TUNE_CACHE", str(tmp_path / "tune.json"))
    ctx = FakeCtx()
    cands = [dict(max_grid=g, pairs_per_thread=1, tail_pairs=1024) for g in (256, 512, 1024)]
    cost = {256: 3e-3, 512: 1e-3, 1024: 2e-3}

    def run():
        import time
        time.sleep(cost[ctx.tune["max_grid"]])

    best, ms, table = autotune.tune(ctx, run, key="w", candidates=cands, reps=2,
                                    check=lambda: 42)
    assert best["max_grid"] == 512 and len(table) == 3 and ms >= 1.0
    assert ctx.tune["max_grid"] == 512
    ctx2 = FakeCtx()
    assert autotune.apply(ctx2, "w") == {"max_grid": 512, "pairs_per_thread": 1,
                                         "tail_pairs": 1024}
    assert autotune.apply(ctx2, "other") is None


def test_tune_rejects_shape_dependent_results(tmp_path, monkeypatch):
    monkeypatch.setenv("ZS_TUNE_CACHE", str(tmp_path / "tune.json"))
    ctx = FakeCtx()
    cands = [dict(max_grid=g, pairs_per_thread=1, tail_pairs=1024) for g in (256, 512)]
    with pytest.raises(AssertionError):
        autotune.tune(ctx, lambda: None, candidates=cands, reps=1,
                      check=lambda: ctx.tune["max_grid"])
    assert ctx.tune["max_grid"] == 2048  # restored
