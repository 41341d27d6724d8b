"""Launch-shape autotuner.

The prover's launch shapes (``Context.set_tune``: the largest grid of a
multi-block pass, the target pairs per thread, the round size below which
a proof's remaining rounds run in one block, and in batched and served proving
how many of those rounds one step runs) never change results, only speed. ``tune`` times a workload under each candidate shape and keeps the
fastest; results are cached per workload key in a JSON file (``ZS_TUNE_CACHE``
or ``~/.cache/zorch_amd/tune.json``) and re-applied with ``apply``.
"""

from __future__ import annotations

import itertools
import json
import os
import statistics
import time
from pathlib import Path
from typing import Callable, Iterable

from zorch_amd.runtime import Context

ARCH = "gfx950"

CANDIDATES = [dict(max_grid=g, pairs_per_thread=p, tail_pairs=t, step_rounds=s)
              for g, p, t, s in itertools.product((256, 512, 1024, 2048, 4096), (1, 2, 4),
                                                  (512, 1024, 2048, 4096), (0, 4))]


def cache_path() -> Path:
    env = os.environ.get("ZS_TUNE_CACHE")
    if env:
        return Path(env)
    return Path.home() / ".cache" / "zorch_amd" / "tune.json"


def _load_all() -> dict:
    p = cache_path()
    try:
        return json.loads(p.read_text())
    except (OSError, ValueError):
        return {}


def load(key: str) -> dict | None:
    return _load_all().get(f"{ARCH}:{key}")


def save(key: str, params: dict, ms: float) -> None:
    p = cache_path()
    data = _load_all()
    data[f"{ARCH}:{key}"] = dict(params, ms=ms)
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        p.write_text(json.dumps(data, indent=1, sort_keys=True) + "\n")
    except OSError:
        pass  # a read-only home: the result is still returned


def apply(ctx: Context, key: str) -> dict | None:
    """Set the cached shape for `key` on `ctx`; returns it (None if untuned)."""
    got = load(key)
    if got is None:
        return None
    params = {k: got[k] for k in ("max_grid", "pairs_per_thread", "tail_pairs", "step_rounds")
              if k in got}
    ctx.set_tune(**params)
    return params


# Cache keys of the workload classes the library applies by itself: batched
# proving (``prove_stage_batch``) and served proving (``ProofServer``).
BATCH_KEY = "batch"
SERVE_KEY = "serve"


class applied:
    """Context manager: the cached shape of `key` on `ctx` for the duration
    (the context's own shape when none is cached), restored afterwards."""

    def __init__(self, ctx: Context, key: str):
        self.ctx, self.key, self.saved = ctx, key, None

    def __enter__(self) -> dict | None:
        self.saved = self.ctx.get_tune()
        return apply(self.ctx, self.key)

    def __exit__(self, *exc) -> None:
        self.ctx.set_tune(**self.saved)


def tune(ctx: Context, run: Callable[[], None], key: str | None = None,
         candidates: Iterable[dict] = CANDIDATES, reps: int = 3,
         check: Callable[[], object] | None = None) -> tuple[dict, float, list]:
    """Time `run` under every candidate shape (one warm-up, median of `reps`)
    and leave the fastest set on `ctx`. `check`, if given, is called after each
    candidate's runs and must return the same value for every candidate (the
    results do not depend on launch shapes). Returns (best, best_ms, table)."""
    base = ctx.get_tune()
    table = []
    ref = None
    best, best_ms = None, float("inf")
    try:
        for cand in candidates:
            ctx.set_tune(**cand)
            run()
            ts = []
            for _ in range(reps):
                t = time.perf_counter()
                run()
                ts.append((time.perf_counter() - t) * 1e3)
            ms = statistics.median(ts)
            if check is not None:
                got = check()
                if ref is None:
                    ref = got
                elif got != ref:
                    raise AssertionError(f"results changed under launch shape {cand}")
            table.append(dict(cand, ms=ms))
            if ms < best_ms:
                best, best_ms = dict(cand), ms
    finally:
        ctx.set_tune(**base)
    ctx.set_tune(**best)
    if key is not None:
        save(key, best, best_ms)
    return best, best_ms, table
