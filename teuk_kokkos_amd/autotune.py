"""Tile-shape autotuner for a Plan.

The stage kernels' radial tile height trades LDS footprint (occupancy)
against halo re-reads and redundant psi_t rows.  The tuner times a few RK4
steps for each (LDS budget, tile rows) candidate with the plan's own HIP-event
kernel timing and keeps the fastest.  Results are cached per workload
signature in a JSON file so later runs start tuned.
"""
from __future__ import annotations

import hashlib
import json
import os

DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_cache.json")

# 160 KiB of LDS per CU: three, two and one workgroups per CU.
BUDGETS = (53 * 1024, 64 * 1024, 80 * 1024, 160 * 1024)
ROWS = (0, 32, 64, 128)


# Bumped whenever the stage kernels' tiling changes, invalidating cached picks.
TUNING_VERSION = 2


def signature(config_texts) -> str:
    h = hashlib.sha256(f"v{TUNING_VERSION}".encode())
    for t in config_texts:
        h.update(t.encode())
    return h.hexdigest()[:16]


def _time(plan, steps: int) -> float:
    """Kernel milliseconds per RK stage (all execution groups)."""
    plan.set_kernel_timing(True)
    plan.advance(steps)
    ms, n = plan.step_timing()
    plan.set_kernel_timing(False)
    return ms / (4 * max(n, 1))


def tune(plan, config_texts, steps: int = 3, cache_path: str = DEFAULT_CACHE, log=None):
    """Pick (lds_budget, tile_rows) for `plan` and apply it.  The state is
    advanced by the timing steps; re-initialise afterwards."""
    key = signature(config_texts)
    cache = {}
    if cache_path and os.path.exists(cache_path):
        try:
            cache = json.load(open(cache_path))
        except (OSError, ValueError):
            cache = {}
    if key in cache:
        best = cache[key]
        plan.set_lds_budget(best["lds_budget"])
        plan.set_tile_rows(best["tile_rows"])
        return best
    results = []
    for budget in BUDGETS:
        for rows in ROWS:
            try:
                plan.set_lds_budget(budget)
                plan.set_tile_rows(rows)
            except Exception:
                continue
            _time(plan, 1)
            ms = _time(plan, steps)
            tiles, lds = plan.tiles()
            results.append({"lds_budget": budget, "tile_rows": rows, "ms_per_stage": ms,
                            "tiles": tiles, "max_lds": lds})
            if log:
                log(f"[autotune] budget {budget} rows {rows}: {ms:.3f} ms/launch, "
                    f"{tiles} tiles, lds {lds}")
    best = min(results, key=lambda r: r["ms_per_stage"])
    plan.set_lds_budget(best["lds_budget"])
    plan.set_tile_rows(best["tile_rows"])
    if cache_path:
        cache[key] = best
        try:
            with open(cache_path, "w") as f:
                json.dump(cache, f, indent=1)
        except OSError:
            pass
    return best
