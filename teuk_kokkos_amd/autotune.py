"""Per-execution-group tile-shape autotuner for a Plan.

A plan launches one stage kernel per execution group (radial scheme x ell
tile class x line variant) and RK stage, the groups running concurrently on
streams of their own.  Each group's tile shape trades LDS footprint
(workgroups per CU) against halo re-reads and redundant psi_t rows (tile
rows, LDS budget).  The tuner walks the groups, largest first, and for each sweeps
its shapes with the others fixed, keeping the one with the shortest whole
RK4 step (GPU time between the fork and the join of the group streams,
replayed as a hipGraph).  Results do not depend on the shape, so the pick
is free to change between devices; picks are cached per workload signature
in a JSON file that later runs apply without timing anything.
"""
from __future__ import annotations

import hashlib
import json
import os

DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_cache.json")

# LDS budgets per workgroup (160 KiB per CU): 0 is the plan's 80 KiB (two
# workgroups per CU); 53 KiB allows three, 160 KiB one.
BUDGETS = (0, 53 * 1024, 64 * 1024, 160 * 1024)
# Tile rows: 0 picks the largest fitting the budget, capped for parallelism.
ROWS = (0, 32, 64, 96, 128, 192)

DEFAULT_SHAPE = {"lds_budget": 0, "tile_rows": 0}

# Bumped whenever the stage kernels' tiling or the shape space changes,
# invalidating cached picks.
TUNING_VERSION = 3


def signature(config_texts) -> str:
    h = hashlib.sha256(f"v{TUNING_VERSION}".encode())
    for t in config_texts:
        h.update(t.encode())
    return h.hexdigest()[:16]


def group_key(g) -> str:
    """Cache key of an execution group: scheme/ell_tiles/variant."""
    return f"{g['scheme']}/{g['ell_tiles']}/{g['variant']}"


def _parse_key(key: str):
    scheme, ub, variant = (int(x) for x in key.split("/"))
    return scheme, ub, variant


def step_ms(plan, steps: int) -> float:
    """GPU milliseconds per RK4 step over `steps` steps (one untimed first)."""
    plan.advance(1)
    plan.set_kernel_timing(True)
    plan.advance(steps)
    ms, n = plan.step_timing()
    plan.set_kernel_timing(False)
    return ms / max(n, 1)


def load_cache(path: str) -> dict:
    if path and os.path.exists(path):
        try:
            with open(path) as f:
                data = json.load(f)
            return data if isinstance(data, dict) else {}
        except (OSError, ValueError):
            return {}
    return {}


def apply(plan, shapes: dict):
    """Set every group shape of a pick ({group key: shape dict})."""
    for key, shape in shapes.items():
        plan.set_group_shape(*_parse_key(key), shape["lds_budget"], shape["tile_rows"])


def cached(config_texts, cache_path: str = DEFAULT_CACHE):
    """The cached pick for this workload, or None."""
    return load_cache(cache_path).get(signature(config_texts))


def apply_cached(plan, config_texts, cache_path: str = DEFAULT_CACHE):
    """Apply the cached pick if there is one; returns it (or None)."""
    pick = cached(config_texts, cache_path)
    if pick is not None:
        apply(plan, pick["shapes"])
    return pick


def tune(plan, config_texts, steps: int = 4, cache_path: str = DEFAULT_CACHE, log=None,
         use_cache: bool = True, measure=step_ms):
    """Pick a shape per execution group of `plan` and apply it.  The plan's
    state is advanced by the timing steps; re-initialise afterwards.  Returns
    {"shapes": {group key: shape}, "step_ms": ..., "default_step_ms": ...},
    the two step times from the closing comparison."""
    key = signature(config_texts)
    cache = load_cache(cache_path)
    if use_cache and key in cache:
        apply(plan, cache[key]["shapes"])
        return cache[key]
    plan.set_graph(True)
    default_ms = measure(plan, steps)
    groups = plan.groups()
    order = sorted(range(len(groups)), key=lambda i: -groups[i]["tiles"])
    shapes = {}
    best_ms = default_ms
    for gi in order:
        gkey = group_key(groups[gi])
        scheme, ub, variant = _parse_key(gkey)
        best = (best_ms, dict(DEFAULT_SHAPE))
        for b in BUDGETS:
            for r in ROWS:
                cand = {"lds_budget": b, "tile_rows": r}
                if cand == DEFAULT_SHAPE:
                    continue
                try:
                    plan.set_group_shape(scheme, ub, variant, b, r)
                except Exception as e:  # a shape the plan cannot tile
                    if log:
                        log(f"[autotune] group {gkey} {cand}: refused ({e})")
                    continue
                ms = measure(plan, steps)
                if log:
                    log(f"[autotune] group {gkey} {cand}: {ms:.4f} ms/step")
                if ms < best[0]:
                    best = (ms, cand)
        shapes[gkey] = best[1]
        best_ms = best[0]
        plan.set_group_shape(scheme, ub, variant, best[1]["lds_budget"], best[1]["tile_rows"])
    # Confirm against the default shapes over longer runs, alternating, so a
    # lucky short measurement cannot keep a slower pick.
    default_shapes = {k: dict(DEFAULT_SHAPE) for k in shapes}
    totals = {"pick": 0.0, "default": 0.0}
    for _ in range(2):
        for name, sh in (("default", default_shapes), ("pick", shapes)):
            apply(plan, sh)
            totals[name] += measure(plan, 4 * steps)
    if totals["pick"] >= totals["default"]:
        shapes = default_shapes
        apply(plan, shapes)
    pick = {"shapes": shapes, "step_ms": min(totals.values()) / 2,
            "default_step_ms": totals["default"] / 2}
    if log:
        log(f"[autotune] confirmed: pick {totals['pick'] / 2:.4f} ms/step, default "
            f"{totals['default'] / 2:.4f} ms/step")
    if cache_path:
        cache[key] = pick
        try:
            with open(cache_path, "w") as f:
                json.dump(cache, f, indent=1, sort_keys=True)
        except OSError:
            pass
    return pick
