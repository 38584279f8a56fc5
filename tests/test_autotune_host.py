"""The tile-shape autotuner's selection and caching, on the CPU.

A stand-in plan answers groups() and records set_group_shape calls; the
measured step time is a fixed function of the shapes, so the pick the tuner
must make is known.  The GPU side (that tuned and untuned plans step to the
same bits) is tests/test_gpu_tuning.py."""
from __future__ import annotations

import json

import pytest

from teuk_kokkos_amd import autotune as A

DEFAULT = {"lds_budget": 0, "tile_rows": 0}


class FakePlan:
    def __init__(self, keys, refuse=()):
        self.keys = keys
        self.shapes = {k: dict(DEFAULT) for k in keys}
        self.refuse = set(refuse)
        self.graph = False

    def groups(self):
        out = []
        for i, k in enumerate(self.keys):
            scheme, ub, variant = (int(x) for x in k.split("/"))
            out.append({"scheme": scheme, "ell_tiles": ub, "variant": variant,
                        "tiles": 100 * (i + 1), "lds_bytes": 0})
        return out

    def set_group_shape(self, scheme, ub, variant, lds, rows):
        key = f"{scheme}/{ub}/{variant}"
        if (key, lds, rows) in self.refuse:
            raise ValueError("cannot tile")
        self.shapes[key] = {"lds_budget": lds, "tile_rows": rows}

    def set_graph(self, on):
        self.graph = on


def cost_model(best):
    """Step time: 1 ms plus, per group, the distance of its shape from
    best[group] (0 at the optimum)."""
    calls = []

    def measure(plan, steps):
        calls.append(steps)
        t = 1.0
        for k, s in plan.shapes.items():
            b = best.get(k, DEFAULT)
            t += 0.1 * (s["lds_budget"] != b["lds_budget"])
            t += 0.05 * abs(s["tile_rows"] - b["tile_rows"]) / 32
        return t

    return measure, calls


def test_tuner_picks_each_groups_fastest_shape_and_caches_it(tmp_path):
    keys = ["2/1/1", "2/2/2", "2/1/0"]
    best = {"2/1/1": {"lds_budget": 53 * 1024, "tile_rows": 32},
            "2/2/2": {"lds_budget": 160 * 1024, "tile_rows": 96}}
    plan = FakePlan(keys)
    measure, calls = cost_model(best)
    cache = tmp_path / "cache.json"
    pick = A.tune(plan, ["cfg a", "cfg b"], cache_path=str(cache), measure=measure)
    assert plan.graph
    assert pick["shapes"]["2/1/1"] == best["2/1/1"]
    assert pick["shapes"]["2/2/2"] == best["2/2/2"]
    assert pick["shapes"]["2/1/0"] == DEFAULT
    assert plan.shapes == pick["shapes"]
    assert pick["step_ms"] == pytest.approx(1.0)
    assert pick["default_step_ms"] > pick["step_ms"]
    stored = json.loads(cache.read_text())
    assert stored[A.signature(["cfg a", "cfg b"])]["shapes"] == pick["shapes"]

    # a later run applies the cached pick without timing anything
    n_calls = len(calls)
    fresh = FakePlan(keys)
    again = A.tune(fresh, ["cfg a", "cfg b"], cache_path=str(cache), measure=measure)
    assert len(calls) == n_calls
    assert again["shapes"] == pick["shapes"] and fresh.shapes == pick["shapes"]
    other = FakePlan(keys)
    assert A.apply_cached(other, ["cfg a", "cfg b"], str(cache)) is not None
    assert other.shapes == pick["shapes"]
    # another workload (or tuner version) misses the cache
    assert A.cached(["cfg a"], str(cache)) is None


def test_refused_shapes_are_skipped(tmp_path):
    keys = ["0/1/0"]
    best = {"0/1/0": {"lds_budget": 64 * 1024, "tile_rows": 64}}
    plan = FakePlan(keys, refuse={("0/1/0", 64 * 1024, 64)})
    measure, _ = cost_model(best)
    pick = A.tune(plan, ["x"], cache_path=str(tmp_path / "c.json"), measure=measure)
    # the refused optimum is skipped; of its accepted neighbours (rows 32
    # and 96 tie) the first swept wins
    assert pick["shapes"]["0/1/0"] == {"lds_budget": 64 * 1024, "tile_rows": 32}


def test_a_lucky_short_measurement_does_not_survive_the_confirmation(tmp_path):
    """One spuriously fast reading of a worse shape wins the sweep; the
    closing alternating comparison against the defaults undoes it."""
    keys = ["2/1/1"]
    plan = FakePlan(keys)
    state = {"lucky": True}

    def measure(p, steps):
        s = p.shapes["2/1/1"]
        if s == {"lds_budget": 53 * 1024, "tile_rows": 32} and \
                state["lucky"]:
            state["lucky"] = False
            return 0.5
        return 1.0 if s == DEFAULT else 1.2

    pick = A.tune(plan, ["y"], cache_path=str(tmp_path / "c.json"), measure=measure)
    assert pick["shapes"]["2/1/1"] == DEFAULT
    assert plan.shapes["2/1/1"] == DEFAULT


def test_the_committed_cache_is_well_formed():
    data = A.load_cache(A.DEFAULT_CACHE)
    for sig, pick in data.items():
        assert len(sig) == 16
        for key, shape in pick["shapes"].items():
            scheme, ub, variant = (int(x) for x in key.split("/"))
            assert scheme in (0, 1, 2) and ub in (1, 2, 4) and variant in (0, 1, 2)
            assert shape["lds_budget"] in A.BUDGETS
            assert shape["tile_rows"] in A.ROWS
        assert pick["step_ms"] <= pick["default_step_ms"]
