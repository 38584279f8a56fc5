"""Tile shapes change speed, never results: plans with tuned per-group shapes
(LDS budget, tile rows) step to the same bits as the default plan, eagerly
and as graph replays, and the autotuner run on the GPU leaves a plan that
does too."""
from __future__ import annotations

import numpy as np
import pytest

import helpers as H
import teuk_kokkos_amd as T
from teuk_kokkos_amd import autotune as A
from teuk_kokkos_amd import workloads as W

pytestmark = pytest.mark.gpu

# three execution groups: a Kerr ell-16 line (G = I), a Kerr ell-24 line
# with the Gram projection and a dissipative line (the general kernel)
CASES = [H.small_case(spin=0.7, nr=97, radial_discretization="d10-5", ellmax_first=16,
                      ellmax_second=16, seeds=[(2, 2, 1e-3)]),
         H.small_case(spin=0.5, nr=81, radial_discretization="d10-5", ellmax_first=24,
                      ellmax_second=24, seeds=[(3, 2, 2e-3j)]),
         H.small_case(spin=0.3, nr=65, radial_discretization="d10-5", ellmax_first=8,
                      ellmax_second=8, dissipation=0.05, steps=20, seeds=[(2, 2, 1e-3)])]
STEPS = 6


def _run(shapes=None, graph=False, tune_path=None):
    with T.Plan(W.texts(CASES)) as plan:
        if shapes:
            for key, (lds, rows) in shapes.items():
                plan.set_group_shape(*key, lds, rows)
        pick = None
        if tune_path is not None:
            plan.initialize()
            pick = A.tune(plan, W.texts(CASES), steps=2, cache_path=tune_path,
                          use_cache=False)
        plan.set_graph(graph)
        plan.initialize()
        plan.advance(STEPS)
        groups = plan.groups()
        return [plan.export_state(i) for i in range(len(CASES))], groups, pick


def _same(a, b):
    return all(np.array_equal(x.view(np.float64), y.view(np.float64)) for x, y in zip(a, b))


@pytest.mark.parametrize("graph", [False, True])
def test_group_shapes_do_not_change_results(graph):
    ref, groups, _ = _run(graph=graph)
    keys = [(g["scheme"], g["ell_tiles"], g["variant"]) for g in groups]
    print(f"[gpu] groups {keys}", flush=True)
    assert len(keys) == 3
    shapes = {keys[0]: (53 * 1024, 16), keys[1]: (160 * 1024, 32), keys[2]: (64 * 1024, 0)}
    out, tuned_groups, _ = _run(shapes, graph=graph)
    for g in tuned_groups:
        want = shapes[(g["scheme"], g["ell_tiles"], g["variant"])]
        assert (g["lds_budget"], g["tile_rows"]) == want
    assert _same(out, ref)


def test_autotuned_plan_matches_the_default_plan(tmp_path):
    ref, _, _ = _run(graph=True)
    out, groups, pick = _run(graph=True, tune_path=str(tmp_path / "cache.json"))
    print(f"[gpu] autotune pick {pick}", flush=True)
    assert set(pick["shapes"]) == {A.group_key(g) for g in groups}
    assert A.cached(W.texts(CASES), str(tmp_path / "cache.json"))["shapes"] == pick["shapes"]
    assert _same(out, ref)
