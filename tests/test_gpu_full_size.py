"""HIP path vs the CPU oracle at the full sizes of the standard workloads.

The small cases of test_gpu_parity.py cover every code path; these run the
production shapes themselves: the Schwarzschild D8-4 run (nr 257), the Kerr
a = 0.7 ringdown (nr 1025, ell <= 16), the near-extremal run (nr 4097,
ell <= 32, the Gram path with the 2-tile MFMA class) and instances of the
512-instance Kerr sweep advanced inside the full batch.  FP64 bar: <= 1e-12
relative on the state.  Batch composition must not change an instance's
result at all (bitwise), since every output value is computed from the same
inputs in the same order whatever the tiling."""
from __future__ import annotations

import numpy as np
import pytest

import helpers as H
import teuk_kokkos_amd as T
from teuk_kokkos_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1.0e-12
STEPS = 3


def _oracle_after(params, steps):
    ref = H.oracle_initialized(params)
    ref.advance(steps, params["final_time"] / params["steps"])
    return ref.get_state()


@pytest.mark.parametrize("name", ["case1", "case2", "case4"])
def test_full_size_case_matches_oracle(name):
    params = getattr(W, name)()
    print(f"[gpu] {name}: nr={params['nr']} ellmax={params['ellmax_first']} "
          f"a={params['spin']}", flush=True)
    with T.Plan(W.config_text(**params)) as plan:
        plan.initialize()
        init = plan.export_state()
        plan.advance(STEPS)
        got = plan.export_state()
    ref0 = H.oracle_initialized(params).get_state()
    assert H.rel_err(init, ref0) <= TOL
    err = H.rel_err(got, _oracle_after(params, STEPS))
    print(f"[gpu] {name} {STEPS} steps rel err {err:.3e}", flush=True)
    assert err <= TOL


def _sweep_probe_indices(cases):
    """One instance per ellmax class plus an m = 0 instance."""
    picks = {}
    for k, c in enumerate(cases):
        key = c["ellmax_first"]
        if key not in picks:
            picks[key] = k
    zero = next(k for k, c in enumerate(cases) if c["first_order_modes"] == [0])
    return sorted(set(picks.values()) | {zero})


def test_sweep_instances_in_full_batch_match_oracle_and_solo_runs():
    cases = W.sweep(512)
    probes = _sweep_probe_indices(cases)
    print(f"[gpu] sweep512 probes {probes}", flush=True)
    with T.Plan(W.texts(cases)) as plan:
        plan.initialize()
        plan.advance(STEPS)
        batch = {k: plan.export_state(k) for k in probes}
    for k in probes:
        with T.Plan(W.config_text(**cases[k])) as solo:
            solo.initialize()
            solo.advance(STEPS)
            alone = solo.export_state()
        assert np.array_equal(batch[k], alone), f"instance {k} depends on its batch"
        err = H.rel_err(batch[k], _oracle_after(cases[k], STEPS))
        print(f"[gpu] sweep instance {k} (ellmax {cases[k]['ellmax_first']}, "
              f"modes {cases[k]['first_order_modes']}) rel err {err:.3e}", flush=True)
        assert err <= TOL
