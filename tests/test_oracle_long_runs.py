"""The oracle's long-run fixtures and the CFL of the Schwarzschild D8-4 run
(CPU, no GPU).

The D8-4 boundary closure has a large boundary eigenvalue, so classical RK4
on the nr = 257 Schwarzschild run is unstable at the CFL 0.5 of the standard
case description (and at 0.05).  The run uses CFL 0.02: the oracle stays
bounded for the full 2000 steps there (the committed fixture), while CFL
0.025, the next value up, diverges within 300 steps."""
from __future__ import annotations

import os

import numpy as np

import helpers as H
from teuk_kokkos_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))


def test_schwarzschild_d84_diverges_just_above_the_chosen_cfl():
    params = W.case1(cfl=0.025)
    dt = params["final_time"] / params["steps"]
    pipe = H.oracle_initialized(params)
    pipe.advance(300, dt)
    s = pipe.get_state()
    assert not np.isfinite(s).all() or np.abs(s).max() > 1.0


def test_long_run_fixtures_are_bounded_and_complete():
    fix = np.load(os.path.join(HERE, "golden", "long_runs.npz"))
    for name, checkpoints in (("case1", [2000]), ("case2", [20000]),
                              ("case4", [100, 1000, 10000])):
        dt, every, stride, total = fix[f"{name}_meta"]
        params = getattr(W, name)()
        assert dt == params["final_time"] / params["steps"]
        assert fix[f"{name}_scri"].shape[0] == int(total) // int(every) + 1
        for n in checkpoints:
            s = fix[f"{name}_state_{n}"]
            assert np.isfinite(s).all() and 0 < np.abs(s).max() < 0.1, (name, n)
