"""The teuk_solver driver (src/main.cpp:26-55, solver_driver.hpp:86-353) on
the HIP path: output files and their schemas, and diagnostics lines equal to
the same run driven through the Python binding.  Runs the built executable
as a child process on the GPU."""
from __future__ import annotations

import csv
import os
import subprocess

import numpy as np
import pytest

import helpers as H
import teuk_kokkos_amd as T
from teuk_kokkos_amd import workloads as W

pytestmark = pytest.mark.gpu

SOLVER = os.path.join(os.path.dirname(T.__file__), "teuk_solver")

DIAG_HEADER = (
    "step,time,kappa_time,first_psi_rms,second_psi_rms,first_constraint_rms,"
    "second_constraint_rms,source_rms,forcing_rms,source_active,source_activation_time,"
    "source_constraint_abs_max,source_constraint_normalized_max,"
    "source_constraint_normalized_l2,transport_residual_rms,psi3_bianchi_rms,"
    "psi2_bianchi_rms,hll_reality_rms,first_horizon_d0,first_horizon_d1,first_horizon_d2,"
    "first_horizon_d3,first_horizon_d4,second_horizon_d0,second_horizon_d1,"
    "second_horizon_d2,second_horizon_d3,second_horizon_d4")

CASES = {
    "linear": H.small_case(spin=0.7, nr=49, radial_discretization="d10-5", ellmax_first=6,
                           ellmax_second=6, steps=6, final_time=6e-3),
    "second": H.small_case(spin=0.6, compactification_length=1.5, nr=41,
                           radial_discretization="d10-5", ellmax_first=4, ellmax_second=6,
                           first_order_modes=[-2, 2], second_order_modes=[-4, 0, 4],
                           initial_data__add_sharp_partner=True, seeds=[(2, 2, 1e-3)],
                           second_order__enabled=True,
                           second_order__source_mode="unrestricted", steps=6, final_time=6e-3),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_solver_writes_reference_outputs(name, tmp_path):
    params = CASES[name]
    cfg = tmp_path / "run.cfg"
    cfg.write_text(W.config_text(**dict(params)))
    out = tmp_path / "out"
    print(f"[gpu] teuk_solver {name}", flush=True)
    res = subprocess.run([SOLVER, "--config", str(cfg), f"output.directory={out}",
                          "output.diagnostic_every=3"], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stderr
    lines = (out / "diagnostics.csv").read_text().splitlines()
    assert lines[0] == DIAG_HEADER
    rows = list(csv.DictReader(lines))
    assert [int(r["step"]) for r in rows] == [0, 3, 6]
    assert res.stdout.splitlines()[0] == DIAG_HEADER
    assert f"source_policy={'unrestricted' if name == 'second' else 'disabled'}" in res.stdout
    # the same run through the binding gives the same last line
    dt = params["final_time"] / params["steps"]
    with T.Plan(W.config_text(**dict(params))) as plan:
        plan.initialize()
        plan.advance(6)
        t, steps = plan.progress()
        d = plan.sample_diagnostics(t)
        pairs = plan.source_pairs()
    last = rows[-1]
    assert abs(float(last["time"]) - 6 * dt) <= 1e-15
    assert float(last["first_psi_rms"]) == d["state_rms"][2]
    assert float(last["second_psi_rms"]) == d["state_rms"][12]
    assert float(last["transport_residual_rms"]) == d["transport_combined"][0]
    assert float(last["psi3_bianchi_rms"]) == d["psi3_bianchi"][0]
    assert [float(last[f"first_horizon_d{k}"]) for k in range(5)] == list(d["first_horizon"])
    assert int(last["source_active"]) == int(d["source_active"])
    prows = list(csv.DictReader((out / "source_pairs.csv").read_text().splitlines()))
    if name == "second":
        assert len(prows) == 3 * len(pairs)
        tail = prows[-len(pairs):]
        for p, r in zip(pairs, tail):
            assert (int(r["m1"]), int(r["m2"]), int(r["target"])) == (p["m1"], p["m2"], p["target"])
            assert float(r["D_rms"]) == p["D_rms"] and float(r["T_max"]) == p["T_max"]
        assert np.isfinite(float(last["forcing_rms"])) and float(last["forcing_rms"]) > 0
    else:
        assert prows == []
    wf = (out / "waveforms.csv").read_text().splitlines()
    assert wf[0] == "step,time,boundary,order,m,ell,rescaled_re,rescaled_im,endpoint_re,endpoint_im"
