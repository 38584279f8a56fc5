"""Integrated evolution on the GPU: the reference's production QNM runs.

The ringdown runs of tests/test_qnm.cpp:142-215 (and their documented
results, docs/PRODUCTION_QNM_VALIDATION.md:58-82) are advanced through the C
ABI for thousands of RK4 steps.  Two checks per run:
  * the extracted (l = 2) signal agrees with the CPU oracle's along the whole
    run (the FP64 parity bar, now over 5000-8000 steps instead of ten), and
  * the fitted quasinormal frequency reproduces the reference's documented
    production values (Schwarzschild convergence sequence, Kerr a = 0.7 vs
    Leaver)."""
from __future__ import annotations

import math

import numpy as np
import pytest

import helpers as H
import oracle as O
import teuk_kokkos_amd as T
from teuk_kokkos_amd import workloads as W
from test_oracle_known_answers import GOLD, _fit_complex, _fit_real, qnm_run

pytestmark = pytest.mark.gpu

DT = 0.01


def gpu_qnm_run(nr, ell_max, ntheta, spin, t_end):
    """qnm_run of the oracle tests, stepped on the GPU."""
    params = H.small_case(spin=spin, nr=nr, ellmax_first=ell_max, ellmax_second=ell_max,
                          ntheta=ntheta, dissipation=0.005, radial_discretization="d4-2",
                          initial_data__center=0.675, initial_data__width=0.08,
                          seeds=[(2, 2, 1.0)])
    params["final_time"] = DT
    params["steps"] = 1
    print(f"[gpu] qnm run nr={nr} ellmax={ell_max} a={spin} t_end={t_end}", flush=True)
    _, A = O.transform_matrices(-2, 2, ell_max, ntheta)
    modes = sorted(set(params["first_order_modes"]) | set(params["second_order_modes"]))
    mi = modes.index(2)
    with T.Plan(W.config_text(**params)) as plan:
        plan.initialize()
        steps = int(round(t_end / DT))
        times, sig = [], []
        t = 0.0
        for s in range(0, steps + 1, 5):
            state = plan.export_state()
            times.append(s * DT)
            sig.append(state[mi, 2])
            if s != steps:
                for _ in range(5):
                    plan.step(t, DT)
                    t += DT
    return np.array(times), np.array(sig), A


def _extract(sig_full, A, obs):
    return np.array([(A @ row[obs])[0] for row in sig_full])


def _observation_index(nr, spin):
    """R = 1 on the compactified grid R_H = L^2 / r_+ (L = 2, M = 1)."""
    rp = 1.0 + math.sqrt(max(1.0 - spin * spin, 0.0))
    horizon = 4.0 / rp
    return int(round((1.0 / horizon) * (nr - 1)))


def test_schwarzschild_qnm_production_sequence_on_gpu():
    g = GOLD["schwarzschild_qnm"]
    seq = g["production_sequence"]
    for nr, want in zip(seq["nr"], seq["complex_frequency_error"]):
        t, full, A = gpu_qnm_run(nr, 3, 5, 0.0, 65.0)
        s = _extract(full, A, _observation_index(nr, 0.0))
        t_ref, s_ref = qnm_run(nr, 3, 5, 0.0, 65.0)
        assert np.array_equal(t, t_ref)
        rel = np.abs(s - s_ref).max() / np.abs(s_ref).max()
        print(f"[gpu] nr={nr} signal rel err vs oracle over {len(t)} samples: {rel:.3e}",
              flush=True)
        assert rel <= 1e-12
        f, d = _fit_real(t, s.real, 0.05, 45.0, 65.0)
        e = math.hypot(f - g["omega"][0], d - g["omega"][1])
        assert abs(e - want) < 0.006 * want


def test_kerr_qnm_on_gpu_matches_leaver():
    g = GOLD["kerr_qnm_a07"]
    pr = g["production"]
    t, full, A = gpu_qnm_run(pr["nr"], pr["ell_max"], pr["ntheta"], pr["spin"],
                             pr["fit_window"][1])
    s = _extract(full, A, _observation_index(pr["nr"], pr["spin"]))
    t_ref, s_ref = qnm_run(pr["nr"], pr["ell_max"], pr["ntheta"], pr["spin"],
                           pr["fit_window"][1])
    rel = np.abs(s - s_ref).max() / np.abs(s_ref).max()
    print(f"[gpu] Kerr a=0.7 signal rel err vs oracle over {len(t)} samples: {rel:.3e}",
          flush=True)
    assert rel <= 1e-12
    f, d = _fit_complex(t, s, 0.05, *pr["fit_window"])
    err = math.hypot(f - g["leaver_omega"][0], d - g["leaver_omega"][1])
    assert err < pr["max_error"]
    assert abs(f - pr["omega"][0]) < 2e-6 and abs(d - pr["omega"][1]) < 2e-6
