"""Shared test helpers: one parameter dict drives both the HIP plan (through
its configuration text) and the CPU oracle (through its plain setup)."""
from __future__ import annotations

import numpy as np

import oracle
from teuk_kokkos_amd import workloads


def oracle_pipeline(params: dict, threads: int = 0) -> oracle.Pipeline:
    p = dict(params)
    st = oracle.Setup(
        mass=p.get("mass", 1.0), spin=p.get("spin", 0.0),
        length=p.get("compactification_length", 1.0), nr=p["nr"], ntheta=p["ntheta"],
        ell_max_first=p["ellmax_first"], ell_max_second=p.get("ellmax_second", p["ellmax_first"]),
        first_order_modes=list(p.get("first_order_modes", [-2, 2])),
        second_order_modes=list(p.get("second_order_modes", [-2, 2])),
        reduction_damping=p.get("reduction_damping", 0.1),
        dissipation=p.get("dissipation", 0.01), reduction=p.get("reduction_mode", "free_damped"),
        scheme=p.get("radial_discretization", "d4-2"),
        source_mode=(p.get("second_order__source_mode", "constraint_aware")
                     if p.get("second_order__enabled", False) else "disabled"),
        source_start_time=p.get("second_order__source_start_time", 0.0),
        constraint_tolerance=p.get("second_order__constraint_tolerance", 1.0e-10),
        required_passes=p.get("second_order__required_consecutive_passes", 1))
    return oracle.Pipeline(st, threads=threads)


def expanded_seeds(params: dict):
    """initial_data_factory.hpp:45-74: sharp partners then (m, ell) order."""
    seeds = []
    for ell, m, amp in params["seeds"]:
        seeds.append((ell, m, complex(amp)))
        if params.get("initial_data__add_sharp_partner", False) and m != 0:
            seeds.append((ell, -m, complex(amp).conjugate()))
    return sorted(seeds, key=lambda s: (s[1], s[0]))


def oracle_initialized(params: dict, threads: int = 0) -> oracle.Pipeline:
    pipe = oracle_pipeline(params, threads)
    h = pipe.setup.horizon
    pipe.init_gaussian(params.get("initial_data__center", 0.45) * h,
                       params.get("initial_data__width", 0.12) * h, expanded_seeds(params),
                       compact_support=params.get("initial_data__compact_support", False))
    return pipe


def small_case(**over) -> dict:
    """A small linear case: both codes finish it in well under a second."""
    base = dict(mass=1.0, spin=0.0, compactification_length=2.0, nr=41, ntheta=7,
                ellmax_first=4, ellmax_second=4, first_order_modes=[-2, 2],
                second_order_modes=[-2, 2], final_time=1.0e-3, steps=1, cfl=0.5,
                reduction_damping=0.1, dissipation=0.0, radial_discretization="d4-2",
                reduction_mode="free_damped", initial_data__center=0.5,
                initial_data__width=0.1, seeds=[(2, 2, 1.0e-3)])
    base.update(over)
    if "ntheta" not in over:
        base["ntheta"] = workloads.minimum_ntheta(base["ellmax_first"], base["ellmax_second"])
    return base


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    """max |a - b| / max |b| (state-wide relative error)."""
    scale = np.abs(b).max()
    return float(np.abs(a - b).max() / (scale if scale > 0 else 1.0))


def horizon_weights(d: int, h: float) -> np.ndarray:
    """Backward 9-point stencil for the d-th radial derivative at the last
    grid point: the moment equations sum_j w_j (-j)^k = k! delta_kd,
    k = 0..8 (horizon_diagnostics.hpp:18-79), solved in float64."""
    import math
    A = np.array([[float(-j) ** k for j in range(9)] for k in range(9)])
    b = np.zeros(9)
    b[d] = math.factorial(d)
    return np.linalg.solve(A, b) * h ** (-d)
