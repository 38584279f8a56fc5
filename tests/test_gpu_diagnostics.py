"""Sampled diagnostics on the GPU (solver_driver.hpp:205-242) against numpy
reductions of the oracle's state, RHS and RHS intermediates: RMS and maximum
of every field and of its RHS, the reduction constraints Q - D Psi with the
oracle's SBP matrix, source_over_r3 and the latched forcing, the independent
reconstruction constraints, the transport-equation residuals
(reconstruction_diagnostics.hpp:46-140) and the horizon derivative maxima
(horizon_diagnostics.hpp:18-160).  The reference reduces with atomics, so
only rounding-level agreement is defined; the bars are written per
quantity.  All calls go through the C ABI."""
from __future__ import annotations

import math

import numpy as np
import pytest

import helpers as H
import oracle as O
import teuk_kokkos_amd as T
from teuk_kokkos_amd import workloads as W

pytestmark = pytest.mark.gpu

SCHEME = {"d4-2": 0, "d8-4": 1, "d10-5": 2}

CASES = {
    "linear_kerr": H.small_case(spin=0.7, nr=49, radial_discretization="d10-5", ellmax_first=6,
                                ellmax_second=6, seeds=[(2, 2, 1e-3), (3, 2, 2e-4 - 5e-4j)]),
    "linear_d84_diss": H.small_case(spin=0.6, radial_discretization="d8-4", dissipation=0.004),
    "second_kerr": H.small_case(spin=0.6, compactification_length=1.5, nr=41,
                                radial_discretization="d10-5", ellmax_first=4, ellmax_second=6,
                                first_order_modes=[-2, 2], second_order_modes=[-4, 0, 4],
                                initial_data__add_sharp_partner=True,
                                seeds=[(2, 2, 1e-3), (3, 2, 4e-4 - 2e-4j)],
                                second_order__enabled=True,
                                second_order__source_mode="unrestricted"),
    "second_d42_latched_off": H.small_case(
        spin=0.8, compactification_length=1.5, nr=41, radial_discretization="d4-2",
        ellmax_first=4, ellmax_second=6, first_order_modes=[-2, 2],
        second_order_modes=[-4, 0, 4], initial_data__add_sharp_partner=True,
        seeds=[(2, 2, 1e-3)], second_order__enabled=True,
        second_order__source_mode="unrestricted", second_order__source_start_time=1.0),
}


def rms_max(x):
    a = np.abs(x) ** 2
    return math.sqrt(a.sum() / a.size), math.sqrt(a.max())


def expected(ref, params):
    state = ref.get_state()
    rhs = ref.evaluate_rhs(state)
    M, _, N, T_ = state.shape
    horizon = ref.setup.horizon
    h = horizon / (N - 1)
    D, _ = O.sbp_matrix(SCHEME[params["radial_discretization"]], N, h)
    want = {"state": [rms_max(state[:, f]) for f in range(13)],
            "rhs": [rms_max(rhs[:, f]) for f in range(13)]}
    want["first_constraint"] = rms_max(state[:, 1] - np.einsum("rs,mst->mrt", D, state[:, 2]))
    want["second_constraint"] = rms_max(state[:, 11] - np.einsum("rs,mst->mrt", D, state[:, 12]))
    second = params.get("second_order__enabled", False)
    zero = np.zeros((M, N, T_), dtype=np.complex128)
    if second:
        src, frc = ref.intermediate("source_over_r3"), ref.intermediate("forcing")
        cons, terms = ref.intermediate("constraints"), ref.intermediate("constraint_terms")
        rdr, ang = ref.intermediate("radial_derivatives"), ref.intermediate("angular_inputs")
    else:
        src = frc = zero
        cons = np.zeros((M, 3, N, T_), dtype=np.complex128)
        terms = np.zeros((M, 6, N, T_), dtype=np.complex128)
        rdr = np.zeros((M, 7, N, T_), dtype=np.complex128)
        ang = np.zeros((M, 6, N, T_), dtype=np.complex128)
    want["source_over_r3"] = rms_max(src)
    want["forcing"] = rms_max(frc)
    want["independent"] = [rms_max(cons[:, k]) for k in range(3)]
    want["independent_scale"] = np.abs(terms).max()
    # transport residuals L_X - R_X (reconstruction_diagnostics.hpp:46-140)
    x, _ = O.gauss_legendre(T_)
    mass, spin, L = params["mass"], params["spin"], params["compactification_length"]
    L2 = L * L
    modes = ref.setup.modes
    sharp = [modes.index(-m) if -m in modes else i for i, m in enumerate(modes)]
    res = np.zeros((7, M, N, T_), dtype=np.complex128)
    lhs_scale = np.zeros(7)
    for r in range(N):
        R = h * r
        for t in range(T_):
            bg = O.background(mass, spin, L, R, x[t], math.sqrt(1.0 - x[t] ** 2))
            mu, tau, pi = bg["mu0"], bg["tau0"], bg["pi0"]
            mub, pib, taub = np.conj(mu), np.conj(pi), np.conj(tau)
            for mi in range(M):
                F, G, La, Hh, B, Pi, C, U = state[mi, 2:10, r, t]
                Pis, Bs, Cs = np.conj(state[sharp[mi], [7, 6, 8], r, t])
                a = ang[mi, :, r, t]
                rX = [-4 * R * mu * G + a[0] - R * tau * F,
                      -R * (mu + mub) * La - F,
                      -3 * R * mu * Hh + a[1] - 2 * R * tau * G,
                      R * (mu - mub) * B - 2 * La,
                      -G - R * (pib + tau) * La + 0.5 * R * R * mu * (pib + tau) * B,
                      -R * mub * C - 2 * Pi - R * tau * B,
                      (-R * mub * U - R * mu * a[2] - R * R * mu * (pib + 2 * tau) * C
                       - 2 * a[3] - 2 * R * pib * Pi - 2 * Hh - 2 * R * pi * Pis
                       - R * pi * a[4] + R * R * pi * pi * Bs + R * mu * a[5]
                       + R * R * (-3 * mu * pi + 2 * mub * pi - 2 * mu * taub) * Cs)]
                for k, (v, n) in enumerate(zip((G, La, Hh, B, Pi, C, U), (2, 1, 3, 1, 2, 2, 3))):
                    lhs = ((2 + 4 * mass * R / L2) * rhs[mi, 3 + k, r, t]
                           + (R / L2) * (R * rdr[mi, k, r, t] + n * v))
                    res[k, mi, r, t] = lhs - rX[k]
                    lhs_scale[k] = max(lhs_scale[k], abs(lhs), abs(rX[k]))
    want["transport"] = [rms_max(res[k]) for k in range(7)]
    want["transport_scale"] = lhs_scale
    for key, f in (("first_horizon", 2), ("second_horizon", 12)):
        rows = state[:, f, N - 9:, :][:, ::-1, :]  # point j at radial N-1-j
        want[key] = [np.abs(np.einsum("j,mjt->mt", H.horizon_weights(d, h), rows)).max()
                     for d in range(5)]
    return want


def close(got, want, rtol, atol=0.0):
    return abs(got - want) <= rtol * abs(want) + atol


@pytest.mark.parametrize("name", sorted(CASES))
def test_diagnostics_match_oracle(name):
    params = CASES[name]
    dt = params["final_time"] / params["steps"]
    ref = H.oracle_initialized(params)
    ref.advance(3, dt)
    state = ref.get_state()
    print(f"[gpu] diagnostics {name}", flush=True)
    with T.Plan(W.config_text(**dict(params))) as plan:
        plan.initialize()
        for s in range(3):
            plan.step(s * dt, dt)
        plan.import_state(state)
        got = plan.sample_diagnostics(3 * dt)
        act = plan.source_activation()
    want = expected(ref, params)
    scale = np.abs(state).max()
    for f in range(13):
        fs = np.abs(state[:, f]).max()
        assert close(got["state_rms"][f], want["state"][f][0], 1e-12, 1e-14 * fs), f
        assert close(got["state_max"][f], want["state"][f][1], 1e-12, 1e-14 * fs), f
        # RHS fields 3..12 agree with the oracle to 1e-11 of their own scale
        rs = max(want["rhs"][f][1], 1e-300)
        assert close(got["rhs_rms"][f], want["rhs"][f][0], 1e-10, 1e-11 * rs), f
        assert close(got["rhs_max"][f], want["rhs"][f][1], 1e-10, 1e-11 * rs), f
    for key in ("first_constraint", "second_constraint"):
        f = 1 if key == "first_constraint" else 11
        qs = np.abs(state[:, f]).max()
        for k in range(2):
            assert close(got[key][k], want[key][k], 1e-9, 1e-13 * max(qs, 1e-300)), key
    for key in ("source_over_r3", "forcing"):
        s = max(want[key][1], 1e-300)
        for k in range(2):
            assert close(got[key][k], want[key][k], 1e-9, 1e-11 * s), key
    if name == "second_d42_latched_off":
        assert not act.active and got["forcing"] == (0.0, 0.0)
        assert want["source_over_r3"][1] > 0
    for k, key in enumerate(("psi3_bianchi", "psi2_bianchi", "hll_reality")):
        for j in range(2):
            assert close(got[key][j], want["independent"][k][j], 1e-8,
                         1e-11 * want["independent_scale"]), key
    for k, key in enumerate(("G", "Lambda", "H", "B", "Pi", "C", "U")):
        for j in range(2):
            assert close(got["transport"][key][j], want["transport"][k][j], 1e-8,
                         1e-11 * want["transport_scale"][k]), key
    for key in ("first_horizon", "second_horizon"):
        for d in range(5):
            assert close(got[key][d], want[key][d], 1e-9, 1e-13 * scale * (1e2 ** d)), (key, d)
    assert got["source_active"] == bool(act.active)
    if not params.get("second_order__enabled", False):
        # linear plans: no reconstruction, so the transport residuals reduce
        # to R tau0 F (G equation) and F (Lambda equation)
        assert got["transport"]["H"] == (0.0, 0.0) and got["psi3_bianchi"] == (0.0, 0.0)
        Fs = np.abs(state[:, 2]).max()
        assert close(got["transport"]["Lambda"][1], Fs, 1e-12)


def test_source_pair_report_matches_oracle_single_pair_targets():
    """Targets reached by one ordered pair (m = +-4 from (+-2, +-2)) carry
    that pair's D and T: its report equals the RMS and maximum of the
    oracle's ordered-pair sums at the target mode (solver_driver.hpp:251-280).
    Target 0 sums two pairs, whose reports add up to at least its sum."""
    params = CASES["second_kerr"]
    dt = params["final_time"] / params["steps"]
    ref = H.oracle_initialized(params)
    ref.advance(3, dt)
    state = ref.get_state()
    ref.evaluate_rhs(state)
    sums = ref.intermediate("pair_sums")  # (M, 2, N, T)
    modes = ref.setup.modes
    with T.Plan(W.config_text(**dict(params))) as plan:
        plan.initialize()
        plan.import_state(state)
        plan.sample_diagnostics(3 * dt)
        pairs = plan.source_pairs()
    assert [(p["m1"], p["m2"], p["target"]) for p in pairs] == \
        [(-2, -2, -4), (-2, 2, 0), (2, -2, 0), (2, 2, 4)]
    for p in pairs:
        v = sums[modes.index(p["target"])]
        for c, key in enumerate(("D", "T")):
            a = np.abs(v[c])
            rms, mx = math.sqrt((a ** 2).mean()), a.max()
            if p["target"] != 0:
                assert close(p[key + "_rms"], rms, 1e-11), (p, key)
                assert close(p[key + "_max"], mx, 1e-11), (p, key)
            else:
                assert p[key + "_max"] > 0.0
    zero = [p for p in pairs if p["target"] == 0]
    for c, key in enumerate(("D", "T")):
        a = np.abs(sums[modes.index(0), c]).max()
        assert zero[0][key + "_max"] + zero[1][key + "_max"] >= a * (1 - 1e-12)
