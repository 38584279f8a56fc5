"""Pin the CPU oracle before trusting it: the reference's own known answers
(tests/golden/reference_known_values.json), exact identities of the
operators, and closed forms.  Each check is chosen so that a plausible
mistake (a wrong coefficient, sign, index or transposed operand) fails."""
from __future__ import annotations

import json
import math
import os

import numpy as np
import pytest

import helpers as H
import oracle as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_known_values.json")))


# ------------------------------------------------------------- coefficients
@pytest.mark.parametrize("key", ["coefficient_point_oracle", "scri_limit"])
def test_coefficients_match_reference_point_values(key):
    g = GOLD[key]
    i = g["inputs"]
    c = O.teukolsky_coefficients(i["mass"], i["spin"], i["length"], i["spin_weight"], i["m"],
                                 i["radius"], i["theta"])
    tol = g["tolerance"]
    for name in ("time", "radial_advection", "radial_principal"):
        assert abs(c[name] - g[name]) <= tol
    for name in ("definition", "q", "psi"):
        assert abs(c[name] - complex(*g[name])) <= tol


def test_principal_coefficient_vanishes_at_scri_and_horizon():
    # tests/test_linear.cpp:28-46
    mass, spin, L = 1.0, 0.6, 2.0
    horizon = L * L / (mass + math.sqrt(mass * mass - spin * spin))
    assert abs(O.teukolsky_coefficients(mass, spin, L, -2, 0, 0.0, 0.4)["radial_principal"]) < 1e-14
    assert abs(O.teukolsky_coefficients(mass, spin, L, -2, 0, horizon, 0.4)["radial_principal"]) < 2e-14


# ------------------------------------------------------------------ SBP
SIZES = {0: [8, 9, 17, 30], 1: [16, 17, 25, 40], 2: [22, 23, 31, 48]}


@pytest.mark.parametrize("scheme", [0, 1, 2])
def test_sbp_identity(scheme):
    # H D + D^T H = diag(-1, 0, ..., 0, 1)  (tests/test_sbp105.cpp:75-96)
    for n in SIZES[scheme]:
        D, Hm = O.sbp_matrix(scheme, n, 1.4 / (n - 1))
        B = np.zeros((n, n))
        B[0, 0], B[-1, -1] = -1.0, 1.0
        assert np.abs(Hm @ D + D.T @ Hm - B).max() < 2e-13
        assert np.all(np.diag(Hm) > 0)


@pytest.mark.parametrize("scheme,boundary_degree,interior_degree", [(0, 2, 4), (1, 4, 8), (2, 5, 10)])
def test_sbp_polynomial_exactness(scheme, boundary_degree, interior_degree):
    n = 49
    x = np.linspace(-0.25, 1.1, n)
    h = x[1] - x[0]
    width = {0: 4, 1: 8, 2: 11}[scheme]
    for deg in range(0, interior_degree + 1):
        v = x ** deg
        d = O.sbp_derivative(scheme, v.astype(complex), h).real
        exact = deg * x ** (deg - 1) if deg else 0 * x
        if deg <= boundary_degree:
            assert np.abs(d - exact).max() < 2e-9
        inner = slice(width, n - width)
        assert np.abs(d[inner] - exact[inner]).max() < 2e-8 * max(1, deg)
    # the next interior degree is not exact
    v = x ** (interior_degree + 1)
    d = O.sbp_derivative(scheme, v.astype(complex), h).real
    assert np.abs(d[width:n - width] - (interior_degree + 1) * x[width:n - width] ** interior_degree).max() > 1e-12


@pytest.mark.parametrize("scheme,lo,hi,grid", [(0, 3.5, 3.5, (0.0, 1.2)), (1, 13.0, 14.0, (-0.1, 1.2)),
                                                (2, 25.0, 27.0, (-0.1, 1.2))])
def test_sbp_boundary_convergence_order(scheme, lo, hi, grid):
    # tests/test_sbp105.cpp:135-141, test_sbp84.cpp:91-97, test_sbp.cpp:139-147
    def err(n):
        x = np.linspace(grid[0], grid[1], n)
        if scheme == 0:
            v, dv = np.sin(2.1 * x) + 0.3 * np.cos(1.3 * x), 2.1 * np.cos(2.1 * x) - 0.39 * np.sin(1.3 * x)
        else:
            v, dv = np.sin(1.7 * x) + 0.2 * np.cos(2.3 * x), 1.7 * np.cos(1.7 * x) - 0.46 * np.sin(2.3 * x)
        d = O.sbp_derivative(scheme, v.astype(complex), x[1] - x[0]).real
        return np.abs(d - dv).max()
    e = [err(n) for n in (33, 65, 129)]
    assert e[0] / e[1] > lo and e[1] / e[2] > hi


@pytest.mark.parametrize("scheme,order", [(0, 3), (1, 5), (2, 6)])
def test_dissipation_energy_identity_and_exactness(scheme, order):
    # u* H Q u = -eps ||A u||^2 (tests/test_sbp105.cpp:143-178)
    n, eps = 31, 0.11
    x = np.linspace(-0.2, 1.0, n)
    h = x[1] - x[0]
    v = np.sin(8 * x) + 0.3 * np.cos(5 * x) + 0.4j * np.sin(3 * x)
    diss = O.sbp_dissipation(scheme, v, h, eps)
    weights = np.array([h * O.norm_weight(scheme, n, i) for i in range(n)])
    energy = np.sum(weights * np.conj(v) * diss)
    stencil = np.array([math.comb(order, k) * (-1) ** (order - k) for k in range(order + 1)])
    expected = -eps * sum(abs(np.dot(stencil, v[i:i + order + 1])) ** 2 for i in range(n - order))
    assert energy.real <= 0
    assert abs(energy.real - expected) < 4e-12
    assert abs(energy.imag) < 4e-12
    poly = sum(c * x ** k for k, c in enumerate([0.2, -1, 0.3, -0.4, 0.7, -0.2][:order]))
    assert np.abs(O.sbp_dissipation(scheme, poly.astype(complex), h, eps)).max() < 5e-8


def test_d105_dissipation_spectral_radius_and_rk4_bounds():
    g = GOLD["d105_dissipation_spectral_radius_n25"]
    n = 25
    w = np.array([O.norm_weight(2, n, i) for i in range(n)])
    Q = np.zeros((n, n))
    for c in range(n):
        e = np.zeros(n, complex)
        e[c] = 1.0
        Q[:, c] = -O.sbp_dissipation(2, e, 1.0, 1.0).real
    # symmetric in the H inner product: eigenvalues of H^1/2 Q H^-1/2
    Hs = np.sqrt(w)
    lam = np.linalg.eigvalsh((Hs[:, None] * Q) / Hs[None, :]).max()
    assert g["lower"] < lam < g["upper"]
    spacing = 0.64 / (n - 1)
    ref_dt = 2.785293563405282 * spacing / (0.006 * lam)
    assert g["reference_dt_range"][0] < ref_dt < g["reference_dt_range"][1]
    guarded = spacing * 0.8 * 2.785293563405282 / (0.006 * 5500.0)
    assert g["guarded_dt_range"][0] < guarded < g["guarded_dt_range"][1]


def test_dissipation_guard_bounds():
    g = GOLD["dissipation_spectral_radius_bounds"]
    for scheme, name in enumerate(("d4-2", "d8-4", "d10-5")):
        assert O.lib().oracle_dissipation_spectral_radius_bound(scheme) == g[name]


# ---------------------------------------------------------------- angular
def test_gauss_legendre_moments_and_symmetry():
    x, w = O.gauss_legendre(8)
    assert np.all(np.diff(x) > 0)
    assert np.abs(x + x[::-1]).max() < 2e-15 and np.abs(w - w[::-1]).max() < 2e-15
    for p in range(16):
        exact = 2.0 / (p + 1) if p % 2 == 0 else 0.0
        assert abs(np.sum(w * x ** p) - exact) < 3e-14
    xn, wn = np.polynomial.legendre.leggauss(8)
    assert np.abs(x - xn).max() < 1e-15 and np.abs(w - wn).max() < 1e-14


def test_harmonics_closed_forms():
    # tests/test_angular.cpp:97-113
    th = 0.73
    x = math.cos(th)
    assert abs(O.swsh_theta(0, 0, 0, th) - 1 / math.sqrt(4 * math.pi)) < 2e-15
    assert abs(O.swsh_theta(1, 0, 0, th) - math.sqrt(3 / (4 * math.pi)) * x) < 2e-15
    assert abs(O.swsh_theta(1, 1, 0, th) + math.sqrt(3 / (8 * math.pi)) * math.sin(th)) < 2e-15
    norm = math.sqrt(5 / (64 * math.pi))
    assert abs(O.swsh_theta(2, 2, -2, th) - norm * (1 + x) ** 2) < 3e-15
    assert abs(O.swsh_theta(2, 2, -2, 0.0) - 4 * norm) < 3e-15
    assert abs(O.swsh_theta(2, 2, -2, math.pi)) < 3e-15


@pytest.mark.parametrize("spin,m,ell_max,nodes", [(-2, 2, 8, 14), (-1, -2, 7, 12), (-2, 0, 10, 16),
                                                  (-2, -4, 24, 37)])
def test_transform_orthonormality(spin, m, ell_max, nodes):
    S, A = O.transform_matrices(spin, m, ell_max, nodes)
    # long-double factorial sums lose digits as ell grows (angular.hpp:170-206)
    assert np.abs(A @ S - np.eye(A.shape[0])).max() < (2e-12 if ell_max <= 16 else 5e-12)
    x, w = O.gauss_legendre(nodes)
    assert np.abs(A - 2 * math.pi * w[None, :] * S.T).max() < 1e-15


def test_dense_solve_matches_numpy():
    rng = np.random.default_rng(3)
    a = rng.normal(size=(9, 9)) + 1j * rng.normal(size=(9, 9))
    b = rng.normal(size=9) + 1j * rng.normal(size=9)
    x, res = O.dense_solve(a, b)
    assert np.abs(x - np.linalg.solve(a, b)).max() < 1e-12
    assert res < 1e-14


# ------------------------------------------------------------- pipeline
def test_gaussian_initial_data_is_consistent():
    # tests/test_pipeline_initial_data.cpp:27-140: seeded modal amplitudes,
    # Q = D psi, retained modal psi_t = 0, every field in band.
    params = H.small_case(spin=0.6, nr=21, compactification_length=1.2, ellmax_first=3,
                          ellmax_second=3, ntheta=6, first_order_modes=[-2, 0, 2],
                          second_order_modes=[-2, 0, 2], initial_data__center=0.4,
                          initial_data__width=0.14,
                          seeds=[(3, 2, 0.7 - 0.2j), (3, -2, -0.35 + 0.5j)])
    pipe = H.oracle_initialized(params)
    st = pipe.get_state()
    setup = pipe.setup
    h = setup.spacing
    x, _ = O.gauss_legendre(setup.ntheta)
    center = np.argmin(np.abs(np.arange(setup.nr) * h - 0.4 * setup.horizon))
    prof = math.exp(-((center * h - 0.4 * setup.horizon) / (0.14 * setup.horizon)) ** 2)
    for mi, m in enumerate(setup.modes):
        S, A = O.transform_matrices(-2, m, 3, setup.ntheta)
        modal = A @ st[mi, 2, center]
        want = np.zeros(A.shape[0], complex)
        for ell, mm, amp in params["seeds"]:
            if mm == m:
                want[ell - max(2, abs(m))] = amp * prof
        assert np.abs(modal - want).max() < 2e-12
        for f in (0, 1, 2):
            for r in range(setup.nr):
                v = st[mi, f, r]
                assert np.abs(S @ (A @ v) - v).max() < 3e-12
        for t in range(setup.ntheta):
            d = O.sbp_derivative(0, st[mi, 2, :, t], h)
            assert np.abs(d - st[mi, 1, :, t]).max() < 2e-14
        for r in range(setup.nr):
            nod = []
            for t in range(setup.ntheta):
                c = O.teukolsky_coefficients(1.0, 0.6, 1.2, -2, m, r * h, math.acos(x[t]))
                P, Q, Psi = st[mi, 0, r, t], st[mi, 1, r, t], st[mi, 2, r, t]
                nod.append((P + 2 * c["radial_advection"] * Q - c["definition"] * Psi) / c["time"])
            assert np.abs(A @ np.array(nod)).max() < 8e-13


def test_rhs_and_step_stay_in_band():
    params = H.small_case(spin=0.8, nr=41, radial_discretization="d8-4", dissipation=0.003,
                          ellmax_first=4, ellmax_second=4,
                          seeds=[(2, 2, 1.0), (4, 2, 0.3 - 0.1j)])
    pipe = H.oracle_initialized(params)
    st = pipe.get_state()
    rhs = pipe.evaluate_rhs(st)
    for out in (rhs,):
        for mi, m in enumerate(pipe.setup.modes):
            S, A = O.transform_matrices(-2, m, 4, pipe.setup.ntheta)
            for f in range(3):
                for r in range(pipe.setup.nr):
                    v = out[mi, f, r]
                    n = max(np.linalg.norm(v), 1e-300)
                    assert np.linalg.norm(S @ (A @ v) - v) / n < GOLD["band_invariance"]["rhs_offband_relative"]
    assert np.all(rhs[:, 3:] == 0)


def test_rk4_is_fourth_order_in_time():
    params = H.small_case(spin=0.5, nr=33, radial_discretization="d4-2", dissipation=0.0)
    T_end = 0.4
    def run(nsteps):
        p = H.oracle_initialized(params)
        p.advance(nsteps, T_end / nsteps)
        return p.get_state()
    ref = run(256)
    e1 = np.abs(run(16) - ref).max()
    e2 = np.abs(run(32) - ref).max()
    assert e1 / e2 > 13.0


def _fit_real(times, sig, dt, t0, t1):
    s11 = s12 = s22 = b1 = b2 = 0.0
    for k in range(len(sig) - 2):
        if times[k] < t0 or times[k + 2] > t1:
            continue
        y0, y1, y2 = sig[k], sig[k + 1], sig[k + 2]
        s11 += y1 * y1; s12 += y1 * y0; s22 += y0 * y0; b1 += y1 * y2; b2 += y0 * y2
    det = s11 * s22 - s12 * s12
    c1 = (b1 * s22 - b2 * s12) / det
    c2 = (s11 * b2 - s12 * b1) / det
    r = math.sqrt(-c2)
    return math.acos(max(-1.0, min(1.0, c1 / (2 * r)))) / dt, -math.log(r) / dt


def _fit_complex(times, sig, dt, t0, t1):
    # two-pole recurrence fit (tests/test_qnm.cpp:77-140)
    h11 = h12 = h22 = g1 = g2 = 0j
    for k in range(len(sig) - 2):
        if times[k] < t0 or times[k + 2] > t1:
            continue
        y0, y1, y2 = sig[k], sig[k + 1], sig[k + 2]
        h11 += np.conj(y1) * y1; h12 += np.conj(y1) * y0; h22 += np.conj(y0) * y0
        g1 += np.conj(y1) * y2; g2 += np.conj(y0) * y2
    det = h11 * h22 - h12 * np.conj(h12)
    c1 = (g1 * h22 - h12 * g2) / det
    c2 = (h11 * g2 - np.conj(h12) * g1) / det
    disc = np.sqrt(c1 * c1 + 4 * c2)
    roots = [0.5 * (c1 + disc), 0.5 * (c1 - disc)]
    freqs = [-np.angle(r) / dt for r in roots]
    damps = [-math.log(abs(r)) / dt for r in roots]
    sel = 0
    if freqs[0] <= 0 < freqs[1]:
        sel = 1
    elif freqs[0] > 0 and freqs[1] > 0 and damps[1] < damps[0]:
        sel = 1
    return freqs[sel], damps[sel]


def qnm_run(nr, ell_max, ntheta, spin, t_end, step_fn=None):
    """Production QNM ringdown (tests/test_qnm.cpp:142-215) on the oracle."""
    params = H.small_case(spin=spin, nr=nr, ellmax_first=ell_max, ellmax_second=ell_max,
                          ntheta=ntheta, dissipation=0.005, radial_discretization="d4-2",
                          initial_data__center=0.675, initial_data__width=0.08,
                          seeds=[(2, 2, 1.0)])
    pipe = H.oracle_initialized(params, threads=4)
    horizon = pipe.setup.horizon
    obs = int(round((1.0 / horizon) * (nr - 1)))
    S, A = O.transform_matrices(-2, 2, ell_max, ntheta)
    mi = pipe.setup.modes.index(2)
    steps = int(round(t_end / 0.01))
    times, sig = [], []
    for s in range(0, steps + 1, 5):
        times.append(s * 0.01)
        sig.append((A @ pipe.get_state()[mi, 2, obs])[0])
        if s != steps:
            pipe.advance(5, 0.01)
    return np.array(times), np.array(sig)


def test_schwarzschild_qnm_production_sequence():
    g = GOLD["schwarzschild_qnm"]
    seq = g["production_sequence"]
    errs = []
    for nr, want in zip(seq["nr"], seq["complex_frequency_error"]):
        t, s = qnm_run(nr, 3, 5, 0.0, 65.0)
        f, d = _fit_real(t, s.real, 0.05, 45.0, 65.0)
        e = math.hypot(f - g["omega"][0], d - g["omega"][1])
        errs.append(e)
        # the documented production error, to the digits quoted
        assert abs(e - want) < 0.006 * want
    assert errs[1] < 0.25 * errs[0] and errs[2] < 0.70 * errs[1] and errs[2] < 3e-4


def test_kerr_qnm_matches_leaver():
    g = GOLD["kerr_qnm_a07"]
    pr = g["production"]
    t, s = qnm_run(pr["nr"], pr["ell_max"], pr["ntheta"], pr["spin"], pr["fit_window"][1])
    f, d = _fit_complex(t, s, 0.05, *pr["fit_window"])
    err = math.hypot(f - g["leaver_omega"][0], d - g["leaver_omega"][1])
    assert err < pr["max_error"]
    assert abs(f - g["leaver_omega"][0]) < pr["max_component_error"]
    assert abs(d - g["leaver_omega"][1]) < pr["max_component_error"]
    # documented production value (docs/PRODUCTION_QNM_VALIDATION.md:72-82)
    assert abs(f - pr["omega"][0]) < 2e-6 and abs(d - pr["omega"][1]) < 2e-6
