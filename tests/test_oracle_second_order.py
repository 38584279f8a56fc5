"""Pins of the oracle's second-order graph (CPU, no GPU).

The oracle restates the reference's reconstruction chain, quadratic source,
outer operator and activation latch (oracle/second_order.inc).  These tests
check it against the reference's own known answers and against what the
mathematics fixes, so that a dropped term, a wrong sign or a transposed
operand fails one of them:
  * background golden values (tests/test_background.cpp:8-34) and the
    Schwarzschild connection (tests/test_second_order.cpp:77-82);
  * the corrected half connection term (tests/test_second_order.cpp:55-75);
  * Jet1 tangents against exact central differences (the source is a
    quadratic form, so the central difference is exact up to rounding:
    tests/test_source_tangent_spatial.cpp), and quadratic scaling;
  * eth / eth' on spin-weighted harmonics against the closed-form raising
    and lowering factors (tests/test_ghp_device.cpp "Schwarzschild GHP plan
    reduces to modal raising and lowering factors"), and the rotating point
    formula (tests/test_background.cpp:83-115);
  * the reconstruction time solve inverting Delta_n
    (tests/test_reconstruction.cpp);
  * pipeline invariants: the disabled policy is exactly first order
    (tests/test_spatial_pipeline.cpp "disabled source policy ..."), every
    RHS field stays in its spin band ("full pipeline RHS and RK updates stay
    in every retained spin band"), the forcing scales quadratically, the
    inner-source tangents equal exact directional differences along the
    flow, the latch is fixed across RK stages and an endpoint event
    contributes nothing to the preceding step
    (tests/test_source_activation.cpp), fourth-order RK4 convergence of the
    complete graph (tests/test_full_graph_convergence.cpp), and the
    constraint gate on Gaussian data (tests/test_reconstruction_constraints.cpp
    "default Gaussian exposes inconsistent reconstruction constraints",
    tests/test_source_activation.cpp "normalized source eligibility is
    invariant under amplitude scaling")."""
from __future__ import annotations

import math

import numpy as np
import pytest

import helpers as H
import oracle as O

SPINS = [-2, -2, -2, -1, -2, 0, -2, -1, -1, 0, -2, -2, -2]


def second_order_case(**over):
    base = dict(spin=0.6, compactification_length=1.5, nr=41, radial_discretization="d10-5",
                ellmax_first=4, ellmax_second=6, first_order_modes=[-2, 2],
                second_order_modes=[-4, 0, 4], initial_data__add_sharp_partner=True,
                seeds=[(2, 2, 1e-3), (3, 2, 4e-4 - 2e-4j)], second_order__enabled=True,
                second_order__source_mode="unrestricted")
    base.update(over)
    return H.small_case(**base)


# ------------------------------------------------------------ background
def test_background_schwarzschild_exact_limit():
    bg = O.background(1.25, 0.0, 2.0, 0.3, 0.4, math.sqrt(0.84))
    want = {"mu0": -0.25, "tau0": 0.0, "pi0": 0.0, "rho0": -0.1015625,
            "epsilon0": 0.0390625, "psi20": -1.25 / 64.0}
    for k, v in want.items():
        assert abs(bg[k] - v) <= 1e-14, k


def test_background_rotating_equatorial_connection():
    spin, length = 0.8, 1.7
    bg = O.background(1.0, spin, length, 0.45, 0.0, 1.0)
    want = 1j * 3.0 * spin / (math.sqrt(2.0) * length ** 4)
    assert abs(np.conj(bg["pi0"]) + 2.0 * bg["tau0"] - want) <= 1e-14


def test_background_schwarzschild_connection_vanishes():
    bg = O.background(1.0, 0.0, 1.2, 0.5, 0.3, 0.7)
    assert abs(np.conj(bg["pi0"]) + 2.0 * bg["tau0"]) <= 1e-15


# ------------------------------------------------------------ pair source
F_NAMES = ["F1", "G1", "Lambda1", "Pi1", "B1", "C1", "U1", "F2", "G2", "H2", "B2", "C2", "U2",
           "U2s", "C2s", "C1s", "B1s", "Pi2s", "B2s"]
D_NAMES = ["d1F2", "d3U2", "eth2C2", "ethp2C2s", "eth1B2", "d2C2", "d2G2", "eth2G2", "ethp1F2",
           "d2C2s", "ethp1B2s"]


def _jets(values, tangents):
    return [(v, t) for v, t in zip(values, tangents)]


def test_pair_source_half_connection_term():
    radius = 0.6
    bg = O.background(1.0, 0.8, 1.4, radius, 0.0, 1.0)
    f = [0j] * 19
    d = [0j] * 11
    f[F_NAMES.index("F1")] = 0.7 - 0.2j
    f[F_NAMES.index("C2")] = -0.3 + 0.9j
    d[D_NAMES.index("eth2C2")] = 0.4 + 0.1j
    (D, _), _ = O.pair_source(radius, bg, _jets(f, [0j] * 19), _jets(d, [0j] * 11))
    conn = radius ** 2 * (np.conj(bg["pi0"]) + 2.0 * bg["tau0"]) * f[F_NAMES.index("C2")]
    F1 = f[F_NAMES.index("F1")]
    correct = F1 * (0.5 * radius * d[D_NAMES.index("eth2C2")] + 0.5 * conn)
    legacy = F1 * (0.5 * radius * d[D_NAMES.index("eth2C2")] + conn)
    assert abs(D - correct) <= 1e-14
    assert abs(D - legacy) > 1e-4


def _random_jets(rng, n):
    v = rng.normal(size=n) + 1j * rng.normal(size=n)
    t = rng.normal(size=n) + 1j * rng.normal(size=n)
    return v, t


def test_pair_source_tangent_is_exact_directional_difference():
    rng = np.random.default_rng(7)
    bg = O.background(1.0, 0.63, 1.2, 0.31, 0.2, math.sqrt(0.96))
    fv, ft = _random_jets(rng, 19)
    dv, dt = _random_jets(rng, 11)
    (D, Dt), (T, Tt) = O.pair_source(0.31, bg, _jets(fv, ft), _jets(dv, dt))
    z19, z11 = [0j] * 19, [0j] * 11
    for eps in (0.5, 1e-3):
        (Dp, _), (Tp, _) = O.pair_source(0.31, bg, _jets(fv + eps * ft, z19),
                                         _jets(dv + eps * dt, z11))
        (Dm, _), (Tm, _) = O.pair_source(0.31, bg, _jets(fv - eps * ft, z19),
                                         _jets(dv - eps * dt, z11))
        assert abs((Dp - Dm) / (2 * eps) - Dt) <= 1e-11 * (abs(Dt) + abs(D))
        assert abs((Tp - Tm) / (2 * eps) - Tt) <= 1e-11 * (abs(Tt) + abs(T))
    # quadratic form: scaling every input by lam scales the source by lam^2
    lam = 3.25
    (Ds, Dts), (Ts, Tts) = O.pair_source(0.31, bg, _jets(lam * fv, lam * ft),
                                         _jets(lam * dv, lam * dt))
    assert abs(Ds - lam ** 2 * D) <= 1e-13 * abs(lam ** 2 * D)
    assert abs(Ts - lam ** 2 * T) <= 1e-13 * abs(lam ** 2 * T)
    assert abs(Dts - lam ** 2 * Dt) <= 1e-13 * abs(lam ** 2 * Dt)


def test_pair_source_every_term_is_live():
    """Each field and derivative reaches D or T: zeroing any one input
    changes the source (a dropped term would leave it unchanged).  U2 is the
    one member of OrderedPairFields the reference's formula never reads
    (second_order.hpp:74-142 uses U1 and U2_sharp), so it must stay dead."""
    rng = np.random.default_rng(11)
    bg = O.background(1.0, 0.7, 1.3, 0.4, 0.35, math.sqrt(1 - 0.35 ** 2))
    fv, ft = _random_jets(rng, 19)
    dv, dt = _random_jets(rng, 11)
    base = O.pair_source(0.4, bg, _jets(fv, ft), _jets(dv, dt))
    for k in range(19):
        f2 = fv.copy()
        f2[k] = 0
        got = O.pair_source(0.4, bg, _jets(f2, ft), _jets(dv, dt))
        change = abs(got[0][0] - base[0][0]) + abs(got[1][0] - base[1][0])
        if F_NAMES[k] == "U2":
            assert change == 0
        else:
            assert change > 1e-8, F_NAMES[k]
    for k in range(11):
        d2 = dv.copy()
        d2[k] = 0
        got = O.pair_source(0.4, bg, _jets(fv, ft), _jets(d2, dt))
        assert abs(got[0][0] - base[0][0]) + abs(got[1][0] - base[1][0]) > 1e-8, D_NAMES[k]


# ------------------------------------------------------------ GHP operators
@pytest.mark.parametrize("spin,boost,m", [(-2, -2, 2), (-1, -1, -1), (-1, 1, 0), (0, 0, 2),
                                          (2, 0, -2), (1, 1, 1), (-2, 0, 0)])
def test_schwarzschild_eth_reduces_to_raising_and_lowering(spin, boost, m):
    E, T, L = 6, 9, 1.3
    x, _ = O.gauss_legendre(T)
    theta = np.arccos(x)
    lmin = max(abs(spin), abs(m))
    rng = np.random.default_rng(3)
    coef = {l: complex(rng.normal(), rng.normal()) for l in range(lmin, E + 1)}
    row = np.array([sum(c * O.swsh_theta(l, m, spin, th) for l, c in coef.items())
                    for th in theta])
    dt = rng.normal(size=T) + 1j * rng.normal(size=T)
    eth = O.ghp_row(spin, boost, m, E, row, dt, prime=False, kerr_spin=0.0, length=L)
    want = np.zeros(T, complex)
    for l, c in coef.items():
        if l >= max(abs(spin + 1), abs(m)):
            f = math.sqrt((l - spin) * (l + spin + 1))
            want += c * f * np.array([O.swsh_theta(l, m, spin + 1, th) for th in theta])
    assert np.abs(eth - want / (math.sqrt(2) * L * L)).max() <= 1e-13
    ethp = O.ghp_row(spin, boost, m, E, row, dt, prime=True, kerr_spin=0.0, length=L)
    want = np.zeros(T, complex)
    for l, c in coef.items():
        if l >= max(abs(spin - 1), abs(m)):
            f = -math.sqrt((l + spin) * (l - spin + 1))
            want += c * f * np.array([O.swsh_theta(l, m, spin - 1, th) for th in theta])
    assert np.abs(ethp - want / (math.sqrt(2) * L * L)).max() <= 1e-13


def test_rotating_eth_point_formula_with_connection():
    spin, boost, m, E, T = -1, 0, 1, 5, 8
    a, L, R = 0.71, 1.1, 0.27
    x, _ = O.gauss_legendre(T)
    theta = np.arccos(x)
    row = np.array([O.swsh_theta(3, m, spin, th) for th in theta]) * (0.3 - 0.8j)
    dt = np.array([O.swsh_theta(2, m, spin, th) for th in theta]) * (0.4 + 0.2j)
    got = O.ghp_row(spin, boost, m, E, row, dt, kerr_spin=a, length=L, radius=R)
    raised = (0.3 - 0.8j) * math.sqrt((3 - spin) * (3 + spin + 1)) * np.array(
        [O.swsh_theta(3, m, spin + 1, th) for th in theta])
    sin = np.sqrt(1 - x * x)
    den = L * L - 1j * a * R * x
    want = (-1j * a * sin * dt + raised) / (math.sqrt(2) * den) - \
        1j * (spin + boost) * a * R * sin * row / (math.sqrt(2) * den * den)
    assert np.abs(got - want).max() <= 1e-13


# ------------------------------------------------------------ reconstruction
@pytest.mark.parametrize("n", [1, 2, 3])
def test_reconstruction_time_solve_inverts_delta_n(n):
    value, dr, rhs = 0.7 - 0.2j, 0.3 + 0.8j, -0.4 + 0.6j
    R, M, L = 0.35, 1.0, 1.4
    dt = O.recon_dt(value, dr, rhs, n, R, M, L)
    assert abs(O.delta_n(value, dt, dr, n, R, M, L) - rhs) <= 1e-15
    L2 = L * L
    assert abs(O.delta_n(value, dt, dr, n, R, M, L) -
               ((2 + 4 * M * R / L2) * dt + (R / L2) * (R * dr + n * value))) <= 1e-15


# ------------------------------------------------------------ pipeline
def _evolved(params, steps=3):
    pipe = H.oracle_initialized(params)
    pipe.advance(steps, 2e-3)
    return pipe


def test_disabled_policy_is_exactly_first_order():
    on = _evolved(second_order_case())
    state = on.get_state()
    off = H.oracle_pipeline(second_order_case(second_order__enabled=False))
    k_on = on.evaluate_rhs(state)
    k_off = off.evaluate_rhs(state)
    assert np.array_equal(k_on[:, :3], k_off[:, :3])
    assert np.all(k_off[:, 3:] == 0)
    assert np.abs(k_on[:, 3:10]).max() > 0 and np.abs(k_on[:, 10:]).max() > 0


def test_rhs_stays_in_every_spin_band():
    p = second_order_case()
    pipe = _evolved(p)
    k = pipe.evaluate_rhs(pipe.get_state())
    modes = pipe.setup.modes
    T = p["ntheta"]
    for mi, m in enumerate(modes):
        for f in range(13):
            targets = f >= 10
            active = m in (p["second_order_modes"] if targets else p["first_order_modes"])
            v = k[mi, f]
            if not active:
                assert np.all(v == 0), (m, f)
                continue
            E = p["ellmax_second"] if targets else p["ellmax_first"]
            S, A = O.transform_matrices(SPINS[f], m, E, T)
            proj = v @ (S @ A).T
            assert np.abs(proj - v).max() <= 1e-12 * max(np.abs(v).max(), 1e-300), (m, f)


def test_forcing_scales_quadratically_with_first_order_data():
    p = second_order_case()
    pipe = _evolved(p)
    y = pipe.get_state()
    y[:, 10:] = 0
    pipe.evaluate_rhs(y)
    f1 = pipe.intermediate("forcing")
    pipe.evaluate_rhs(2.5 * y)
    f2 = pipe.intermediate("forcing")
    assert np.abs(f1).max() > 0
    assert np.abs(f2 - 6.25 * f1).max() <= 1e-12 * np.abs(6.25 * f1).max()


def test_inner_source_tangent_is_exact_flow_difference():
    """D and T are quadratic forms of the stage (fields 0..9 evolve
    linearly), so their derivative along the flow k = RHS(y) equals the
    central difference of D(y +- d k) for any d, up to rounding."""
    p = second_order_case()
    pipe = _evolved(p)
    y = pipe.get_state()
    k = pipe.evaluate_rhs(y)
    inner = pipe.intermediate("inner")
    tangent = pipe.intermediate("inner_tangent")
    d = 0.05
    pipe.evaluate_rhs(y + d * k)
    plus = pipe.intermediate("inner")
    pipe.evaluate_rhs(y - d * k)
    minus = pipe.intermediate("inner")
    fd = (plus - minus) / (2 * d)
    scale = np.abs(tangent).max() + np.abs(inner).max() / d
    assert np.abs(tangent).max() > 0
    assert np.abs(fd - tangent).max() <= 1e-9 * scale


def test_latch_is_fixed_across_stages_and_split_at_start_time():
    dt = 2e-3
    pipe = H.oracle_initialized(second_order_case(second_order__source_start_time=2.5 * dt))
    assert not pipe.activation()["active"]
    flags = []
    for s in range(4):
        pipe.step(s * dt, dt)
        flags.append(pipe.last_stage_sources())
    assert flags[:2] == [[0, 0, 0, 0]] * 2
    assert flags[2] == [0, 0, 0, 0, 1, 1, 1, 1]
    assert flags[3] == [1, 1, 1, 1]
    act = pipe.activation()
    assert act["active"] and act["activation_time"] == 2.5 * dt


def test_endpoint_event_contributes_nothing_to_preceding_step():
    dt = 2e-3
    a = H.oracle_initialized(second_order_case(second_order__source_start_time=2 * dt))
    b = H.oracle_initialized(second_order_case(second_order__enabled=False))
    for s in range(2):
        a.step(s * dt, dt)
        b.step(s * dt, dt)
        assert a.last_stage_sources() == [0, 0, 0, 0]
    assert a.activation()["active"]
    # first order identical, second order untouched by the inactive source
    assert np.array_equal(a.get_state()[:, :3], b.get_state()[:, :3])
    assert np.all(a.get_state()[:, 10:] == 0)
    a.step(2 * dt, dt)
    assert a.last_stage_sources() == [1, 1, 1, 1]
    assert np.abs(a.get_state()[:, 10:]).max() > 0


def test_complete_graph_rk4_is_fourth_order():
    p = second_order_case(nr=29)
    T_end = 0.04
    finals = {}
    for n in (4, 8, 16, 32):
        pipe = H.oracle_initialized(p)
        pipe.advance(n, T_end / n)
        finals[n] = pipe.get_state()
    e1 = np.abs(finals[4][:, 10:] - finals[32][:, 10:]).max()
    e2 = np.abs(finals[8][:, 10:] - finals[32][:, 10:]).max()
    e3 = np.abs(finals[16][:, 10:] - finals[32][:, 10:]).max()
    # Richardson-corrected ratios against the finest run
    r1 = e1 / e2 * (1 - 1 / 8 ** 4) / (1 - 1 / 4 ** 4) if e2 > 0 else 0
    r2 = e2 / e3 * (1 - 1 / 4 ** 4) / (1 - 1 / 2 ** 4) if e3 > 0 else 0
    assert 13.0 < r1 < 19.0 and 13.0 < r2 < 19.0, (r1, r2)


def test_constraint_gate_rejects_gaussian_data_and_is_scale_invariant():
    p = second_order_case(second_order__source_mode="constraint_aware")
    pipe = H.oracle_initialized(p)
    assert not pipe.consider_activation(0.0)
    n1 = pipe.constraint_norms()
    assert n1["all_finite"] and n1["controlling_normalized"] > 1e-10
    act = pipe.activation()
    assert not act["active"] and act["last_eligibility_time"] == 0.0
    y = pipe.get_state()
    other = H.oracle_pipeline(p)
    other.set_state(3.0 * y)
    other.consider_activation(0.0)
    n2 = other.constraint_norms()
    assert abs(n2["controlling_normalized"] - n1["controlling_normalized"]) <= \
        1e-12 * n1["controlling_normalized"]
    for f in range(3):
        assert abs(n2["absolute_maximum"][f] - 3.0 * n1["absolute_maximum"][f]) <= \
            1e-12 * 3.0 * n1["absolute_maximum"][f] + 1e-300
