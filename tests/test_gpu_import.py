"""Band-validated import on the GPU (pipeline_bands.hpp:108-232) against a
numpy statement of the same check built from the oracle's spin-weighted
transform matrices: every (mode, field, r) theta row x is projected onto its
field's band, |x - S A x| / |x| is reduced to its maximum, imports above the
tolerance are refused and leave the state untouched.  All calls go through
the C ABI."""
from __future__ import annotations

import numpy as np
import pytest

import helpers as H
import oracle as O
import teuk_kokkos_amd as T
from teuk_kokkos_amd import workloads as W

pytestmark = pytest.mark.gpu

# spin weight of each of the 13 pipeline fields (pipeline_bands.hpp:25-42)
SPINS = [-2, -2, -2, -1, -2, 0, -2, -1, -1, 0, -2, -2, -2]


def expected_offband(state, params, modes):
    """Largest relative off-band residual, each row against its own norm
    (pipeline_bands.hpp:127-153), each row measured against its field's band
    (oracle states carry content only on active fields)."""
    T_ = state.shape[-1]
    worst = 0.0
    for mi, m in enumerate(modes):
        for f in range(13):
            x = state[mi, f]
            if not np.any(x):
                continue
            E = params["ellmax_second"] if f >= 10 else params["ellmax_first"]
            S, A = O.transform_matrices(SPINS[f], m, E, T_)
            proj = x @ A.T @ S.T
            res = np.sqrt((np.abs(x - proj) ** 2).sum(axis=1))
            line = np.sqrt((np.abs(x) ** 2).sum(axis=1))
            worst = max(worst, float((res / np.maximum(line, np.finfo(float).tiny)).max()))
    return worst


def noisy(state, mode, field, r, level):
    """state with an off-band pattern of relative size `level` on one row"""
    out = state.copy()
    row = out[mode, field, r]
    pattern = np.cos(np.arange(row.size) * 2.1) + 1j * np.sin(np.arange(row.size) * 0.7)
    out[mode, field, r] = row + level * np.abs(row).max() * pattern
    return out


LINEAR = H.small_case(spin=0.7, nr=49, radial_discretization="d10-5", ellmax_first=6,
                      ellmax_second=6, seeds=[(2, 2, 1e-3), (3, 2, 2e-4 - 5e-4j)])
GRAM = H.small_case(spin=0.95, nr=45, radial_discretization="d10-5", ellmax_first=24,
                    ellmax_second=24, first_order_modes=[0], second_order_modes=[0],
                    seeds=[(2, 0, 1e-3 + 1e-3j)])
SECOND = H.small_case(spin=0.6, compactification_length=1.5, nr=41,
                      radial_discretization="d10-5", ellmax_first=4, ellmax_second=6,
                      first_order_modes=[-2, 2], second_order_modes=[-4, 0, 4],
                      initial_data__add_sharp_partner=True, seeds=[(2, 2, 1e-3)],
                      second_order__enabled=True, second_order__source_mode="unrestricted")


@pytest.mark.parametrize("name,params", [("linear", LINEAR), ("gram", GRAM), ("second", SECOND)])
def test_band_check_matches_numpy_and_roundtrips(name, params):
    dt = params["final_time"] / params["steps"]
    ref = H.oracle_initialized(params)
    ref.advance(3, dt)
    clean = ref.get_state()
    modes = ref.setup.modes
    r = params["nr"] // 3
    mi = modes.index(params["first_order_modes"][-1])  # a parent mode
    dirty = noisy(clean, mi, 2, r, 1e-9)
    print(f"[gpu] import {name}", flush=True)
    with T.Plan(W.config_text(**dict(params))) as plan:
        got_clean = plan.import_state(clean)
        want_clean = expected_offband(clean, params, modes)
        assert got_clean <= 5e-11 and want_clean <= 5e-11
        assert abs(got_clean - want_clean) <= 1e-14
        back = plan.export_state()
        assert H.rel_err(back, clean) <= 1e-14
        got = plan.import_state(dirty, tolerance=1e-6)
        want = expected_offband(dirty, params, modes)
        assert 1e-10 < want < 1e-8
        assert abs(got - want) <= 1e-6 * want, (got, want)
        # the split keeps the off-band part, so export returns the import
        assert H.rel_err(plan.export_state(), dirty) <= 1e-14
        # above the tolerance: refused, state untouched
        with pytest.raises(T.TeukError):
            plan.import_state(noisy(clean, mi, 2, r, 1e-6))
        assert H.rel_err(plan.export_state(), dirty) <= 1e-14


def test_import_into_one_instance_of_a_batch():
    cases = [LINEAR, GRAM, H.small_case(spin=0.3, nr=37, seeds=[(2, 2, 5e-4)])]
    with T.Plan([W.config_text(**dict(c)) for c in cases]) as plan:
        plan.initialize()
        plan.advance(2)
        before = [plan.export_state(i) for i in range(len(cases))]
        ref = H.oracle_initialized(cases[1])
        ref.advance(3, cases[1]["final_time"] / cases[1]["steps"])
        plan.import_state(ref.get_state(), 1)
        assert np.array_equal(plan.export_state(0), before[0])
        assert np.array_equal(plan.export_state(2), before[2])
        assert H.rel_err(plan.export_state(1), ref.get_state()) <= 1e-14


def test_band_check_is_strict_per_row():
    """Every row is measured against its own norm, as the reference does
    (pipeline_bands.hpp:127-153): an off-band row 1e-25 below its field's
    largest row is refused like one at 1e-9 of it, and a row of the same
    tiny size that is a band synthesis is accepted."""
    params = LINEAR
    ref = H.oracle_initialized(params)
    ref.advance(3, params["final_time"] / params["steps"])
    clean = ref.get_state()
    modes = ref.setup.modes
    mi = modes.index(2)
    T_ = clean.shape[-1]
    pattern = np.cos(np.arange(T_) * 2.1)
    scale = np.abs(clean[mi, 2]).max()
    S, _ = O.transform_matrices(-2, 2, params["ellmax_first"], T_)
    tiny_in_band = clean.copy()
    tiny_in_band[mi, 2, 0] = 1e-25 * scale * (S @ np.linspace(1.0, 2.0, S.shape[1]))
    with T.Plan(W.config_text(**dict(params))) as plan:
        assert plan.import_state(tiny_in_band) <= 5e-11
        assert H.rel_err(plan.export_state(), tiny_in_band) <= 1e-14
        for level in (1e-25, 1e-9):
            bad = clean.copy()
            bad[mi, 2, 0] = level * scale * pattern
            with pytest.raises(T.TeukError):
                plan.import_state(bad)


def test_second_order_checkpoint_state_passes_the_strict_check():
    """The state of a second-order plan after the source switch-on, scri
    rows of the reconstruction fields included (remainders of cancellations
    there), is a band synthesis on every row: the oracle's strict per-row
    measure stays at rounding, far below the reference's 5e-11 restart
    tolerance, and the plan re-imports it."""
    params = SECOND
    with T.Plan(W.config_text(**dict(params))) as plan:
        plan.initialize()
        plan.advance(8)
        state = plan.export_state()
        modes = sorted(set(params["first_order_modes"]) | set(params["second_order_modes"]))
        measured = expected_offband(state, params, modes)
        print(f"[gpu] second-order state after 8 steps: strict off-band {measured:.3e}", flush=True)
        assert measured <= 1e-13
        assert plan.import_state(state) <= 1e-13
