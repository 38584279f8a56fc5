"""HIP path vs CPU oracle on identical inputs (FP64 bar: <= 1e-12 relative
on the state).  All calls go through the C ABI."""
from __future__ import annotations

import numpy as np
import pytest

import helpers as H
import teuk_kokkos_amd as T
from teuk_kokkos_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1.0e-12  # relative, FP64 state parity

CASES = {
    "schw_d42": H.small_case(),
    "schw_d42_diss": H.small_case(dissipation=0.005),
    "kerr_d42": H.small_case(spin=0.7),
    "kerr_d84_diss": H.small_case(spin=0.6, radial_discretization="d8-4", dissipation=0.004),
    "kerr_d105": H.small_case(spin=0.7, nr=49, radial_discretization="d10-5", ellmax_first=6,
                              ellmax_second=6, seeds=[(2, 2, 1e-3), (3, 2, 2e-4 - 5e-4j)]),
    "kerr_d105_diss_constrained": H.small_case(
        spin=0.9, nr=53, radial_discretization="d10-5", dissipation=0.003,
        reduction_mode="stage_constrained", seeds=[(2, 2, 1e-3 + 2e-4j)]),
    "kerr_m0_m3_sharp": H.small_case(
        spin=0.5, nr=37, radial_discretization="d4-2", ellmax_first=5, ellmax_second=5,
        first_order_modes=[-3, 0, 3], second_order_modes=[-3, 0, 3],
        initial_data__add_sharp_partner=True, seeds=[(3, 3, 1e-3 - 1e-3j), (2, 0, 5e-4)]),
    # ell_max 24 / 32: the reference's A S differs from I by 6e-12 / 2e-9, so
    # these exercise the Gram-corrected modal projection.
    "kerr_ell24_gram": H.small_case(spin=0.95, nr=45, radial_discretization="d10-5",
                                    ellmax_first=24, ellmax_second=24, first_order_modes=[0],
                                    second_order_modes=[0], seeds=[(2, 0, 1e-3 + 1e-3j)]),
    "kerr_ell32_gram_diss": H.small_case(
        spin=0.999, compactification_length=1.0, nr=41, radial_discretization="d10-5",
        dissipation=0.001, ellmax_first=32, ellmax_second=32,
        seeds=[(2, 2, 1e-3), (5, 2, -3e-4j)]),
    # 35 coefficients per line: the Kerr kernels' widest (4 x 16) ell tiling.
    "kerr_ell36_wide": H.small_case(spin=0.8, nr=33, radial_discretization="d4-2",
                                    ellmax_first=36, ellmax_second=36,
                                    seeds=[(2, 2, 1e-3), (7, 2, 4e-4 + 1e-4j)]),
    # compact-support bump (the Galerkin P of the GPU initial data at its
    # edges, where the profile's derivatives are steep)
    "kerr_d84_compact": H.small_case(spin=0.5, nr=57, radial_discretization="d8-4",
                                     initial_data__compact_support=True, initial_data__width=0.3,
                                     seeds=[(2, 2, 1e-3), (4, 2, -5e-4j)]),
    "targets_only_modes": H.small_case(
        spin=0.3, ellmax_first=4, ellmax_second=6, first_order_modes=[-2, 2],
        second_order_modes=[-4, 0, 4]),
}


def _plan(params):
    print(f"[gpu] plan {params.get('radial_discretization')} nr={params['nr']} "
          f"a={params.get('spin')}", flush=True)
    return T.Plan(W.config_text(**dict(params)))


@pytest.mark.parametrize("name", sorted(CASES))
def test_initial_state_matches_oracle(name):
    params = CASES[name]
    with _plan(params) as plan:
        plan.initialize()
        gpu = plan.export_state()
    ref = H.oracle_initialized(params).get_state()
    assert gpu.shape == ref.shape
    assert H.rel_err(gpu, ref) <= TOL


@pytest.mark.parametrize("name", sorted(CASES))
def test_steps_match_oracle(name):
    params = CASES[name]
    dt = 0.2 * W.cfl_time_step(params["mass"], params["spin"], params["compactification_length"],
                               params["nr"], params["ntheta"], 1.0)
    ref = H.oracle_initialized(params)
    with _plan(params) as plan:
        plan.initialize()
        t = 0.0
        for _ in range(10):
            plan.step(t, dt)
            t += dt
        gpu = plan.export_state()
    ref.advance(10, dt)
    err = H.rel_err(gpu, ref.get_state())
    print(f"[gpu] {name} 10 steps rel err {err:.3e}")
    assert err <= TOL


@pytest.mark.parametrize("name", ["kerr_d105", "kerr_d105_diss_constrained", "kerr_m0_m3_sharp",
                                  "kerr_ell32_gram_diss"])
def test_evaluate_rhs_matches_oracle(name):
    params = CASES[name]
    ref = H.oracle_initialized(params)
    state = ref.get_state()
    expected = ref.evaluate_rhs(state)
    with _plan(params) as plan:
        plan.initialize()
        d_in = T.DeviceArray(state)
        d_out = T.DeviceArray(np.full_like(state, complex(7.0, 7.0)))
        plan.evaluate_rhs_device(d_in.ptr, d_out.ptr)
        plan.sync()
        got = d_out.download()
    assert np.all(got[:, 3:] == 0)
    assert H.rel_err(got, expected) <= TOL


def test_batch_equals_independent_runs():
    cases = [CASES["kerr_d105"], CASES["kerr_m0_m3_sharp"], CASES["schw_d42_diss"],
             CASES["kerr_d84_diss"]]
    with T.Plan([W.config_text(**dict(c)) for c in cases]) as plan:
        plan.initialize()
        plan.advance(7)
        got = [plan.export_state(i) for i in range(len(cases))]
        steps = [plan.progress(i)[1] for i in range(len(cases))]
    assert steps == [7] * len(cases)
    for c, g in zip(cases, got):
        ref = H.oracle_initialized(c)
        ref.advance(7, c["final_time"] / c["steps"])
        assert H.rel_err(g, ref.get_state()) <= TOL


def test_graph_replay_equals_eager():
    params = CASES["kerr_d105"]
    out = []
    for graph in (False, True):
        with _plan(params) as plan:
            plan.initialize()
            plan.set_graph(graph)
            plan.advance(3)
            plan.sync()
            plan.advance(4)
            out.append(plan.export_state())
    assert np.array_equal(out[0], out[1])


def test_tiling_does_not_change_results():
    params = CASES["kerr_d105_diss_constrained"]
    out = []
    for rows in (0, 1, 7, 16, 53):
        with _plan(params) as plan:
            plan.set_tile_rows(rows)
            plan.initialize()
            plan.advance(3)
            out.append(plan.export_state())
    for o in out[1:]:
        assert H.rel_err(o, out[0]) <= 1e-14


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("ell", [24, 32])
def test_gram_split_and_group_outputs_agree_and_match_oracle(ell, mixed):
    """Gram bands apply G = A S as I + E with the correction E X (and N Psi)
    on FP32 MFMA.  Tiles of at most five rows per thread take the
    one-point-per-thread routine, larger ones the accumulator-position routine;
    both must give the same bits, and at ell 32, where |E| ~ 2e-9, a missing
    or misplaced correction would exceed the FP64 bar against the oracle.
    With mixed precision on, the ell-24 band's correction runs on BF16 in
    both routines (ell 32 keeps FP32), again to the same bits."""
    params = H.small_case(spin=0.99, compactification_length=1.0, nr=61,
                          radial_discretization="d10-5", ellmax_first=ell, ellmax_second=ell,
                          seeds=[(2, 2, 1e-3), (6, 2, 2e-4 - 3e-4j)])
    out = {}
    # 16-row tiles fit the split routine (n_ell 23 or 31: 11 or 8 rows per
    # thread pass, five passes); one 61-row tile does not
    for rows in (16, 61):
        with _plan(params) as plan:
            plan.set_mixed_gram(mixed)
            plan.set_tile_rows(rows)
            assert {g["variant"] for g in plan.groups()} == {2}
            plan.initialize()
            plan.advance(4)
            out[rows] = plan.export_state()
    assert np.array_equal(out[16], out[61])
    ref = H.oracle_initialized(params)
    ref.advance(4, params["final_time"] / params["steps"])
    err = H.rel_err(out[16], ref.get_state())
    print(f"[gpu] Gram ell {ell} (mixed {mixed}) split/group outputs, 4 steps rel err {err:.3e}",
          flush=True)
    assert err <= TOL


@pytest.mark.parametrize("scheme,ell", [("d10-5", 16), ("d10-5", 24), ("d8-4", 16), ("d4-2", 24)])
def test_specialised_stage_kernels_match_general_kernel_and_oracle(scheme, ell):
    """Kerr lines without dissipation in free-damped mode run stage kernels
    with the line flags fixed at compile time (G = I for ell <= 16, the Gram
    projection at ell 24).  They must reproduce the general kernel bit for
    bit (the same arithmetic in the same order) and the oracle to the FP64
    bar; at ell 24 the mixed-precision Gram correction (BF16) agrees with
    the FP32 one to 1e-13."""
    params = H.small_case(spin=0.8, nr=229, radial_discretization=scheme, ellmax_first=ell,
                          ellmax_second=ell, seeds=[(2, 2, 1e-3), (5, 2, 3e-4j)])
    out = {}
    for special in (True, False):
        with _plan(params) as plan:
            plan.set_specialized(special)
            variants = {g["variant"] for g in plan.groups()}
            assert variants == ({2 if ell > 16 else 1} if special else {0}), variants
            plan.initialize()
            plan.advance(3)
            out[special] = plan.export_state()
    assert np.array_equal(out[True], out[False])
    if ell > 16:
        # mixed precision: the Kerr-Gram kernel takes the ell-24 band's Gram
        # correction on BF16 operands (the general kernel keeps FP32): the
        # correction is 6e-12 of k and BF16 holds it to 4e-3, so 1e-13
        with _plan(params) as plan:
            plan.set_mixed_gram(True)
            plan.initialize()
            plan.advance(3)
            mixed = plan.export_state()
        diff = H.rel_err(mixed, out[False])
        print(f"[gpu] mixed (BF16 correction) vs FP32 correction {diff:.3e}", flush=True)
        assert 0.0 < diff <= 1e-13
    ref = H.oracle_initialized(params)
    ref.advance(3, params["final_time"] / params["steps"])
    err = H.rel_err(out[True], ref.get_state())
    print(f"[gpu] specialised {scheme} ell {ell} 3 steps rel err {err:.3e}", flush=True)
    assert err <= TOL


def test_import_export_roundtrip_and_band_rejection():
    params = CASES["kerr_d105"]
    ref = H.oracle_initialized(params)
    ref.advance(3, 1e-3)
    state = ref.get_state()
    with _plan(params) as plan:
        off = plan.import_state(state)
        assert off < 5e-11
        back = plan.export_state()
        assert H.rel_err(back, state) <= 1e-14
        bad = state.copy()
        bad[1, 2, 10, :] += 1e-3 * np.cos(np.arange(bad.shape[-1]) * 2.1)
        with pytest.raises(T.TeukError):
            plan.import_state(bad)
        # failure leaves the state untouched
        assert H.rel_err(plan.export_state(), state) <= 1e-14


def test_waveform_records_follow_reference_order():
    params = CASES["targets_only_modes"]
    ref = H.oracle_initialized(params)
    with _plan(params) as plan:
        plan.initialize()
        recs = plan.sample_waveform()
    keys = [(r.order, r.boundary, r.m, r.ell) for r in recs]
    want = []
    for order, modes, emax in ((1, [-2, 2], 4), (2, [-4, 0, 4], 6)):
        for b in (0, 1):
            for m in modes:
                want += [(order, b, m, ell) for ell in range(max(2, abs(m)), emax + 1)]
    assert keys == want
    scri, hor = ref.endpoint_modal(ref.setup.modes.index(2))
    first = [r for r in recs if r.order == 1 and r.m == 2]
    got_scri = np.array([complex(r.rescaled_re, r.rescaled_im) for r in first if r.boundary == 0])
    got_hor = np.array([complex(r.rescaled_re, r.rescaled_im) for r in first if r.boundary == 1])
    assert np.abs(got_scri - scri).max() <= TOL * max(np.abs(scri).max(), 1e-300) + 1e-300
    assert np.abs(got_hor - hor).max() <= TOL * max(np.abs(hor).max(), 1e-30) + 1e-300
    L2 = params["compactification_length"] ** 2
    for r in recs:
        if r.boundary == 0:
            assert r.endpoint_re == L2 * r.rescaled_re


def test_invalid_inputs_fail_closed():
    with pytest.raises(T.TeukInvalidArgument):
        T.Plan(W.config_text(**dict(H.small_case(nr=7))))
    p = H.small_case()
    p["final_time"] = 10.0  # dt far above the CFL limit
    with pytest.raises(T.TeukInvalidArgument):
        T.Plan(W.config_text(**p))
    with _plan(H.small_case()) as plan:
        with pytest.raises(T.TeukInvalidArgument):
            plan.step(0.0, -1.0)
        with pytest.raises(T.TeukInvalidArgument):
            plan.step(float("nan"), 1e-3)


def _two_group_cases():
    """Instances in different execution groups (ell-tile classes 1 and 2)."""
    return [CASES["kerr_d105"], CASES["kerr_ell24_gram"], CASES["kerr_m0_m3_sharp"]]


def test_graph_replay_equals_eager_across_execution_groups():
    cases = _two_group_cases()
    out = []
    for graph in (False, True):
        with T.Plan([W.config_text(**dict(c)) for c in cases]) as plan:
            assert len(plan.groups()) >= 2
            plan.initialize()
            plan.set_graph(graph)
            plan.advance(2)
            plan.sync()
            plan.advance(3)
            out.append([plan.export_state(i) for i in range(len(cases))])
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def test_evaluate_rhs_for_an_instance_of_a_later_group():
    cases = _two_group_cases()
    i = 1  # the ell_max 24 instance
    ref = H.oracle_initialized(cases[i])
    state = ref.get_state()
    expected = ref.evaluate_rhs(state)
    with T.Plan([W.config_text(**dict(c)) for c in cases]) as plan:
        plan.initialize()
        d_in = T.DeviceArray(state)
        d_out = T.DeviceArray(np.zeros_like(state))
        plan.evaluate_rhs_device(d_in.ptr, d_out.ptr, instance=i)
        plan.sync()
        got = d_out.download()
    assert H.rel_err(got, expected) <= TOL


def test_step_and_group_timing_account_for_every_launch():
    cases = _two_group_cases()
    with T.Plan([W.config_text(**dict(c)) for c in cases]) as plan:
        plan.initialize()
        groups = plan.groups()
        plan.set_kernel_timing(True)
        plan.advance(3)
        ms, steps = plan.step_timing()
        total, launches = plan.kernel_timing()
        per_group = plan.groups()
        plan.set_kernel_timing(False)
    assert steps == 3 and ms > 0.0
    assert launches == 3 * 4 * len(groups)
    assert abs(sum(g["kernel_ms"] for g in per_group) - total) <= 1e-9 * max(total, 1.0)
    assert sum(g["tiles"] for g in groups) == sum(g["tiles"] for g in per_group)


def test_explicit_steps_are_timed_in_graph_mode():
    """teuk_step launches directly even in graph mode; with timing on its
    steps count in step_timing next to the replays of teuk_advance."""
    cases = _two_group_cases()
    with T.Plan([W.config_text(**dict(c)) for c in cases]) as plan:
        plan.initialize()
        plan.set_graph(True)
        plan.set_kernel_timing(True)
        plan.advance(2)
        t, _ = plan.progress()
        plan.step(t, plan.time_step())
        ms, steps = plan.step_timing()
        plan.set_kernel_timing(False)
        replays = plan.graph_replays()
    assert steps == 3 and ms > 0.0 and replays == 2


def test_multi_step_graph_replays_match_eager_steps():
    """teuk_advance replays a graph of eight RK4 steps (each execution
    group's stage chain running through all eight between one fork and one
    join) while eight or more steps remain, then one-step graphs: 19 steps
    are two eight-step and three one-step replays.  The state equals the
    eager steps' bit for bit, and step_timing counts every step."""
    cases = _two_group_cases()
    texts = [W.config_text(**dict(c)) for c in cases]
    with T.Plan(texts) as plan:
        plan.initialize()
        plan.set_graph(True)
        plan.set_kernel_timing(True)
        plan.advance(19)
        ms, steps = plan.step_timing()
        plan.set_kernel_timing(False)
        replays = plan.graph_replays()
        t, done = plan.progress()
        graph_state = [plan.export_state(i) for i in range(len(cases))]
    assert steps == 19 and replays == 19 and ms > 0.0 and done == 19
    with T.Plan(texts) as plan:
        plan.initialize()
        plan.advance(19)
        assert plan.progress() == (t, done)
        for i, g in enumerate(graph_state):
            assert np.array_equal(plan.export_state(i), g)


def test_graph_replays_are_timed_and_match_eager_steps():
    """With timing on, each graph replay is bracketed by HIP events on the
    plan's stream (no per-launch events inside the graph), and the replayed
    steps leave the state the eager steps leave."""
    cases = _two_group_cases()
    texts = [W.config_text(**dict(c)) for c in cases]
    with T.Plan(texts) as plan:
        plan.initialize()
        plan.set_graph(True)
        plan.set_kernel_timing(True)
        plan.advance(3)
        ms, steps = plan.step_timing()
        _, launches = plan.kernel_timing()
        plan.set_kernel_timing(False)
        graph_state = [plan.export_state(i) for i in range(len(cases))]
    assert steps == 3 and ms > 0.0 and launches == 0
    with T.Plan(texts) as plan:
        plan.initialize()
        plan.advance(3)
        for i, g in enumerate(graph_state):
            assert np.array_equal(plan.export_state(i), g)
