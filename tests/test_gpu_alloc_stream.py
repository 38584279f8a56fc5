"""Allocation-free calls and the caller's execution stream (C ABI).

The reference asserts that its stage operator and stepper neither allocate
nor deep-copy once the pipeline exists (tests/test_spatial_pipeline.cpp:
102-125, tests/test_angular_device.cpp:65-92), and that they enqueue on the
execution-space instance the caller passes (include/teuk/
spatial_pipeline.hpp:458-462, 590).  Here a plan makes its device memory,
streams and events at creation, in the set_* calls and in teuk_initialize;
teuk_allocation_count counts every one the library makes, and must not move
across steps, graph replays, RHS evaluations, waveform and diagnostics
samples, imports, exports and checkpoint writes and loads.
teuk_plan_set_stream puts a plan's work on the caller's stream, ordered with
the caller's own work there."""
from __future__ import annotations

import numpy as np
import pytest

import helpers as H
import teuk_kokkos_amd as T
from teuk_kokkos_amd import workloads as W

pytestmark = pytest.mark.gpu

# two execution groups (ell 16 and ell 24 bands), so the step forks and
# joins the group streams
LINEAR = [H.small_case(spin=0.7, nr=61, radial_discretization="d10-5", ellmax_first=16,
                       ellmax_second=16, seeds=[(2, 2, 1e-3)]),
          H.small_case(spin=0.5, nr=53, radial_discretization="d10-5", ellmax_first=24,
                       ellmax_second=24, seeds=[(3, 2, 2e-3j)])]
SECOND = H.small_case(spin=0.6, compactification_length=1.5, nr=41,
                      radial_discretization="d10-5", ellmax_first=4, ellmax_second=6,
                      first_order_modes=[-2, 2], second_order_modes=[-4, 0, 4],
                      initial_data__add_sharp_partner=True, seeds=[(2, 2, 1e-3)],
                      second_order__enabled=True, second_order__source_mode="constraint_aware")


def _exercise(plan, graph, tmp_path):
    """Every steady-state call of the ABI once (buffers made beforehand)."""
    state = plan.export_state()
    d_in = T.DeviceArray(state)
    d_out = T.DeviceArray(np.zeros_like(state))
    if graph:
        plan.set_graph(True)
    plan.advance(1)
    plan.sync()
    before = T.allocation_count()
    plan.step(0.0, 1e-3)
    plan.advance(3)
    plan.evaluate_rhs_device(d_in.ptr, d_out.ptr)
    plan.sync()
    plan.sample_waveform()
    plan.sample_diagnostics(0.01)
    s = plan.export_state()
    plan.import_state(s)
    plan.write_checkpoint(str(tmp_path / f"ck_{graph}"))
    plan.load_checkpoint(str(tmp_path / f"ck_{graph}"))
    plan.advance(2)
    if plan.modal_points > 0:
        plan.set_mixed_gram(True)
        plan.advance(9)
        plan.set_mixed_gram(False)
    plan.sync()
    after = T.allocation_count()
    d_in.free()
    d_out.free()
    return before, after


@pytest.mark.parametrize("graph", [False, True])
def test_linear_calls_allocate_nothing(graph, tmp_path):
    with T.Plan(W.texts(LINEAR)) as plan:
        plan.initialize()
        assert len(plan.groups()) == 2
        before, after = _exercise(plan, graph, tmp_path)
    print(f"[gpu] linear graph={graph}: allocation count {before} -> {after}", flush=True)
    assert after == before


def test_second_order_calls_allocate_nothing(tmp_path):
    with T.Plan(W.config_text(**SECOND)) as plan:
        plan.initialize()
        before, after = _exercise(plan, False, tmp_path)
    print(f"[gpu] second order: allocation count {before} -> {after}", flush=True)
    assert after == before


def _hip():
    """The HIP runtime the library links (already loaded in this process)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipMemsetAsync.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p]
    hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                   ctypes.c_void_p]
    return hip


def test_work_runs_on_the_callers_stream_in_order():
    """RHS evaluations and steps enqueued on a stream the caller owns run in
    that stream's order with the caller's own work there: a fill with NaN
    bytes before and a copy after, on the same stream, bracket the plan's
    output; and the steps equal those on the plan's own stream bit for bit."""
    import ctypes
    hip = _hip()
    params = LINEAR[1]
    with T.Plan(W.config_text(**params)) as plan:
        plan.initialize()
        state = plan.export_state()
        d_in = T.DeviceArray(state)
        d_out = T.DeviceArray(np.zeros_like(state))
        plan.evaluate_rhs_device(d_in.ptr, d_out.ptr)
        plan.sync()
        own_rhs = d_out.download()
        plan.advance(4)
        own = plan.export_state()
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        with T.Plan(W.config_text(**params)) as plan:
            plan.initialize()
            d_copy = T.DeviceArray(np.zeros_like(state))
            plan.set_stream(stream.value)
            nbytes = state.nbytes
            assert hip.hipMemsetAsync(d_out.ptr, 0xFF, nbytes, stream) == 0
            plan.evaluate_rhs_device(d_in.ptr, d_out.ptr)
            assert hip.hipMemcpyAsync(d_copy.ptr, d_out.ptr, nbytes, 3, stream) == 0
            plan.advance(4)
            assert hip.hipStreamSynchronize(stream) == 0
            rhs = d_copy.download()
            stepped = plan.export_state()
            plan.set_stream(None)
            d_copy.free()
    finally:
        hip.hipStreamDestroy(stream)
        d_in.free()
        d_out.free()
    assert np.array_equal(rhs, own_rhs)
    assert np.array_equal(stepped, own)
