"""teuk_evaluate_rhs_noalloc against teuk_evaluate_rhs (C ABI).

The no-allocation entry point runs the linear RHS kernels with their modal
staging in a fixed per-device array instead of the plan's buffers.  Its
output must equal teuk_evaluate_rhs bit for bit, on two calls in a row (the
second on a smaller plan, so the array still holds the first call's larger
staging), and it must refuse plans it cannot serve without touching the
output."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import helpers as H
import teuk_kokkos_amd as T
from teuk_kokkos_amd import workloads as W

pytestmark = pytest.mark.gpu

SCRATCH_HALF = 1 << 20  # complex doubles per staging buffer


def _rhs_both(plan):
    """The RHS of the plan's current state from both entry points."""
    state = plan.export_state()
    d_in = T.DeviceArray(state)
    d_ref = T.DeviceArray(np.full_like(state, np.nan))
    d_new = T.DeviceArray(np.full_like(state, np.nan))
    try:
        lib = T.lib()
        rc = lib.teuk_evaluate_rhs(plan._h, 0, ctypes.c_void_p(d_in.ptr),
                                   ctypes.c_void_p(d_ref.ptr), 0.0)
        assert rc == T.TEUK_OK, T.lib().teuk_last_error(plan._h).decode()
        before = T.allocation_count()
        rc = lib.teuk_evaluate_rhs_noalloc(plan._h, 0, ctypes.c_void_p(d_in.ptr),
                                           ctypes.c_void_p(d_new.ptr), 0.0)
        assert rc == T.TEUK_OK, T.lib().teuk_last_error(plan._h).decode()
        plan.sync()
        assert T.allocation_count() == before
        return d_ref.download(), d_new.download(), state
    finally:
        for d in (d_in, d_ref, d_new):
            d.free()


@pytest.mark.parametrize("first,second", [("case4", "case1")])
def test_noalloc_rhs_matches_the_allocating_entry_point(first, second):
    for name in (first, second):
        params = getattr(W, name)()
        print(f"[gpu] noalloc rhs: {name} nr {params['nr']} ell {params['ellmax_first']}",
              flush=True)
        with T.Plan(W.config_text(**params)) as plan:
            plan.initialize()
            plan.advance(3)
            extent = 3 * plan.modal_points
            assert extent <= SCRATCH_HALF
            print(f"[gpu]   {extent:.0f} modal values per staging buffer", flush=True)
            ref, new, state = _rhs_both(plan)
        assert np.isfinite(ref).all() and np.abs(ref).max() > 0
        assert np.array_equal(new.view(np.float64), ref.view(np.float64)), name


def test_noalloc_rhs_refuses_plans_beyond_the_scratch_and_second_order_plans():
    big = dict(W.case4(), first_order_modes=[-3, -2, 2, 3],
               second_order_modes=[-3, -2, 2, 3])
    second = H.small_case(spin=0.6, compactification_length=1.5, nr=41,
                          radial_discretization="d10-5", ellmax_first=4, ellmax_second=6,
                          first_order_modes=[-2, 2], second_order_modes=[-4, 0, 4],
                          initial_data__add_sharp_partner=True, seeds=[(2, 2, 1e-3)],
                          second_order__enabled=True,
                          second_order__source_mode="constraint_aware")
    for params, words in ((big, "fixed scratch"), (second, "second-order")):
        with T.Plan(W.config_text(**params)) as plan:
            if words == "fixed scratch":
                assert 3 * plan.modal_points > SCRATCH_HALF
            shape = plan.shape()
            n = int(np.prod(shape))
            fill = np.full(n, 7.0 + 3.0j)
            d_in = T.DeviceArray(np.zeros(n, dtype=np.complex128))
            d_out = T.DeviceArray(fill)
            try:
                rc = T.lib().teuk_evaluate_rhs_noalloc(plan._h, 0, ctypes.c_void_p(d_in.ptr),
                                                       ctypes.c_void_p(d_out.ptr), 0.0)
                assert rc == T.TEUK_EINVAL
                assert words in T.lib().teuk_last_error(plan._h).decode()
                plan.sync()
                assert np.array_equal(d_out.download(), fill)
            finally:
                d_in.free()
                d_out.free()
