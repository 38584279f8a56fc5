"""Second-order steps replayed as a captured hipGraph give the same bits as
direct launches, through the source switch-on.

A batch of second-order instances, each with its own source policy and
start: the unrestricted source switching on exactly at an accepted state
(10 dt, as the 64-instance bench batch does), an unrestricted start inside a
step (the exact two-part split of spatial_pipeline.hpp:590-620), a
constraint-aware gate (a constraint evaluation at every eligible accepted
state) and a source that stays off.  Graph mode replays the captured step
wherever no instance needs a device-side latch decision or a split; the
states, latches and constraint norms must equal those of direct launches
after every step, and the replay count shows which steps were replays."""
from __future__ import annotations

import numpy as np
import pytest

import helpers as H
import teuk_kokkos_amd as T
from teuk_kokkos_amd import workloads as W

pytestmark = pytest.mark.gpu


def so_case(**over):
    base = dict(spin=0.6, compactification_length=1.5, nr=41, radial_discretization="d10-5",
                ellmax_first=4, ellmax_second=6, first_order_modes=[-2, 2],
                second_order_modes=[-4, 0, 4], initial_data__add_sharp_partner=True,
                seeds=[(2, 2, 1e-3), (3, 2, 4e-4 - 2e-4j)], second_order__enabled=True,
                second_order__source_mode="unrestricted", steps=40, final_time=40e-3)
    base.update(over)
    return H.small_case(**base)


DT = 1e-3
BATCHES = {
    # the bench batch's policy alone: every step but the switch-on a replay
    "unrestricted_at_10dt": [so_case(second_order__source_start_time=10 * DT),
                             so_case(spin=0.3, second_order__source_start_time=10 * DT)],
    # a split step and device-side gates mixed in
    "mixed": [so_case(second_order__source_start_time=10 * DT),
              so_case(spin=0.2, second_order__source_start_time=6.5 * DT),
              so_case(spin=0.9, nr=45, second_order__source_mode="constraint_aware",
                      second_order__source_start_time=14 * DT),
              so_case(spin=0.0, second_order__source_start_time=30 * DT)],
}
STEPS = 20


def _flat(n):
    return np.array(n["absolute_maximum"] + n["normalized_maximum"] + n["weighted_rms"] +
                    n["normalized_weighted"] + [n["controlling_normalized"],
                                                float(n["all_finite"])])


def _run(cases, graph):
    states, latches, norms = [], [], []
    with T.Plan(W.texts(cases)) as plan:
        plan.set_graph(graph)
        plan.initialize()
        for s in range(STEPS):
            plan.advance(1)
            states.append([plan.export_state(i) for i in range(len(cases))])
            latches.append([plan.source_activation(i) for i in range(len(cases))])
            norms.append([plan.source_constraint_norms(i) for i in range(len(cases))])
        return states, latches, norms, plan.graph_replays()


@pytest.mark.parametrize("name", sorted(BATCHES))
def test_graph_replays_match_direct_launches_through_switch_on(name):
    cases = BATCHES[name]
    print(f"[gpu] second-order graph: {name}", flush=True)
    ref_states, ref_latch, ref_norms, ref_replays = _run(cases, graph=False)
    states, latch, norms, replays = _run(cases, graph=True)
    print(f"[gpu]   {replays} of {STEPS} steps replayed", flush=True)
    assert ref_replays == 0
    if name == "unrestricted_at_10dt":
        # the accumulated time of 9 steps plus dt may pass 10 dt by an ulp,
        # which makes that step an exact split (launched directly)
        assert replays >= STEPS - 1
    else:
        # step 6 (the split at 6.5 dt) and the steps whose accepted states
        # the gated instance evaluates (from 14 dt until it switches on)
        # launch directly
        assert 0 < replays < STEPS
    for s in range(STEPS):
        for i in range(len(cases)):
            assert np.array_equal(states[s][i].view(np.float64), ref_states[s][i].view(np.float64)), \
                (s, i)
            a, b = latch[s][i], ref_latch[s][i]
            assert (a.active, a.consecutive_passes, a.activation_time, a.last_eligibility_time) == \
                (b.active, b.consecutive_passes, b.activation_time, b.last_eligibility_time), (s, i)
            assert np.array_equal(_flat(norms[s][i]), _flat(ref_norms[s][i]), equal_nan=True), \
                (s, i)
    # the source is on where it should be by the end
    final = latch[-1]
    assert final[0].active and final[1].active
    if name == "mixed":
        assert not final[3].active
