"""Host logic of libteuk_hip.so through the C ABI, no GPU needed: exported
symbols, the schema-v1 configuration (cases of the reference's
tests/test_config.cpp), SBP and angular tables against the oracle, mode
bookkeeping (tests/test_modes.cpp) and the synthetic workloads."""
from __future__ import annotations

import os
import re

import numpy as np
import pytest

import helpers as H
import oracle as O
import teuk_kokkos_amd as T
from teuk_kokkos_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "teuk_hip.h")).read()
    declared = set(re.findall(r"\b(teuk_[a-z_]+)\s*\(", header))
    assert declared, "no declarations found"
    assert declared == set(T.SIGNATURES), declared ^ set(T.SIGNATURES)
    lib = T.lib()
    for name in declared:
        assert hasattr(lib, name), name


def rejects(text, overrides=()):
    with pytest.raises(T.TeukInvalidArgument):
        T.resolve_config(text, overrides)


def test_config_rejects_duplicate_unknown_malformed():
    rejects("config_version = 1\nconfig_version = 1\n")
    rejects("config_version = 1\nspni = 0.7\n")
    rejects("config_version = 1\nspin = 0.7oops\n")
    rejects("config_version = 2\n")
    rejects("config_version = 1\ninitial_data.add_sharp_partner = maybe\n")
    rejects("config_version = 1\ninitial_data.mode.1.ell = 3\n")
    rejects("config_version = 1\n", ["spin=0.2", "spin=0.3"])
    rejects("config_version = 1\nspin = = 0.2\n")
    rejects("config_version = 1\nnr = 65.0\n")


def parse(text, overrides=()):
    out = {}
    for line in T.resolve_config(text, overrides).splitlines():
        k, v = line.split(" = ", 1)
        out[k] = v
    return out


def test_config_defaults_and_round_trip():
    d = parse("config_version = 1\n")
    assert d["mass"] == "1" and d["ellmax_first"] == "4" and d["ellmax_second"] == "4"
    assert d["second_order.enabled"] == "false" and d["plus2.mode"] == "disabled"
    assert d["plus2.ell_max_first"] == "4" and d["radial_discretization"] == "d4-2"
    assert d["initial_data.mode.0.ell"] == "2" and d["initial_data.mode.0.amplitude_real"] == "0.001"
    cfg = """
    config_version = 1
    spin = 0.55
    ellmax_first = 5
    ntheta = 9
    first_order_modes = -2,2
    second_order_modes = -4,0,4
    initial_data.mode.0.ell = 4
    initial_data.mode.0.m = 2
    initial_data.mode.0.amplitude_real = 1.25e-4
    initial_data.mode.0.amplitude_imag = -2.5e-5
    initial_data.compact_support = true
    radial_discretization = d8-4
    nr = 17
    second_order.enabled = true
    second_order.source_mode = unrestricted
    second_order.source_start_time = 0.125
    second_order.constraint_tolerance = 3e-9
    second_order.required_consecutive_passes = 3
    plus2.enabled = true
    plus2.mode = concurrent
    plus2.linear.method = both
    plus2.linear.evolve_validation = true
    plus2.second.method = sourced_companion
    plus2.second.initial_policy = checkpoint
    plus2.second.checkpoint = companion-start.bin
    plus2.ell_max_first = 5
    plus2.ell_max_second = 5
    plus2.output.regularized = false
    plus2.output.physical_tetrad_field = true
    plus2.output.source_families = true
    plus2.output.ordered_pairs = true
    output.directory = roundtrip-output
    """
    resolved = T.resolve_config(cfg)
    again = T.resolve_config(resolved)
    assert again == resolved
    d = parse(cfg)
    assert d["ellmax_second"] == "5" and d["second_order_modes"] == "-4,0,4"
    assert d["plus2.source_normalization.version"] == "2"
    assert d["plus2.source_normalization.name"] == "plus2_s0_over_r7_complete_field_factor_v2"
    assert float(d["initial_data.mode.0.amplitude_imag"]) == -2.5e-5


def test_config_radial_discretization_and_dissipation_guard():
    rejects("config_version = 1\nradial_discretization = d6-3\n")
    rejects("config_version = 1\nradial_discretization = d8-4\nnr = 15\n")
    rejects("config_version = 1\nradial_discretization = d10-5\nnr = 21\n")
    assert parse("config_version = 1\nradial_discretization = d10-5\nnr = 22\n")[
        "radial_discretization"] == "d10-5"
    near = ("config_version = 1\nspin = 0.999\nnr = 513\nntheta = 9\nellmax_first = 4\n"
            "ellmax_second = 4\nfinal_time = 200\nsteps = 400000\ndissipation = 0.005\n")
    assert parse(near + "radial_discretization = d8-4\n")["radial_discretization"] == "d8-4"
    rejects(near + "radial_discretization = d10-5\n")
    assert parse(near + "radial_discretization = d10-5\n", ["steps=1600000"])["steps"] == "1600000"


def test_config_plus2_combinations_fail_closed():
    rejects("config_version = 1\nplus2.mode = concurrent\n")
    rejects("config_version = 1\nplus2.enabled = true\n")
    rejects("config_version = 1\nplus2.linear.method = curvatureish\n")
    rejects("config_version = 1\nplus2.second.method = transformed_psi4\n")
    rejects("config_version = 1\nplus2.second.initial_policy = checkpoint\n")
    rejects("config_version = 1\nplus2.second.checkpoint = state.bin\n")
    rejects("config_version = 1\nplus2.enabled = true\nplus2.mode = concurrent\n")
    rejects("config_version = 1\nplus2.source_normalization.version = 3\n")


def test_config_band_padding_and_modes():
    rejects("config_version = 1\nellmax_first = 2\n")
    rejects("config_version = 1\nellmax_first = 6\nntheta = 9\n")   # (18+6+2)/2 = 13 > 9
    rejects("config_version = 1\nfirst_order_modes = 2\n")           # not closed under m -> -m
    rejects("config_version = 1\nfirst_order_modes = -2,2,2\n")
    rejects("config_version = 1\nsecond_order.enabled = true\n")     # daughters -4,0,4 missing
    ok = parse("config_version = 1\nsecond_order.enabled = true\n"
               "second_order.allow_truncated_daughter_modes = true\n")
    assert ok["second_order.allow_truncated_daughter_modes"] == "true"


@pytest.mark.parametrize("scheme", [0, 1, 2])
def test_product_sbp_tables_equal_oracle(scheme):
    L = T.lib()
    for n in {0: [8, 13, 30], 1: [16, 25, 40], 2: [22, 31, 48]}[scheme]:
        for r in range(n):
            assert L.teuk_sbp_norm_weight(scheme, n, r) == O.norm_weight(scheme, n, r)
            for c in range(n):
                assert (L.teuk_sbp_derivative_entry(scheme, n, r, c) ==
                        O.lib().oracle_sbp_derivative_entry(scheme, n, r, c))


@pytest.mark.parametrize("spin,m,ell_max,nodes", [(-2, 2, 10, 16), (-2, -3, 16, 25), (-1, 0, 6, 10),
                                                  (-2, 4, 24, 37)])
def test_product_band_matrices_equal_oracle(spin, m, ell_max, nodes):
    S, A = T.band_matrices(spin, m, ell_max, nodes)
    So, Ao = O.transform_matrices(spin, m, ell_max, nodes)
    assert np.array_equal(S, So) and np.array_equal(A, Ao)


def test_mode_registry_matches_reference_rules():
    # tests/test_modes.cpp:8-110
    cfg = ("config_version = 1\nellmax_first = 4\nellmax_second = 8\nntheta = 9\n"
           "first_order_modes = 2,-2\nsecond_order_modes = 4,0,-4\n")
    modes, sharp, pairs = T.config_modes(cfg)
    assert modes == [-4, -2, 0, 2, 4]
    assert sharp == [4, 3, 2, 1, 0]
    assert pairs == [(-2, -2, -4), (-2, 2, 0), (2, -2, 0), (2, 2, 4)]
    modes, sharp, pairs = T.config_modes("config_version = 1\n")
    assert modes == [-2, 2] and sharp == [1, 0] and pairs == []


def test_workloads_are_deterministic_and_admissible():
    a = W.texts(W.sweep(16))
    b = W.texts(W.sweep(16))
    assert a == b
    cases = W.sweep(512)
    assert {c["nr"] for c in cases} == {257, 513, 1025, 2049}
    assert {c["ellmax_first"] for c in cases} == {8, 12, 16, 24}
    assert all(0.0 <= c["spin"] <= 0.99 for c in cases)
    for c in cases[:64]:
        T.resolve_config(W.config_text(**dict(c)))
        dt = c["final_time"] / c["steps"]
        vmax = O.maximum_coordinate_speed(c["mass"], c["spin"], c["compactification_length"],
                                          c["nr"], c["ntheta"])
        h = W.cfl_time_step(c["mass"], c["spin"], c["compactification_length"], c["nr"],
                            c["ntheta"], 1.0) * W.maximum_coordinate_speed(
            c["mass"], c["spin"], c["compactification_length"], c["nr"], c["ntheta"])
        assert dt <= 0.5 * h / vmax


def test_plan_creation_without_gpu_fails_loudly():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises((T.TeukError, T.TeukInvalidArgument)):
        T.Plan("config_version = 1\n")


def _so_config(**over):
    base = dict(spin=0.6, compactification_length=1.5, nr=41, radial_discretization="d10-5",
                ellmax_first=8, ellmax_second=12, first_order_modes=[-2, 2],
                second_order_modes=[-4, 0, 4], initial_data__add_sharp_partner=True,
                seeds=[(2, 2, 1e-3)], second_order__enabled=True,
                second_order__source_mode="unrestricted", final_time=1e-9, steps=1)
    base.update(over)
    return W.config_text(**H.small_case(**base))


def test_second_order_plane_past_the_32_bit_background_offsets_is_refused():
    """The source kernel addresses a plane's background records (96 bytes
    each) with 32-bit offsets: a plane one node past 2^32 / 96 points is
    refused at plan creation, before any device work."""
    limit = (1 << 32) // 96
    nr = limit // 15 + 1  # ntheta 15
    with pytest.raises(T.TeukInvalidArgument, match="nr x ntheta <"):
        T.Plan(_so_config(nr=nr, ntheta=15))


def test_second_order_source_projection_past_the_lds_limit_is_refused():
    """Nine targets at ntheta 32 need 9 x 4 x 9 x 32 x 16 bytes of LDS in the
    source projection, more than a workgroup's 160 KiB: refused at plan
    creation."""
    with pytest.raises(T.TeukInvalidArgument, match="source projection"):
        T.Plan(_so_config(first_order_modes=[-2, -1, 0, 1, 2],
                          second_order_modes=[-4, -3, -2, -1, 0, 1, 2, 3, 4],
                          ellmax_second=24, ntheta=32, seeds=[(2, 2, 1e-3)]))
