"""Pins of the oracle's second-order point functions to the reference's own
equation data (CPU, no GPU).

tests/golden/second_order_spec_vectors.json holds numeric vectors that
tests/golden/gen_spec_vectors.py evaluated, at 50 digits, from the
reference's machine-readable equations (equation_spec_v2.yaml) and its term
ledger (SOURCE_TERM_LEDGER.csv):
  * the rescaled Kerr background mu0, tau0, pi0, psi20 (spec: background);
  * the coordinate forms of Delta_n, eth_n and eth'_n (rescaled_operators);
  * the ordered-pair source D_12, T_12 with Jet1 tangents
    (compact_ordered_pair_source; second_order.hpp:74-142);
  * every ledger row sd01-sd15, st01-st07, S01-S04 alone, at two radii with
    the background held fixed, and its explicit R power;
  * S_over_R3 and the coordinate forcing 2(L^4 + a^2 R^2 cos^2) S / R^3
    (second_order.hpp:152-182);
  * the seven rescaled reconstruction equations Delta_n X = R_X and their
    time solve (metric_reconstruction_rescaled; reconstruction.hpp:48-121);
  * the linear reduction coefficients (linear_teukolsky_reduction);
  * the GHP weights and falloffs of every field (weights_and_falloffs).
The constraint point functions are pinned against the reference's own
known-answer tests (tests/test_reconstruction_constraints.cpp:207-283).
A swapped coefficient, a sign, a wrong R power or a wrong operator weight in
any oracle function fails one of these at 1e-13."""
from __future__ import annotations

import json
import math
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "second_order_spec_vectors.json")) as _f:
    SPEC = json.load(_f)

TOL = 1e-13


def c(v):
    return complex(v[0], v[1])


def close(got, want, scale=None, tol=TOL):
    scale = max(abs(want), 1e-300) if scale is None else scale
    return abs(got - want) <= tol * scale


# oracle pair-source slots -> spec symbols (second_order.hpp:74-142 names)
PAIR_FIELDS = {"F1": "F1", "G1": "G1", "Lambda1": "Lambda1", "Pi1": "Pi1", "B1": "B1",
               "C1": "C1", "U1": "U1", "F2": "F2", "G2": "G2", "H2": "H2", "B2": "B2",
               "C2": "C2", "U2": None, "U2s": "U2_s", "C2s": "C2_s", "C1s": "C1_s",
               "B1s": "B1_s", "Pi2s": "Pi2_s", "B2s": "B2_s"}
PAIR_DERIVS = {"d1F2": "Delta_1[F2]", "d3U2": "Delta_3[U2]", "eth2C2": "eth_2[C2]",
               "ethp2C2s": "ethprime_2[C2_s]", "eth1B2": "eth_1[B2]", "d2C2": "Delta_2[C2]",
               "d2G2": "Delta_2[G2]", "eth2G2": "eth_2[G2]", "ethp1F2": "ethprime_1[F2]",
               "d2C2s": "Delta_2[C2_s]", "ethp1B2s": "ethprime_1[B2_s]"}
RECON_INPUTS = {"F": "F", "G": "G", "Lambda": "Lambda", "H": "H", "B": "B", "Pi": "Pi", "C": "C",
                "U": "U", "Pi_s": "Pi_s", "B_s": "B_s", "C_s": "C_s", "eth1F": "eth_1[F]",
                "eth2G": "eth_2[G]", "eth2C": "eth_2[C]", "eth2Pi": "eth_2[Pi]",
                "ethp1B_s": "ethprime_1[B_s]", "ethp2C_s": "ethprime_2[C_s]"}
OUTER_INPUTS = ("Delta_3[D_mt]", "D_mt", "ethprime_3[T_mt]", "T_mt")


def test_every_spec_input_has_an_oracle_slot():
    """The oracle's argument lists cover the spec's symbols exactly: no
    spec input is left unfed, and the one oracle slot the spec never reads
    (U2, OrderedPairFields) is the only extra."""
    spec_pair = set(SPEC["pair_source_inputs"])
    mapped = {v for v in PAIR_FIELDS.values() if v} | set(PAIR_DERIVS.values())
    assert mapped == spec_pair
    assert set(RECON_INPUTS.values()) == set(SPEC["reconstruction_inputs"])
    assert set(O.RECON_INPUTS) == set(RECON_INPUTS)


def _bg(v):
    return {"mu0": c(v["mu0"]), "tau0": c(v["tau0"]), "pi0": c(v["pi0"])}


# ------------------------------------------------------------ background
def test_background_matches_spec():
    for v in SPEC["background"]:
        got = O.background(v["M"], v["a"], v["L"], v["R"], v["ct"], v["st"])
        for k in ("mu0", "tau0", "pi0", "psi20"):
            assert close(got[k], c(v[k]), scale=abs(c(v[k])) + 1e-300), (k, v)


def test_field_weights_match_spec():
    """Spin, boost and falloff of every rescaled field, as the oracle's
    coordinators, Delta_n, eth_n and thorn_n use them (spec
    weights_and_falloffs; ledger spin_boost_weight for D and T; Delta_3 D and
    eth'_3 T in S_over_R3)."""
    want = {"F": (-2, -2, 1), "G": (-1, -1, 2), "H": (0, 0, 3), "Lambda": (-2, -1, 1),
            "Pi": (-1, 0, 2), "B": (-2, 0, 1), "C": (-1, 1, 2), "U": (0, 1, 3)}
    w = SPEC["weights"]
    for name, (s, b, n) in want.items():
        assert (w[name]["spin"], w[name]["boost"], w[name]["falloff"]) == (s, b, n)
    got = O.field_weights()
    for name in want:
        assert got[name] == (w[name]["spin"], w[name]["boost"], w[name]["falloff"]), name
    for name in ("D", "T"):
        assert got[name][:2] == tuple(SPEC["ledger_weights"][name]), name
    # Delta_3 D and eth'_3 T in S / R^3; Delta_n X = R_X with X's falloff
    assert got["D"][2] == 3 and got["T"][2] == 3
    assert "Delta_3[D_mt]" in OUTER_INPUTS and "ethprime_3[T_mt]" in OUTER_INPUTS
    for name, n in SPEC["reconstruction_n"].items():
        assert got[name][2] == n, name


# ------------------------------------------------------------ operators
def test_delta_n_matches_spec():
    for v in SPEC["delta_n"]:
        got = O.delta_n(c(v["X"]), c(v["X_T"]), c(v["X_R"]), v["n"], v["R"], v["M"], v["L"])
        scale = abs(c(v["X"])) + abs(c(v["X_T"])) * 6 + abs(c(v["X_R"]))
        assert close(got, c(v["Delta"]), scale=scale), v


def test_eth_and_ethprime_coordinate_forms_match_spec():
    for v in SPEC["eth"]:
        got = O.eth_point(v["prime"], v["spin"], v["boost"], v["a"], v["L"], v["R"], v["ct"],
                          v["st"], c(v["X"]), c(v["X_T"]), c(v["ang"]))
        scale = (abs(c(v["X"])) + abs(c(v["X_T"])) + abs(c(v["ang"]))) / v["L"] ** 2
        assert close(got, c(v["out"]), scale=scale), v


# ------------------------------------------------------------ pair source
def _pair_args(values, tangents=None):
    tangents = tangents or {}
    f = []
    for slot, name in PAIR_FIELDS.items():
        f.append((c(values[name]) if name in values else 0j,
                  c(tangents[name]) if name in tangents else 0j))
    d = []
    for slot, name in PAIR_DERIVS.items():
        d.append((c(values[name]) if name in values else 0j,
                  c(tangents[name]) if name in tangents else 0j))
    return f, d


def test_pair_source_and_tangents_match_spec():
    for v in SPEC["pair_source"]:
        f, d = _pair_args(v["values"], v["tangents"])
        (D, Dt), (T, Tt) = O.pair_source(v["R"], _bg(v), f, d)
        # terms are O(1) products of O(1) inputs: a scale of the input norm
        scale = sum(abs(c(x)) for x in v["values"].values())
        assert close(D, c(v["D"]), scale=scale), "D"
        assert close(T, c(v["T"]), scale=scale), "T"
        assert close(Dt, c(v["D_t"]), scale=2 * scale), "D_t"
        assert close(Tt, c(v["T_t"]), scale=2 * scale), "T_t"


@pytest.mark.parametrize("row", [r["id"] for r in SPEC["ledger"]])
def test_ledger_term_alone(row):
    """Each ledger row with only its own inputs nonzero: the oracle's D, T
    or S / R^3 equals the ledger's coefficient x product at two radii (the
    background held fixed), so the row's coefficient, sign and explicit R
    power are each fixed."""
    v = next(r for r in SPEC["ledger"] if r["id"] == row)
    bg = _bg(v)
    got = []
    for R, want in zip(v["radii"], v["expected"]):
        if v["target"] == "S":
            x = [c(v["values"][k]) if k in v["values"] else 0j for k in OUTER_INPUTS]
            val = O.outer_source_point(R, bg, *x)
        else:
            f, d = _pair_args(v["values"])
            (D, _), (T, _) = O.pair_source(R, bg, f, d)
            val = D if v["target"] == "D" else T
        scale = max(abs(c(want)), 1e-300)
        assert close(val, c(want), scale=scale), (row, R, val, c(want))
        assert abs(c(want)) > 1e-6, row
        got.append(val)
    if len(v["r_powers"]) == 1:
        # homogeneous in R: the ratio over the two radii is (R2/R1)^p
        p = v["r_powers"][0]
        ratio = (v["radii"][1] / v["radii"][0]) ** p
        assert abs(got[1] - ratio * got[0]) <= 1e-12 * abs(got[1]), row


# ------------------------------------------------------------ outer source
def test_outer_source_and_forcing_match_spec():
    for v in SPEC["outer_source"]:
        x = [c(v["values"][k]) for k in OUTER_INPUTS]
        S = O.outer_source_point(v["R"], _bg(v), *x)
        scale = sum(abs(z) for z in x)
        assert close(S, c(v["S_over_R3"]), scale=scale)
        Fo = O.coordinate_forcing(v["L"], v["a"], v["R"], v["ct"], c(v["S_over_R3"]))
        assert close(Fo, c(v["forcing"]))


# ------------------------------------------------------------ reconstruction
def test_reconstruction_equations_match_spec():
    for v in SPEC["reconstruction"]:
        inputs = {slot: c(v["values"][name]) for slot, name in RECON_INPUTS.items()}
        got = O.recon_rhs(v["R"], _bg(v), inputs)
        scale = sum(abs(z) for z in inputs.values())
        for name, want in v["rhs"].items():
            assert close(got[name], c(want), scale=scale), name
        # the time solve of Delta_n X = R_X with the field's falloff
        w = O.field_weights()
        for name, want in v["time_derivative"].items():
            dt = O.recon_dt(c(v["values"][name]), c(v["radial"][name]), c(v["rhs"][name]),
                            w[name][2], v["R"], v["M"], v["L"])
            assert close(dt, c(want), scale=scale), name


def test_each_reconstruction_input_reaches_its_equations():
    """Zeroing one input changes exactly the equations the spec says read
    it (the spec vectors fix values; this fixes the wiring of every slot)."""
    v = SPEC["reconstruction"][0]
    inputs = {slot: c(v["values"][name]) for slot, name in RECON_INPUTS.items()}
    base = O.recon_rhs(v["R"], _bg(v), inputs)
    readers = SPEC["reconstruction_readers"]
    for slot, name in RECON_INPUTS.items():
        cut = dict(inputs)
        cut[slot] = 0j
        got = O.recon_rhs(v["R"], _bg(v), cut)
        changed = {k for k in base if got[k] != base[k]}
        assert changed == set(readers[name]), (name, changed)


# ------------------------------------------------------------ linear coefficients
def test_linear_reduction_coefficients_match_spec():
    for v in SPEC["linear_coefficients"]:
        got = O.teukolsky_coefficients(v["M"], v["a"], v["L"], v["s"], v["m"], v["R"], v["theta"])
        want = {"time": c(v["C"]), "radial_advection": c(v["K"]),
                "radial_principal": c(v["Hrad"]), "definition": c(v["Gm"]), "q": c(v["q"]),
                "psi": c(v["psi"])}
        for k, w in want.items():
            scale = max(abs(w), 1.0)
            assert close(complex(got[k]), w, scale=scale), (k, v)


# ------------------------------------------------------------ constraints
def test_thorn_matches_reference_known_answer():
    """tests/test_reconstruction_constraints.cpp:207-247: thorn_n at
    (M, a, L) = (1, 0.63, 1.7), R = 0.37, cos = 0.41, n = 2, (s, b) = (-1,-1),
    m = 2 against the reference's expansion."""
    M, a, L, R, ct = 1.0, 0.63, 1.7, 0.37, 0.41
    st = math.sqrt(1.0 - ct * ct)
    bg = O.background(M, a, L, R, ct, st)
    value, dt, dr = 0.7 - 0.2j, -0.4 + 0.9j, 0.3 + 0.5j
    n, s, b, m = 2, -1, -1, 2
    L2 = L * L
    den = L2 * L2 + a * a * R * R * ct * ct
    want = (R * 2.0 * M * (2.0 * M - a * a * R / L2) / den * dt
            - 0.5 * (L2 - 2.0 * M * R + a * a * R * R / L2) / den * (R * dr + n * value)
            + R * 1j * a * m / den * value
            - R * ((s + b) * bg["epsilon0"] + (-s + b) * np.conj(bg["epsilon0"])) * value)
    got = O.thorn(value, dt, dr, n, s, b, m, R, ct, M, a, L, bg["epsilon0"])
    assert abs(got - want) <= 3e-15


def test_independent_constraints_match_reference_known_answer():
    """tests/test_reconstruction_constraints.cpp:249-283: the psi3 and psi2
    Bianchi and h_ll reality residuals at (M, a, L) = (1, 0.52, 1.4),
    R = 0.43, theta = 0.81 with the reference's inputs."""
    M, a, L, R, theta = 1.0, 0.52, 1.4, 0.43, 0.81
    bg = O.background(M, a, L, R, math.cos(theta), math.sin(theta))
    x = dict(F=0.2 - 0.7j, G=-0.1 + 0.4j, H=0.5 + 0.3j, Lambda=-0.6 + 0.2j, Pi=0.8 - 0.1j,
             B=0.3 + 0.9j, C=-0.2 - 0.5j, U=0.7 + 0.1j, U_s=-0.4 + 0.6j, thorn1F=0.11 - 0.13j,
             thorn2G=-0.17 + 0.19j, ethp2G=0.23 + 0.29j, ethp3H=-0.31 + 0.37j)
    terms = O.constraint_terms(R, bg, x)
    R2, R3 = R * R, R ** 3
    want3 = (R * x["ethp2G"] + 4.0 * R2 * bg["pi0"] * x["G"] - x["thorn1F"] + bg["rho0"] * x["F"]
             - 3.0 * R2 * bg["psi20"] * x["Lambda"])
    want2 = (bg["psi20"] * (-3.0 * R3 * bg["mu0"] * x["C"] - 1.5 * R3 * bg["tau0"] * x["B"]
                            - 3.0 * R2 * x["Pi"])
             - R * x["ethp3H"] - 3.0 * R2 * bg["pi0"] * x["H"] + x["thorn2G"]
             - 2.0 * bg["rho0"] * x["G"])
    wanth = x["U"] / bg["mu0"] - x["U_s"] / np.conj(bg["mu0"])
    for (left, right), want in zip(terms, (want3, want2, wanth)):
        assert abs((left - right) - want) <= 3e-15
