"""Checkpoint format v5 (pipeline_checkpoint.hpp:40-943): the metadata text
and state bytes of the reference's format, restart equal to an uninterrupted
run (bit for bit for nodal second-order plans, to rounding for modal linear
plans), and the load-side refusals (existing destination, checksum,
size, key set, registry, configuration) leaving the state untouched.  All
calls go through the C ABI; the solver test runs the executable."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

import helpers as H
import teuk_kokkos_amd as T
from teuk_kokkos_amd import workloads as W

pytestmark = pytest.mark.gpu

SOLVER = os.path.join(os.path.dirname(T.__file__), "teuk_solver")

CASES = {
    "linear": H.small_case(spin=0.7, nr=49, radial_discretization="d10-5", ellmax_first=6,
                           ellmax_second=6, seeds=[(2, 2, 1e-3), (3, 2, 2e-4 - 5e-4j)],
                           steps=8, final_time=8e-3),
    "second": H.small_case(spin=0.6, compactification_length=1.5, nr=41,
                           radial_discretization="d10-5", ellmax_first=4, ellmax_second=6,
                           first_order_modes=[-2, 2], second_order_modes=[-4, 0, 4],
                           initial_data__add_sharp_partner=True, seeds=[(2, 2, 1e-3)],
                           second_order__enabled=True, second_order__source_mode="unrestricted",
                           second_order__source_start_time=2.5e-3, steps=8, final_time=8e-3),
}

KEYS = ["format", "version", "ordering", "complex_storage", "byte_order", "state_file", "modes",
        "fields", "radial", "theta", "m_values", "target_values", "parent_values",
        "radial_coordinates", "theta_coordinates", "mass", "spin", "compactification_length",
        "ell_max_first", "ell_max_second", "theta_nodes", "reduction_damping", "dissipation",
        "reduction", "radial_discretization", "time_step", "source_mode", "source_start_time",
        "source_normalized_constraint_tolerance", "source_required_consecutive_passes",
        "source_active", "source_activation_time", "source_consecutive_passes",
        "source_last_eligibility_time", "time", "step", "state_checksum_fnv1a64"]


def fnv1a64(data: bytes) -> int:
    h = 14695981039346656037
    for b in data:
        h ^= b
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def test_fnv1a64_known_vectors():
    # published FNV-1a 64-bit test vectors
    assert fnv1a64(b"") == 0xCBF29CE484222325
    assert fnv1a64(b"a") == 0xAF63DC4C8601EC8C
    assert fnv1a64(b"foobar") == 0x85944171F73967E8


def metadata(path):
    lines = (path / "metadata.txt").read_text().splitlines()
    return [ln.split("=", 1)[0] for ln in lines], dict(ln.split("=", 1) for ln in lines)


@pytest.mark.parametrize("name", sorted(CASES))
def test_checkpoint_format_and_bitwise_restart(name, tmp_path):
    params = CASES[name]
    text = W.config_text(**dict(params))
    ck = tmp_path / "ck"
    print(f"[gpu] checkpoint {name}", flush=True)
    with T.Plan(text) as a:
        a.initialize()
        a.advance(4)
        a.write_checkpoint(str(ck))
        at4 = a.export_state()
        act4 = a.source_activation()
        a.advance(4)
        final = a.export_state()
        t_final, s_final = a.progress()
    order, md = metadata(ck)
    assert order == KEYS
    assert md["format"] == "teuk-kokkos-pipeline-checkpoint" and md["version"] == "5"
    assert md["step"] == "4" and float(md["time"]) == pytest.approx(4e-3, abs=1e-15)
    raw = (ck / "state.bin").read_bytes()
    assert len(raw) == at4.size * 16
    assert int(md["state_checksum_fnv1a64"]) == fnv1a64(raw)
    assert np.array_equal(np.frombuffer(raw, dtype="<c16").reshape(at4.shape), at4)
    M, _, N, T_ = at4.shape
    assert (int(md["modes"]), int(md["fields"]), int(md["radial"]), int(md["theta"])) == (M, 13, N, T_)
    R = [float(v) for v in md["radial_coordinates"].split(",")]
    assert R[0] == 0.0 and len(R) == N
    assert md["source_mode"] == ("unrestricted" if name == "second" else "disabled")
    assert int(md["source_active"]) == int(act4.active)
    with T.Plan(text) as b:
        b.load_checkpoint(str(ck))
        t, s = b.progress()
        assert s == 4 and t == float(md["time"])
        assert b.source_activation().active == act4.active
        if name == "second":
            # nodal state: the restart is the uninterrupted run, bit for bit
            assert np.array_equal(b.export_state(), at4)
            b.advance(4)
            assert np.array_equal(b.export_state(), final)
        else:
            # modal state: the coefficients are re-derived from the stored
            # nodal values, which agree with them to rounding
            assert H.rel_err(b.export_state(), at4) <= 1e-14
            b.advance(4)
            assert H.rel_err(b.export_state(), final) <= 1e-13
        assert b.progress() == (t_final, s_final)


def test_checkpoint_refusals_leave_state_untouched(tmp_path):
    params = CASES["second"]
    text = W.config_text(**dict(params))
    ck = tmp_path / "ck"
    with T.Plan(text) as a:
        a.initialize()
        a.advance(2)
        a.write_checkpoint(str(ck))
        with pytest.raises(T.TeukError, match="already exists"):
            a.write_checkpoint(str(ck))
    bad_sum = tmp_path / "bad_sum"
    bad_size = tmp_path / "bad_size"
    bad_key = tmp_path / "bad_key"
    for d in (bad_sum, bad_size, bad_key):
        d.mkdir()
        (d / "metadata.txt").write_bytes((ck / "metadata.txt").read_bytes())
        (d / "state.bin").write_bytes((ck / "state.bin").read_bytes())
    raw = bytearray((bad_sum / "state.bin").read_bytes())
    raw[100] ^= 1
    (bad_sum / "state.bin").write_bytes(bytes(raw))
    (bad_size / "state.bin").write_bytes((ck / "state.bin").read_bytes()[:-16])
    with open(bad_key / "metadata.txt", "a") as f:
        f.write("extra=1\n")
    other = W.config_text(**dict(params, spin=0.61))
    with T.Plan(text) as b:
        b.initialize()
        before = b.export_state()
        for d, msg in ((bad_sum, "checksum"), (bad_size, "size"), (bad_key, "key set")):
            with pytest.raises(T.TeukError, match=msg):
                b.load_checkpoint(str(d))
            assert np.array_equal(b.export_state(), before)
    with T.Plan(other) as c:
        c.initialize()
        with pytest.raises(T.TeukError, match="configuration"):
            c.load_checkpoint(str(ck))
    wrong_modes = W.config_text(**dict(params, second_order_modes=[-4, 4],
                                       second_order__allow_truncated_daughter_modes=True))
    with T.Plan(wrong_modes) as d:
        d.initialize()
        with pytest.raises(T.TeukError, match="registry"):
            d.load_checkpoint(str(ck))


def test_solver_writes_and_resumes_from_checkpoints(tmp_path):
    params = CASES["second"]
    cfg = tmp_path / "run.cfg"
    cfg.write_text(W.config_text(**dict(params)))
    out = tmp_path / "first"
    res = subprocess.run([SOLVER, "--config", str(cfg), f"output.directory={out}",
                          "output.diagnostic_every=4", "output.checkpoint_every=4"],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    assert sorted(p.name for p in out.iterdir() if p.name.startswith("checkpoint_")) == \
        ["checkpoint_00000004", "checkpoint_00000008"]
    # resume from step 4 for the remaining 4 steps: the same step-8 state
    resumed = tmp_path / "resumed"
    res = subprocess.run([SOLVER, "--config", str(cfg), f"output.directory={resumed}",
                          "initial_data.type=checkpoint",
                          f"initial_data.checkpoint_directory={out / 'checkpoint_00000004'}",
                          "steps=4", "final_time=0.004", "output.diagnostic_every=4",
                          "output.checkpoint_every=4"],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    a = (out / "checkpoint_00000008" / "state.bin").read_bytes()
    b = (resumed / "checkpoint_00000008" / "state.bin").read_bytes()
    assert a == b
    first = (out / "diagnostics.csv").read_text().splitlines()
    second = (resumed / "diagnostics.csv").read_text().splitlines()
    assert second[1].split(",")[0] == "4" and second[-1] == first[-1]
