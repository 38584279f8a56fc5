#!/usr/bin/env python3
"""Throughput of the fused Teukolsky RK4 stage kernels on one MI355X.

Default workload: the batched sweep of 512 independent linear Kerr
evolutions (spin U[0,0.99], m U{-4..4}, ellmax U{8,12,16,24}, nr
U{257,513,1025,2049}, D10-5, CFL 0.5), all advancing together with one
launch per RK stage.  A step is one classical RK4 step of every instance;
the metric counts grid points as the reference does
(include/teuk/solver_driver.hpp:323-341): steps * modes * nr * ntheta.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

METRIC = ("grid-point-updates/s, mixed vs FP64 on GPU, at ≤1e-9 waveform err; "
          "% HBM peak")
HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per retained modal coefficient per RK4 step: three
# complex128 fields, 16 state vectors (stage reads/writes of y, stage value
# and accumulator); one step is four launches.
BYTES_PER_MODAL_STEP = 3 * 16 * 16
# Second-order workloads evolve all 13 nodal fields: 13 complex128 values
# per grid point and 16 state vectors per RK4 step.
BYTES_PER_GRID_STEP_13 = 13 * 16 * 16


def workload(name: str):
    from teuk_kokkos_amd import workloads as W
    if name == "sweep512":
        return W.sweep(512), "kerr_sweep_512: a~U[0,0.99], m~U{-4..4}, ellmax~U{8,12,16,24}, " \
                             "nr~U{257,513,1025,2049}, d10-5, cfl 0.5"
    if name == "case2":
        return [W.case2()], "kerr_a0.7_ringdown: nr=1025 d10-5 ellmax=16 ntheta=25 cfl 0.5"
    if name == "case1":
        return [W.case1()], "schwarzschild: nr=257 d8-4 ellmax=10 ntheta=16 cfl 0.02"
    if name == "case4":
        return [W.case4()], "near_extremal_a0.999: nr=4097 d10-5 ellmax=32 ntheta=49 cfl 0.5"
    if name == "so64":
        return W.second_order_batch(64), \
            "second_order_64: a~U[0,0.95], L=1.5, parents m=+-2 ell<=8, targets {-4,0,4} " \
            "ell<=12, ntheta=15, nr=1025 d10-5, 13 live fields, unrestricted source from 10 dt"
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(cases, target_seconds: float):
    """Oracle (nodal FP64 C++/OpenMP restatement) on a bounded sample."""
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests"))
    import helpers as H
    threads = min(16, os.cpu_count() or 1)
    sample = cases[:8]
    pipes = [H.oracle_initialized(c, threads=threads) for c in sample]
    dts = [c["final_time"] / c["steps"] for c in sample]
    pts = sum(len(set(c["first_order_modes"]) | set(c["second_order_modes"])) * c["nr"] *
              c["ntheta"] for c in sample)
    t0 = time.perf_counter()
    for p, dt in zip(pipes, dts):
        p.advance(1, dt)
    one = time.perf_counter() - t0
    steps = max(1, int(target_seconds / max(one, 1e-6)))
    t0 = time.perf_counter()
    for p, dt in zip(pipes, dts):
        p.advance(steps, dt)
    sec = time.perf_counter() - t0
    return {"value": pts * steps / sec, "unit": "grid-point-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"first {len(sample)} instances of the workload, {steps} RK4 steps each, "
                      f"{sec:.1f} s"}


def _under_profiler() -> bool:
    """True inside a rocprofv3 run (its tool library is preloaded)."""
    return ("rocprofiler" in os.environ.get("LD_PRELOAD", "")
            or any(k.startswith(("ROCPROF", "ROCP_")) for k in os.environ))


def _timed_kernel(name: str, second_order: bool) -> bool:
    if second_order:
        return "so_" in name and "so_background" not in name and "so_reduce" not in name
    return "teuk_stage_kernel" in name


def measure_traffic(workload_name: str, second_order: bool, algorithmic_stage_bytes: float):
    """HBM bytes per RK stage of the timed kernels: two child runs of one
    step of the same workload under `rocprofv3 --pmc` (FETCH_SIZE, then
    WRITE_SIZE, KiB per dispatch), each scaled by what the same counter reads
    for a 512 MiB device-to-device copy made in the same run.  None when
    rocprofv3 is absent or a pass yields no usable counters."""
    exe = shutil.which("rocprofv3")
    if exe is None:
        return None
    copy_bytes = 512 << 20
    total = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        with tempfile.TemporaryDirectory() as d:
            cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "run", "--",
                   sys.executable, os.path.abspath(__file__), "--workload", workload_name,
                   "--steps", "1", "--warmup", "0", "--no-cpu", "--no-traffic", "--eager",
                   "--calibrate-copy-mib", str(copy_bytes >> 20)]
            try:
                # the profiler may exit non-zero after a complete run; the
                # counter file decides
                subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                               timeout=240, check=False)
            except subprocess.TimeoutExpired:
                return None
            files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
            if not files:
                return None
            per = {}
            for row in csv.DictReader(open(files[0])):
                key = (row["Kernel_Name"], row["Dispatch_Id"])
                per[key] = per.get(key, 0.0) + float(row["Counter_Value"])
        copies = [v for (k, _), v in per.items() if "copyBuffer" in k]
        if not copies or max(copies) <= 0:
            return None
        scale = copy_bytes / max(copies)  # bytes per counter unit
        total[counter] = sum(v for (k, _), v in per.items() if _timed_kernel(k, second_order))
        total[counter] *= scale / 4.0  # one step = four RK stages
        total[counter + "_scale"] = scale / 1024.0
    stage = total["FETCH_SIZE"] + total["WRITE_SIZE"]
    return {"bytes_per_stage": stage, "read": total["FETCH_SIZE"], "written": total["WRITE_SIZE"],
            "over_algorithmic": stage / algorithmic_stage_bytes,
            "calibration": {"fetch": total["FETCH_SIZE_scale"],
                            "write": total["WRITE_SIZE_scale"]}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="sweep512")
    ap.add_argument("--eager", action="store_true",
                    help="launch every stage kernel directly instead of replaying the step as "
                         "a captured hipGraph (the default for first-order plans)")
    ap.add_argument("--graph", action="store_true", help=argparse.SUPPRESS)  # the default
    ap.add_argument("--lds-budget", type=int, default=0)
    ap.add_argument("--tile-rows", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 --pmc child runs that measure HBM traffic")
    ap.add_argument("--autotune", action="store_true",
                    help="sweep LDS budget x tile rows before timing (cached per workload)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first run one device-to-device copy of this size (a counter "
                         "calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (got --gpus "
              f"{args.gpus})", file=sys.stderr)
        return 2

    import teuk_kokkos_amd as T
    from teuk_kokkos_amd import workloads as W

    if args.calibrate_copy_mib:
        import ctypes
        import numpy as np
        n = args.calibrate_copy_mib << 20
        src = T.DeviceArray(np.zeros(n // 16, dtype=np.complex128))
        dst = T.DeviceArray(np.zeros(n // 16, dtype=np.complex128))
        T.lib().teuk_memcpy(ctypes.c_void_p(dst.ptr), ctypes.c_void_p(src.ptr), n, 2)
        print(f"[bench] calibration copy {n} bytes", file=sys.stderr, flush=True)
        src.free()
        dst.free()

    cases, desc = workload(args.workload)
    print(f"[bench] building plan: {len(cases)} instances", file=sys.stderr, flush=True)
    plan = T.Plan(W.texts(cases))
    if args.lds_budget:
        plan.set_lds_budget(args.lds_budget)
    if args.tile_rows:
        plan.set_tile_rows(args.tile_rows)
    if args.autotune:
        from teuk_kokkos_amd import autotune
        plan.initialize()
        best = autotune.tune(plan, W.texts(cases), log=lambda m: print(m, file=sys.stderr,
                                                                      flush=True))
        print(f"[bench] autotuned: {best}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    plan.initialize()
    print(f"[bench] initial data {time.perf_counter() - t0:.1f} s; tiles {plan.tiles()}",
          file=sys.stderr, flush=True)
    # first-order plans replay each RK4 step as one captured hipGraph; the
    # second-order graph launches its kernels directly (its latch decisions
    # are taken on the host between steps)
    use_graph = not args.eager and plan.modal_points > 0
    plan.set_graph(use_graph)
    plan.advance(args.warmup)
    plan.sync()

    # HIP events around every step (around every replay in graph mode) and,
    # in eager mode, around every stage launch
    plan.set_kernel_timing(True)
    plan.sync()
    t0 = time.perf_counter()
    plan.advance(args.steps)
    plan.sync()
    seconds = time.perf_counter() - t0
    kernel_ms, launches = plan.kernel_timing()
    step_ms, timed_steps = plan.step_timing()
    groups = plan.groups()
    plan.set_kernel_timing(False)

    grid = plan.grid_points
    modal = plan.modal_points
    value = grid * args.steps / seconds
    roofline = None
    second_order = modal == 0
    if timed_steps and second_order:
        # The 13-field graph: seven RHS launches and one RK combination per
        # stage; algorithmic bytes are the 13-field state vectors only.
        stage_s = step_ms / (4 * timed_steps) * 1e-3
        achieved = BYTES_PER_GRID_STEP_13 / 4.0 * grid / stage_s / 1e9
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                    "kernel": "second-order stage graph (10 launches per RK stage)",
                    "avg_stage_ms": stage_s * 1e3}
    elif timed_steps:
        # One RK stage = one launch per execution group, the groups running
        # concurrently on streams of their own: a stage's algorithmic bytes
        # over a quarter of the GPU time of a step (HIP events on the plan's
        # stream at the fork and the join).
        stage_s = step_ms / (4 * timed_steps) * 1e-3
        achieved = BYTES_PER_MODAL_STEP / 4.0 * modal / stage_s / 1e9
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                    "kernel": "teuk_stage_kernel (4 RK stages)",
                    "avg_stage_ms": stage_s * 1e3,
                    "launches_per_stage": (float(sum(1 for g in groups if g["tiles"] > 0))
                                           if use_graph else launches / (4 * args.steps)),
                    "sum_of_launches_per_stage_ms": (None if use_graph
                                                     else kernel_ms / (4 * args.steps)),
                    "groups": [{k: (None if use_graph else round(v / (4 * args.steps), 4))
                                if k == "kernel_ms" else v
                                for k, v in g.items()} for g in groups]}
    out = {
        "metric": METRIC, "value": value, "unit": "grid-point-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": seconds / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": desc, "instances": len(cases), "grid_points_per_step": grid,
                   "modal_points_per_step": modal, "precision": "fp64",
                   "graph": bool(use_graph)},
        "modal_updates_per_s": modal * args.steps / seconds,
        "roofline": roofline,
        "cpu_baseline": None,
    }
    plan.close()
    if roofline is not None and not args.no_traffic and not _under_profiler():
        algorithmic = (BYTES_PER_GRID_STEP_13 * grid if second_order
                       else BYTES_PER_MODAL_STEP * modal) / 4.0
        tr = measure_traffic(args.workload, second_order, algorithmic)
        if tr is not None:
            # HBM bytes of one RK stage (all its launches), as `achieved`
            roofline["traffic"] = tr["bytes_per_stage"]
            roofline["traffic_detail"] = tr
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cases, args.cpu_seconds)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
