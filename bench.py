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
# complex128 fields, 12 state vectors (stage 1 reads y and writes Y2; stage 2
# reads Y2 and y, writes Y3 and the partial sum C; stage 3 reads Y3 and y,
# writes Y4; stage 4 reads Y4 and C, writes y); one step is four launches.
BYTES_PER_MODAL_STEP = 3 * 16 * 12
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


def _host_cpu():
    """Host CPU model (as lscpu prints it) and the logical CPUs this process
    may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return model, usable


def cpu_baseline(cases, target_seconds: float):
    """Oracle (nodal FP64 C++/OpenMP restatement) on a bounded sample: timed
    with every host core this job may use (OMP_NUM_THREADS when the host sets
    it, else the CPUs in the affinity mask) and with one thread."""
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests"))
    import helpers as H
    model, usable = _host_cpu()
    env_threads = os.environ.get("OMP_NUM_THREADS", "")
    threads = int(env_threads) if env_threads.isdigit() and int(env_threads) > 0 else usable
    sample = cases[:8]
    dts = [c["final_time"] / c["steps"] for c in sample]
    pts = sum(len(set(c["first_order_modes"]) | set(c["second_order_modes"])) * c["nr"] *
              c["ntheta"] for c in sample)

    def rate(n_threads, seconds):
        pipes = [H.oracle_initialized(c, threads=n_threads) for c in sample]
        t0 = time.perf_counter()
        for p, dt in zip(pipes, dts):
            p.advance(1, dt)
        one = time.perf_counter() - t0
        steps = max(1, int(seconds / max(one, 1e-6)))
        t0 = time.perf_counter()
        for p, dt in zip(pipes, dts):
            p.advance(steps, dt)
        sec = time.perf_counter() - t0
        return pts * steps / sec, steps, sec

    value, steps, sec = rate(threads, target_seconds)
    single, steps1, sec1 = rate(1, target_seconds / 2)
    return {"value": value, "unit": "grid-point-updates/s", "cores": threads, "kind": "port",
            "sample": f"first {len(sample)} instances of the workload, {steps} RK4 steps each, "
                      f"{sec:.1f} s on {threads} threads",
            "single_thread": {"value": single, "cores": 1,
                              "sample": f"same instances, {steps1} RK4 steps each, "
                                        f"{sec1:.1f} s"},
            "host": {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": usable,
                     "omp_num_threads": env_threads or None}}


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
                   "--no-compare",
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


def summarize_trace(trace_dir: str, bench_line: str) -> str:
    """Markdown summary of a `rocprofv3 --kernel-trace --stats --output-format
    csv` run of this script: per-kernel statistics, and over the timed steps
    (the last steps x 4 x groups stage-kernel dispatches) the per-step kernel
    time (sum of durations), busy time (union of the dispatch intervals: the
    groups overlap) and span, set against the bench line of the same run."""
    line = json.loads(open(bench_line).read().strip().splitlines()[-1])
    traces = glob.glob(os.path.join(trace_dir, "**", "*kernel_trace.csv"), recursive=True)
    if not traces:
        raise SystemExit(f"no kernel_trace.csv under {trace_dir}")
    rows = []
    for r in csv.DictReader(open(traces[0])):
        rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]))
    rows.sort()
    stats = {}
    for t0, t1, name in rows:
        st = stats.setdefault(name, [0, 0.0, float("inf"), 0.0])
        d = (t1 - t0) / 1e3
        st[0] += 1
        st[1] += d
        st[2] = min(st[2], d)
        st[3] = max(st[3], d)
    total = sum(v[1] for v in stats.values())
    out = ["| kernel | calls | avg us | min us | max us | % |", "|---|---|---|---|---|---|"]
    for name, (n, tot, lo, hi) in sorted(stats.items(), key=lambda kv: -kv[1][1]):
        out.append(f"| `{name}` | {n} | {tot / n:.1f} | {lo:.1f} | {hi:.1f} | "
                   f"{100 * tot / total:.1f} |")
    steps = int(line["steps"])
    rl = line.get("roofline") or {}
    second_order = line["config"].get("modal_points_per_step", 0) == 0
    groups = [g for g in rl.get("groups", []) if g.get("tiles", 0) > 0]
    # launches per RK stage: one per execution group, or the second-order
    # graph's ten (three linear triples, two radial D, two reconstructions,
    # source, source projection, RK combination)
    per_stage = 10 if second_order else max(1, len(groups))
    stage = [r for r in rows if _timed_kernel(r[2], second_order)]
    timed = stage[-steps * 4 * per_stage:]
    busy, cur0, cur1 = 0, None, None
    for t0, t1, _ in timed:
        if cur1 is None or t0 > cur1:
            if cur1 is not None:
                busy += cur1 - cur0
            cur0, cur1 = t0, t1
        else:
            cur1 = max(cur1, t1)
    if cur1 is not None:
        busy += cur1 - cur0
    sum_ms = sum(t1 - t0 for t0, t1, _ in timed) / 1e6 / steps
    busy_ms = busy / 1e6 / steps
    span_ms = (timed[-1][1] - timed[0][0]) / 1e6 / steps if timed else float("nan")
    modal = line["config"].get("modal_points_per_step", 0)
    grid = line["config"].get("grid_points_per_step", 0)
    summary = [
        f"Timed region: the last {len(timed)} stage-kernel dispatches ({steps} steps x 4 RK "
        f"stages x {per_stage} " + ("launches" if second_order else "execution groups") + ").",
        "",
        "| per RK4 step | ms |",
        "|---|---|",
        f"| bench line `ms_per_step` (host clock, this profiled run) | "
        f"{line['ms_per_step']:.4f} |",
        f"| bench line step time from HIP events (4 x `avg_stage_ms`) | "
        f"{4 * rl.get('avg_stage_ms', float('nan')):.4f} |",
        f"| trace: span from the first dispatch's start to the last's end | {span_ms:.4f} |",
        f"| trace: busy time (union of dispatch intervals) | {busy_ms:.4f} |",
        f"| trace: sum of dispatch durations (groups overlap) | {sum_ms:.4f} |",
    ]
    per_unit, units, what = ((BYTES_PER_GRID_STEP_13, grid, "grid points (13 fields)")
                             if second_order else
                             (BYTES_PER_MODAL_STEP, modal, "modal coefficients"))
    if units:
        ach = per_unit * units / (busy_ms * 1e-3) / 1e9
        summary += ["",
                    f"Algorithmic bytes per step: {per_unit} B x {units:,.0f} {what} = "
                    f"{per_unit * units / 1e9:.3f} GB.  From the "
                    f"trace's busy time: {ach:.0f} GB/s = {ach / HBM_PEAK_GBS:.3f} of "
                    f"{HBM_PEAK_GBS:.0f} GB/s; the bench line reports "
                    f"{rl.get('achieved', float('nan')):.0f} GB/s = "
                    f"{rl.get('frac', float('nan')):.3f}."]
    return "\n".join(summary + ["", "## Kernel statistics (kernel_trace.csv)", ""] + out) + "\n"


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
    ap.add_argument("--fp64-gram", action="store_true",
                    help="time the FP32 Gram correction (as accurate as the FP64 product) "
                         "instead of the mixed-precision BF16 one on the ell-24 bands")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the timed run with the FP32 Gram correction after the headline")
    ap.add_argument("--lds-budget", type=int, default=0)
    ap.add_argument("--tile-rows", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 --pmc child runs that measure HBM traffic")
    ap.add_argument("--autotune", action="store_true",
                    help="tune every execution group's tile shape before timing and store the "
                         "pick in the tuning cache (by default a cached pick is applied)")
    ap.add_argument("--no-tuning", action="store_true",
                    help="run the default tile shapes, ignoring the tuning cache")
    ap.add_argument("--tuning-cache", default=None,
                    help="tuning cache file (default: the package's tuning_cache.json)")
    ap.add_argument("--summarize-trace", default=None, metavar="DIR",
                    help="print a markdown summary of a rocprofv3 kernel trace of this script "
                         "(with --bench-line, the JSON line of the same run) and exit")
    ap.add_argument("--bench-line", default=None)
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first run one device-to-device copy of this size (a counter "
                         "calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.summarize_trace:
        print(summarize_trace(args.summarize_trace, args.bench_line), end="")
        return 0
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
    # mixed precision (first-order plans): the ell-24 bands' Gram correction
    # on BF16 MFMA, set before tuning so the tuner times the headline form;
    # the FP32 form is timed after the headline run for the mixed-vs-FP64
    # ratio
    mixed = plan.modal_points > 0 and not args.fp64_gram
    plan.set_mixed_gram(mixed)
    tuning_source = "default"
    if plan.modal_points > 0 and not args.no_tuning and not (args.lds_budget or args.tile_rows):
        from teuk_kokkos_amd import autotune
        cache_path = args.tuning_cache or autotune.DEFAULT_CACHE
        if args.autotune:
            plan.initialize()
            best = autotune.tune(plan, W.texts(cases), cache_path=cache_path, use_cache=False,
                                 log=lambda m: print(m, file=sys.stderr, flush=True))
            tuning_source = "autotune"
            print(f"[bench] autotuned: {best}", file=sys.stderr, flush=True)
        elif autotune.apply_cached(plan, W.texts(cases), cache_path) is not None:
            tuning_source = "cache"
    t0 = time.perf_counter()
    plan.initialize()
    print(f"[bench] initial data {time.perf_counter() - t0:.1f} s; tiles {plan.tiles()}",
          file=sys.stderr, flush=True)
    # each RK4 step replayed as one captured hipGraph (second-order plans
    # launch directly only the steps that split at the source start or need
    # a constraint evaluation for the latch)
    use_graph = not args.eager
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
    fp64_gram = None
    # (not under a profiler: its trace summary takes the last steps' dispatches
    # as the headline run's)
    if mixed and not args.no_compare and not _under_profiler():
        # the same steps with the FP32 Gram correction (graph replays as above)
        plan.set_mixed_gram(False)
        plan.advance(max(1, args.warmup))
        plan.sync()
        t1 = time.perf_counter()
        plan.advance(args.steps)
        plan.sync()
        s64 = time.perf_counter() - t1
        fp64_gram = {"ms_per_step": s64 / args.steps * 1e3, "value": grid * args.steps / s64,
                     "mixed_over_fp64": s64 / seconds}
    roofline = None
    second_order = modal == 0
    if timed_steps and second_order:
        # The 13-field graph: seven RHS launches and one RK combination per
        # stage; algorithmic bytes are the 13-field state vectors only.
        stage_s = step_ms / (4 * timed_steps) * 1e-3
        achieved = BYTES_PER_GRID_STEP_13 / 4.0 * grid / stage_s / 1e9
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                    "algorithmic_bytes": {"per_unit_and_step": BYTES_PER_GRID_STEP_13,
                                          "unit": "grid point (13 fields)", "units": grid},
                    "kernel": "second-order stage graph (10 launches per RK stage)",
                    "avg_stage_ms": stage_s * 1e3,
                    "graph_replays": plan.graph_replays()}
    elif timed_steps:
        # One RK stage = one launch per execution group, the groups running
        # concurrently on streams of their own: a stage's algorithmic bytes
        # over a quarter of the GPU time of a step (HIP events on the plan's
        # stream at the fork and the join).
        stage_s = step_ms / (4 * timed_steps) * 1e-3
        achieved = BYTES_PER_MODAL_STEP / 4.0 * modal / stage_s / 1e9
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                    "algorithmic_bytes": {"per_unit_and_step": BYTES_PER_MODAL_STEP,
                                          "unit": "retained modal coefficient (P, Q, Psi)",
                                          "units": modal},
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
                   "modal_points_per_step": modal,
                   "precision": ("fp64 state; ell-24 Gram correction on bf16 MFMA (mixed)"
                                 if mixed else "fp64"),
                   "graph": bool(use_graph),
                   "tuning": {"source": tuning_source,
                              "shapes": {f"{g['scheme']}/{g['ell_tiles']}/{g['variant']}":
                                         {k: g[k] for k in ("lds_budget", "tile_rows",
                                                            "tiles", "lds_bytes")}
                                         for g in groups}}},
        "modal_updates_per_s": modal * args.steps / seconds,
        "fp64_gram": fp64_gram,
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
