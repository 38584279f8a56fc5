"""Writes tests/golden/long_runs.npz: oracle trajectories of the standard
linear workloads run to their stated lengths.

Runs the CPU oracle (test infrastructure) in the build container; the GPU
tests compare the HIP path with these vectors.  Every stored array holds the
m = 2 mode only: the seeds have no sharp partner, so m = -2 stays exactly
zero in the linear evolution.

  case1  Schwarzschild nr 257 D8-4, ell <= 10, CFL 0.02, 2000 steps:
         scri modal coefficients (all ell) every 10 steps, final state of
         P, Q, Psi at every 8th radial row.
  case2  Kerr a = 0.7 nr 1025 D10-5, ell <= 16, CFL 0.5, 20000 steps:
         scri modal coefficients every 10 steps, final state at every 32nd
         radial row.
  case4  a = 0.999 nr 4097 D10-5, ell <= 32, CFL 0.5: the state at 100,
         1000 and 10000 steps at every 128th radial row, scri modal
         coefficients every 100 steps.
  case4x the same run to its stated 100000 steps: the state at 100000
         steps at every 128th radial row, scri modal coefficients every
         1000 steps (written to long_runs_1e5.npz with --out; about four
         hours of six host threads).

Usage: python tests/golden/gen_long_runs.py [--threads N] [--cases case1,case2,case4]
           [--out FILE]
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import helpers as H  # noqa: E402
from teuk_kokkos_amd import workloads as W  # noqa: E402

OUT = os.path.join(HERE, "long_runs.npz")

PLAN = {
    "case1": dict(every=10, stride=8, checkpoints=[2000]),
    "case2": dict(every=10, stride=32, checkpoints=[20000]),
    "case4": dict(every=100, stride=128, checkpoints=[100, 1000, 10000]),
    "case4x": dict(every=1000, stride=128, checkpoints=[100000], workload="case4"),
}


def run(name: str, threads: int, out: dict) -> None:
    cfg = PLAN[name]
    params = getattr(W, cfg.get("workload", name))()
    dt = params["final_time"] / params["steps"]
    pipe = H.oracle_initialized(params, threads=threads)
    modes = pipe.setup.modes
    mi = modes.index(2)
    total = cfg["checkpoints"][-1]
    every, stride = cfg["every"], cfg["stride"]
    scri = [pipe.endpoint_modal(mi)[0]]
    t0 = time.time()
    done = 0
    while done < total:
        pipe.advance(every, dt, time0=done * dt)
        done += every
        scri.append(pipe.endpoint_modal(mi)[0])
        if done % (10 * every) == 0:
            print(f"{name}: {done} of {total} steps, {time.time() - t0:.0f} s", flush=True)
        if done in cfg["checkpoints"]:
            state = pipe.get_state()
            out[f"{name}_state_{done}"] = state[mi, :3, ::stride, :].copy()
            print(f"{name}: {done} steps, {time.time() - t0:.0f} s", flush=True)
    out[f"{name}_scri"] = np.array(scri)
    out[f"{name}_meta"] = np.array([dt, every, stride, total], dtype=np.float64)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--cases", default="case1,case2,case4")
    ap.add_argument("--out", default=OUT)
    args = ap.parse_args()
    out = dict(np.load(args.out)) if os.path.exists(args.out) else {}
    for name in args.cases.split(","):
        run(name, args.threads, out)
        np.savez_compressed(args.out, **out)
    print(f"wrote {args.out}")


if __name__ == "__main__":
    main()
