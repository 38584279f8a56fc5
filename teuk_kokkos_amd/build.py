"""Build libteuk_hip.so (HIP kernels for gfx950 + host C++ + C ABI) in-tree.

Usage: python -m teuk_kokkos_amd.build   (or teuk_kokkos_amd.build.build())
Also builds the teuk_solver command-line driver next to the library.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libteuk_hip.so")
SOLVER = os.path.join(HERE, "teuk_solver")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++20", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter"]
# Packed FP32 instructions stay off in every device pass.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]

LIB_SOURCES = ["config.cpp", "host_math.cpp", "capi.cpp", "plan.hip", "stage_kernels.hip",
               "so_plan.hip", "so_kernels.hip", "init_kernels.hip", "import_kernels.hip",
               "diag_kernels.hip", "diag_plan.hip", "checkpoint.cpp", "scratch_rhs.hip"]
# The stage kernels, one object per (radial scheme, ell-tile class) so the
# instantiations compile in parallel.
STAGE_INSTANCES = [(scheme, ub) for scheme in (0, 1, 2) for ub in (1, 2, 4)]


def _compile(job) -> str:
    src, defines = job if isinstance(job, tuple) else (job, ())
    stem = os.path.splitext(src)[0] + "".join(f"_{d.split('=')[1]}" for d in defines)
    obj = os.path.join(BUILD, stem + ".o")
    path = os.path.join(CSRC, src)
    flags = [f"-D{d}" for d in defines]
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip", *DEVICE, *COMMON, *flags, "-c", path, "-o", obj]
    else:
        cmd = [HIPCC, *COMMON, *flags, "-c", path, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        jobs = [("stage_instance.hip", (f"TEUK_SCHEME={sc}", f"TEUK_UB={ub}"))
                for sc, ub in STAGE_INSTANCES] + LIB_SOURCES
        objs = list(ex.map(_compile, jobs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    solver_src = os.path.join(CSRC, "teuk_solver.cpp")
    if os.path.exists(solver_src):
        cmd = [HIPCC, *COMMON, solver_src, "-o", SOLVER + ".tmp", f"-L{HERE}", "-lteuk_hip",
               "-Wl,-rpath,$ORIGIN"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"solver build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
        os.replace(SOLVER + ".tmp", SOLVER)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
