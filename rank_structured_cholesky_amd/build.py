"""Build librsc.so (host ordering + gfx950 HIP kernels + C ABI) in-tree.

Usage: python -m rank_structured_cholesky_amd.build  [--jobs N]
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librsc.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HOST_SOURCES = ["host/ordering.cpp", "host/generators.cpp", "host/symbolic.cpp",
                "host/solve_tasks.cpp"]
HIP_SOURCES = ["device/kernels.hip", "device/factor.hip", "device/solve.hip",
               "device/pcg.hip", "capi.cpp", "serve.cpp", "noalloc.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function"]
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wno-unused-command-line-argument"]


def _obj(src):
    return os.path.join(BUILD, src.replace("/", "_") + ".o")


def _compile(src):
    path = os.path.join(CSRC, src)
    out = _obj(src)
    if src in HOST_SOURCES:
        cmd = ["g++"] + COMMON + ["-c", path, "-o", out]
    else:
        cmd = [HIPCC] + COMMON + HIP_FLAGS + ["-x", "hip", "-c", path, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def build(jobs=8, verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = HOST_SOURCES + HIP_SOURCES
    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(srcs)))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "-pthread", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_cli()
    if verbose:
        print(f"built {LIB}")
    return LIB


CLI = os.path.join(HERE, "pcg_solver")


def build_cli():
    """The pcg_solver driver (reference src/pcg_solver.cpp) over librsc.so."""
    inc = os.path.join(os.path.dirname(HERE), "include")
    tmp = CLI + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", inc,
           os.path.join(CSRC, "..", "cli", "pcg_solver.cpp"), "-o", tmp, "-L", HERE, "-lrsc",
           "-lz", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"pcg_solver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, CLI)
    return CLI


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    a = ap.parse_args()
    try:
        build(a.jobs)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
