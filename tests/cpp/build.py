"""Builds the C++ test programs of this directory against the in-tree
librsc.so: test_cpp_api (the C++ mirror exercised through the C ABI) and
emulate_apply (CPU emulation of the apply's task schedules).  Run after
rank_structured_cholesky_amd/build.py."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
PKG = os.path.join(ROOT, "rank_structured_cholesky_amd")
OUT = os.path.join(HERE, "test_cpp_api")


def _one(src, out, extra=()):
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", *extra, "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, src), "-o", out + ".tmp", "-L", PKG, "-lrsc",
           "-Wl,-rpath,$ORIGIN/../../rank_structured_cholesky_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(out + ".tmp", out)
    return out


def build():
    _one("emulate_apply.cpp", os.path.join(HERE, "emulate_apply"), ("-pthread",))
    return _one("test_cpp_api.cpp", OUT)


if __name__ == "__main__":
    try:
        print(build())
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
