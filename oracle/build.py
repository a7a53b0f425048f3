"""Build the CPU oracle (test infrastructure only) into oracle/librsc_oracle.so."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "rsc_oracle.cpp")
LIB = os.path.join(HERE, "librsc_oracle.so")


def build_native(out_dir, verbose=False):
    """The same source built for the machine it runs on (-march=native),
    into out_dir: the CPU baseline's build.  Returns the library path, or
    None when the compiler is missing or fails."""
    os.makedirs(out_dir, exist_ok=True)
    lib = os.path.join(out_dir, "librsc_oracle_native.so")
    cmd = ["g++", "-O3", "-march=native", "-fopenmp", "-fPIC", "-shared", "-std=c++17",
           "-o", lib, SRC]
    try:
        r = subprocess.run(cmd, capture_output=True, text=True)
    except OSError:
        return None
    if r.returncode != 0:
        if verbose:
            print(r.stderr, file=sys.stderr)
        return None
    return lib


def build(verbose=True):
    tmp = LIB + ".tmp"
    # x86-64-v3 (AVX2 + FMA): built here, run on the GPU hosts as well
    cmd = ["g++", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-std=c++17",
           "-o", tmp, SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
