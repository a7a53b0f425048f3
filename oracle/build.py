"""Build the CPU oracle (test infrastructure only) into oracle/librsc_oracle.so."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "rsc_oracle.cpp")
LIB = os.path.join(HERE, "librsc_oracle.so")


def build(verbose=True):
    tmp = LIB + ".tmp"
    cmd = ["g++", "-O3", "-fPIC", "-shared", "-std=c++17", "-o", tmp, SRC]
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
