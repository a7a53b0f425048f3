"""The header-only C++ mirror (include/rschol_amd.hpp): error paths on the
CPU, factor + PCG + refactor on the GPU (tests/cpp/test_cpp_api.cpp)."""
import os
import subprocess

import numpy as np
import pytest

import rank_structured_cholesky_amd as R

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "test_cpp_api")


def _run(mode):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is not built (python tests/cpp/build.py)")
    r = subprocess.run([BIN, mode], capture_output=True, text=True, timeout=100)
    return r.returncode, r.stdout + r.stderr


def test_cpp_mirror_builds_and_rejects_bad_calls():
    import importlib.util
    spec = importlib.util.spec_from_file_location("rsc_cpp_test_build",
                                                  os.path.join(HERE, "cpp", "build.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    mod.build()
    rc, out = _run("cpu")
    assert rc == 0, out
    for key in ("unsupported_mode=rejected", "bad_input=rejected", "solve_before_init=rejected"):
        assert key in out


@pytest.mark.gpu
def test_cpp_mirror_factor_pcg_refactor():
    rc, out = _run("gpu")
    print(out)
    assert rc == 0, out
    line = [l for l in out.splitlines() if l.startswith("solve ")][0]
    its = int(line.split("iterations=")[1].split()[0])
    # same problem through the Python mirror
    A, xyz = R.poisson3d(16, 1.0, 1234)
    b = R.uniform_rhs(A._nrow, 5678)
    M = R.FactorPreconditioner()
    M.initialize(A, xyz, 63, 400)
    S = R.PCG_SOLVER(A, 1e-8, 1000)
    x = np.zeros_like(b)
    S.solvePCG(x, b, M)
    assert its == S.last["iterations"]
