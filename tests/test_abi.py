"""C ABI: the library loads, exports every symbol include/rsc.h declares, and
its host entry points validate input (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import rank_structured_cholesky_amd as R
from rank_structured_cholesky_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "include", "rsc.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(",
                           text, flags=re.M))
    return {n for n in names if n.startswith("rsc_") or n == "pcg_pysolve"}


def test_header_symbols_exported():
    names = declared_functions()
    assert len(names) >= 30
    L = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert names == set(_lib.SIGNATURES), names ^ set(_lib.SIGNATURES)


def test_default_options_are_reference_constants():
    o = _lib.default_options()
    # pcg_solver.cpp:100-110, FactorManager.cpp:58-71, 312, 384, SparseSolver.cpp:76-77
    assert (o.max_dense_block, o.max_diagonal_block) == (200, 400)
    assert (o.rank_constant, o.diagonal_rank_constant) == (0.4, 0.5)
    assert (o.oversampling, o.min_rank, o.power_iterations) == (5, 16, 1)
    assert (o.seed, o.explicit_block_threshold) == (152, 4000)
    assert (o.max_attempts, o.retry_alpha) == (10, 1.25)
    assert (o.extended_factor_type, o.compression_type, o.decomposition_block,
            o.decomposition_type) == (R.CHOLESKY, R.IN_PLACE, R.DECOMPOSE_TRANSPOSE,
                                      R.DECOMPOSE_FACTOR)


def test_bad_input_and_unsupported_modes():
    L = _lib.lib()
    A, xyz = R.poisson2d(8)
    h = C.c_void_p()
    o = _lib.default_options()
    bad = A._i.copy()
    bad[3] = 10 ** 6
    st = L.rsc_symbolic_create(A._nrow, _lib.iptr(A._p), _lib.iptr(bad), None, C.byref(o),
                               C.byref(h))
    assert st == _lib.RSC_BAD_INPUT
    assert b"range" in L.rsc_last_error()
    for field, val in [("compression_type", R.EXTENDED_VARIABLE),
                       ("extended_factor_type", R.LDL),
                       ("decomposition_block", R.DECOMPOSE_STANDARD),
                       ("decomposition_type", R.DECOMPOSE_SCHUR),
                       ("sketch_precision", R.RSC_BF16), ("sketch_precision", R.RSC_FP32)]:
        o2 = _lib.default_options(**{field: val})
        st = L.rsc_symbolic_create(A._nrow, _lib.iptr(A._p), _lib.iptr(A._i), None,
                                   C.byref(o2), C.byref(h))
        assert st == _lib.RSC_UNSUPPORTED_MODE
    o3 = _lib.default_options(factor_precision=R.RSC_FP32, sketch_precision=R.RSC_BF16)
    assert L.rsc_symbolic_create(A._nrow, _lib.iptr(A._p), _lib.iptr(A._i), None,
                                 C.byref(o3), C.byref(h)) == _lib.RSC_OK
    L.rsc_symbolic_destroy(h)
    st = L.rsc_factor_create(A._nrow, _lib.iptr(A._p), _lib.iptr(A._i), None, None,
                             C.byref(o), C.byref(h))
    assert st == _lib.RSC_BAD_INPUT


def test_generators_shapes():
    A, xyz = R.poisson2d(256)
    assert A._nrow == 65536 and A._nzmax == 326656
    A3, _ = R.poisson3d(16, 1.0, 1234)
    assert A3._nrow == 4096 and A3._nzmax == 7 * 4096 - 6 * 256
    S = A3.to_scipy()
    assert abs(S - S.T).max() == 0
    E, exyz = R.elasticity3d(3, 1.0, 0.3, 1234)
    assert E._nrow == 3 * 4 * 4 * 3
    Es = E.to_scipy()
    assert abs(Es - Es.T).max() < 1e-12 * abs(Es).max()
    assert np.all(np.linalg.eigvalsh(Es.toarray()) > 0)


def test_bcsm_roundtrip(tmp_path):
    A, _ = R.poisson3d(6, 1.0, 7)
    p = str(tmp_path / "a_system.bcsm.gz")
    A.writeToBinaryGZ(p)
    B = R.SparseColumnMatrix.readFromBinaryGZ(p)
    assert B._nrow == A._nrow and np.array_equal(B._p, A._p)
    assert np.array_equal(B._i, A._i) and np.array_equal(B._x, A._x)
    v = np.arange(5, dtype=float)
    R.write_vector(str(tmp_path / "v.vector"), v)
    assert np.array_equal(R.read_vector(str(tmp_path / "v.vector")), v)


def test_factor_without_gpu_fails_loudly():
    """Without a device the numeric factor must error, never fall back."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("device present")
    except Exception:
        pass
    A, xyz = R.poisson2d(8)
    M = R.FactorPreconditioner()
    with pytest.raises(R.RscError):
        M.initialize(A, xyz, 200, 400)


def test_server_options_and_bad_input():
    L = _lib.lib()
    o = _lib.ServerOptions()
    _lib.check(L.rsc_server_options_default(C.byref(o)))
    assert (o.eps, o.max_iters, o.slots, o.memory_bytes, o.release_workspace) == (
        1e-8, 10000, 64, 0, 1)
    h = C.c_void_p()
    assert L.rsc_server_create(None, 0, C.byref(o), C.byref(h)) == _lib.RSC_BAD_INPUT
    none = (C.c_void_p * 1)(None)
    assert L.rsc_server_create(none, 1, C.byref(o), C.byref(h)) == _lib.RSC_BAD_INPUT
    # the same factor twice: rejected before any handle is dereferenced (the
    # two lanes would share one factor's apply buffers)
    fake = C.c_void_p(0x1000)
    twice = (C.c_void_p * 2)(fake, fake)
    assert L.rsc_server_create(twice, 2, C.byref(o), C.byref(h)) == _lib.RSC_BAD_INPUT
    assert b"twice" in L.rsc_last_error()
    t = C.c_int64()
    assert L.rsc_server_submit(None, 0, 1, None, 1, None, 1, C.byref(t)) == _lib.RSC_BAD_INPUT
    assert L.rsc_server_wait(None, 1, None) == _lib.RSC_BAD_INPUT
    assert L.rsc_apply_status(None, None) == _lib.RSC_BAD_INPUT
    k = C.c_int()
    assert L.rsc_server_slots(None, 0, C.byref(k)) == _lib.RSC_BAD_INPUT
    assert L.rsc_factor_release_workspace(None) == _lib.RSC_BAD_INPUT
    bad = _lib.ServerOptions()
    _lib.check(L.rsc_server_options_default(C.byref(bad)))
    bad.memory_bytes = -1
    assert L.rsc_server_create(twice, 1, C.byref(bad), C.byref(h)) == _lib.RSC_BAD_INPUT
    L.rsc_server_destroy(None)
