"""rsc_apply_noalloc: the apply with its scratch in a fixed device array
(no allocation during the call) against rsc_apply through the C ABI, bit
for bit, on two calls in a row (the second with a smaller input), and the
rejection of an apply whose scratch exceeds the array."""
import ctypes as C

import numpy as np
import pytest

import rank_structured_cholesky_amd as R
from rank_structured_cholesky_amd import _lib

pytestmark = pytest.mark.gpu


def _factor(N, mdb):
    A, xyz = R.poisson3d(N, 1.0, 1234)
    M = R.FactorPreconditioner()
    M.initialize(A, xyz, mdb, 400)
    return A, M


def _both(torch, M, n, k, seed):
    """(rsc_apply result, rsc_apply_noalloc result) for k uniform columns."""
    L = _lib.lib()
    B = torch.from_numpy(R.uniform_rhs(k * n, seed).reshape(k, n)).to("cuda")
    Xa = torch.full_like(B, float("nan"))
    Xn = torch.full_like(B, float("nan"))
    torch.cuda.synchronize()
    s = torch.cuda.current_stream().cuda_stream
    _lib.check(L.rsc_apply(M.handle, k, C.c_void_p(B.data_ptr()), n, C.c_void_p(Xa.data_ptr()),
                           n, C.c_void_p(s)))
    _lib.check(L.rsc_apply_status(M.handle, C.c_void_p(s)))
    _lib.check(L.rsc_apply_noalloc(M.handle, k, C.c_void_p(B.data_ptr()), n,
                                   C.c_void_p(Xn.data_ptr()), n, C.c_void_p(s)))
    _lib.check(L.rsc_apply_status(M.handle, C.c_void_p(s)))
    torch.cuda.synchronize()
    return Xa.cpu().numpy(), Xn.cpu().numpy()


def test_noalloc_apply_equals_apply():
    torch = pytest.importorskip("torch")
    A1, M1 = _factor(16, 63)
    A2, M2 = _factor(12, 40)
    # two calls in a row, the second with a smaller input (fewer right-hand
    # sides of a smaller system), then the first again over the stale array
    for M, A, k, seed in ((M1, A1, 3, 11), (M2, A2, 1, 12), (M1, A1, 17, 13)):
        xa, xn = _both(torch, M, A._nrow, k, seed)
        assert np.all(np.isfinite(xa))
        assert np.array_equal(xa, xn), (A._nrow, k)


def test_noalloc_rejects_oversized_scratch():
    torch = pytest.importorskip("torch")
    A, M = _factor(64, 255)
    n, k = A._nrow, 16
    L = _lib.lib()
    B = torch.ones((k, n), dtype=torch.float64, device="cuda")
    X = torch.full_like(B, float("nan"))
    torch.cuda.synchronize()
    st = L.rsc_apply_noalloc(M.handle, k, C.c_void_p(B.data_ptr()), n,
                             C.c_void_p(X.data_ptr()), n, None)
    assert st == _lib.RSC_BAD_INPUT
    assert b"scratch" in L.rsc_last_error()
    torch.cuda.synchronize()
    assert torch.isnan(X).all().item()
    # one right-hand side of the same system fits, and matches rsc_apply
    xa, xn = _both(torch, M, n, 1, 21)
    assert np.array_equal(xa, xn)
