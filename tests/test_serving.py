"""Request server (continuous batching): a request's solution, iteration
count and residuals do not depend on what else shares the device with it
(bitwise), and equal a solo block PCG solve of the same right-hand sides:
bitwise for several right-hand sides (both run the matrix-core apply
bodies), to the PCG tolerance for one (the solo apply then runs the scalar
bodies)."""
import ctypes as C
import time

import numpy as np
import pytest

import rank_structured_cholesky_amd as R
from rank_structured_cholesky_amd import _lib

pytestmark = pytest.mark.gpu

EPS = 1e-8


def _factors():
    """Four resident systems of different sizes: three FP64 factors and one
    mixed (FP32 factor, BF16 sketches)."""
    out = []
    for N, sigma, extra in ((16, 1.0, {}), (20, 1.5, {}), (12, 1.0, {}),
                            (14, 2.0, dict(factor_precision=R.RSC_FP32,
                                           sketch_precision=R.RSC_BF16))):
        A, xyz = R.poisson3d(N, sigma, 1234)
        M = R.FactorPreconditioner()
        M.initialize(A, xyz, 63, 400, **extra)
        out.append((A, M))
    return out


def _solo(torch, M, B):
    """rsc_block_pcg_solve from x0 = 0 on device copies of B (n x k)."""
    n, k = B.shape
    dB = torch.from_numpy(np.ascontiguousarray(B.T)).to("cuda")
    dX = torch.zeros_like(dB)
    res = (_lib.PcgResult * k)()
    st = _lib.lib().rsc_block_pcg_solve(M.handle, k, C.c_void_p(dB.data_ptr()), n,
                                        C.c_void_p(dX.data_ptr()), n, EPS, 500, res, None)
    assert st == _lib.RSC_OK
    torch.cuda.synchronize()
    return dX.cpu().numpy().T.copy(), [(r.iterations, r.recurrence_relres, r.true_relres)
                                       for r in res]


def test_served_results_equal_solo_solves():
    torch = pytest.importorskip("torch")
    fac = _factors()
    # (factor, number of right-hand sides): more columns than slots in total,
    # one request larger than the slot pool
    reqs = [(0, 3), (1, 5), (2, 4), (0, 12), (3, 6), (1, 1), (2, 9), (0, 2), (3, 1),
            (1, 7), (0, 1)]
    rng_seed = 100
    data = []
    for f, m in reqs:
        n = fac[f][0]._nrow
        B = R.uniform_rhs(n * m, rng_seed).reshape(m, n).T.copy()
        rng_seed += 1
        data.append(B)
    solo = [_solo(torch, fac[f][1], B) for (f, _), B in zip(reqs, data)]

    def serve(order, stagger):
        srv = R.RequestServer([M for _, M in fac], eps=EPS, max_iters=500, slots=8)
        dev, tickets = {}, {}
        for i in order:
            f, m = reqs[i]
            dB = torch.from_numpy(np.ascontiguousarray(data[i].T)).to("cuda")
            dX = torch.full_like(dB, float("nan"))
            torch.cuda.synchronize()
            n = dB.shape[1]
            tickets[i] = srv.submit(f, m, dB.data_ptr(), n, dX.data_ptr(), n)
            dev[i] = (dB, dX)
            if stagger:
                time.sleep(0.002)
        out = {}
        for i in order:
            res = srv.wait(tickets[i], reqs[i][1])
            torch.cuda.synchronize()
            out[i] = (dev[i][1].cpu().numpy().T.copy(), res)
        st = srv.stats()
        srv.close()
        return out, st

    mixed, st = serve(list(range(len(reqs))), stagger=True)
    assert st["max_busy"] > 1
    alone = {}
    for i in (3, 4, 5, 6, 8):
        o, _ = serve([i], stagger=False)
        alone.update(o)
    reversed_mix, _ = serve(list(reversed(range(len(reqs)))), stagger=False)
    for i in range(len(reqs)):
        x0, res0 = mixed[i]
        for got in (reversed_mix[i],) + ((alone[i],) if i in alone else ()):
            x, res = got
            assert np.array_equal(x, x0), i
            for r, q in zip(res, res0):
                assert (r["iterations"], r["recurrence_relres"], r["true_relres"]) == \
                    (q["iterations"], q["recurrence_relres"], q["true_relres"])
        xs, rs = solo[i]
        for j, (r, (it, rec, tru)) in enumerate(zip(res0, rs)):
            assert r["recurrence_relres"] <= EPS and r["seconds"] > 0
            if reqs[i][1] > 1:
                assert (r["iterations"], r["recurrence_relres"], r["true_relres"]) == \
                    (it, rec, tru)
            else:
                assert abs(r["iterations"] - it) <= 1
        if reqs[i][1] > 1:
            assert np.array_equal(x0, xs), i
        else:
            assert np.linalg.norm(x0 - xs) <= 1e-6 * np.linalg.norm(xs), i


def test_server_memory_budget_and_workspace_release():
    """Slots come from one device-memory budget, divided by each factor's
    bytes per column: a 1-byte budget leaves one slot per factor (a request
    wider than that is admitted column by column and still equals its solo
    solve bitwise), an unlimited one the 64-slot limit, a moderate one fewer
    slots for the largest system than for the smallest.  The server releases
    the factors' numeric workspaces; a later refactorization allocates them
    again and gives the same preconditioner bitwise."""
    torch = pytest.importorskip("torch")
    fac = _factors()
    mats = [A for A, _ in fac]
    Ms = [M for _, M in fac]
    assert all(M.info()["workspace_bytes"] > 0 for M in Ms)
    b = R.uniform_rhs(mats[0]._nrow, 77)
    y0 = np.empty_like(b)
    Ms[0].solve(y0, b)
    B = R.uniform_rhs(mats[0]._nrow * 3, 78).reshape(3, -1).T.copy()
    xs, rs = _solo(torch, Ms[0], B)

    srv = R.RequestServer(Ms, eps=EPS, max_iters=500, slots=64, memory_bytes=1)
    assert [srv.slots(f) for f in range(4)] == [1, 1, 1, 1]
    assert all(M.info()["workspace_bytes"] == 0 for M in Ms)
    dB = torch.from_numpy(np.ascontiguousarray(B.T)).to("cuda")
    dX = torch.full_like(dB, float("nan"))
    torch.cuda.synchronize()
    n = dB.shape[1]
    res = srv.wait(srv.submit(0, 3, dB.data_ptr(), n, dX.data_ptr(), n), 3)
    torch.cuda.synchronize()
    assert srv.stats()["max_busy"] == 1
    srv.close()
    assert np.array_equal(dX.cpu().numpy().T, xs)
    for r, (it, rec, tru) in zip(res, rs):
        assert (r["iterations"], r["recurrence_relres"], r["true_relres"]) == (it, rec, tru)

    big = R.RequestServer(Ms, eps=EPS, max_iters=500, slots=64, memory_bytes=1 << 40)
    assert [big.slots(f) for f in range(4)] == [64, 64, 64, 64]
    big.close()
    # 40 MB: the 20^3 system (largest columns) cannot reach the limit, the
    # 12^3 one (smallest) gets at least as many slots
    mid = R.RequestServer(Ms, eps=EPS, max_iters=500, slots=64, memory_bytes=40 << 20)
    k = [mid.slots(f) for f in range(4)]
    mid.close()
    assert k[1] < 64 and k[2] >= k[1] and min(k) >= 1, k

    _lib.check(_lib.lib().rsc_factor_refactor(Ms[0].handle, _lib.dptr(mats[0]._x)))
    assert Ms[0].info()["workspace_bytes"] > 0
    y1 = np.empty_like(b)
    Ms[0].solve(y1, b)
    assert np.array_equal(y0, y1)
