"""GPU parity: the HIP path (through the C ABI) against the FP64 CPU oracle
on the same ordering and the same sketch stream."""
import ctypes as C

import numpy as np
import pytest

import rank_structured_cholesky_amd as R
from rank_structured_cholesky_amd import _lib
from helpers import oracle_for, relerr, symbolic

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec,tol", [(R.RSC_FP64, 1e-13), (R.RSC_FP32, 1e-5),
                                      (R.RSC_BF16, 1e-2)])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
# the last three split k over several workgroups (few output tiles, long k)
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (70, 33, 19), (128, 64, 200), (5, 130, 64),
                                   (64, 64, 257), (70, 33, 3001), (200, 150, 20000)])
def test_gemm_selftest(prec, tol, ta, tb, m, n, k):
    rng = np.random.default_rng(m * 1000 + n * 10 + k)
    A = rng.standard_normal((k, m) if ta else (m, k))
    B = rng.standard_normal((n, k) if tb else (k, n))
    Cm = rng.standard_normal((m, n))
    Af, Bf, Cf = (np.asfortranarray(x) for x in (A, B, Cm))
    out = Cf.copy(order="F")
    _lib.check(_lib.lib().rsc_selftest_gemm(prec, ta, tb, m, n, k, -1.5, _lib.dptr(Af),
                                            Af.shape[0], _lib.dptr(Bf), Bf.shape[0], 0.5,
                                            _lib.dptr(out), m))
    opA = A.T if ta else A
    opB = B.T if tb else B
    ref = 0.5 * Cm - 1.5 * opA @ opB
    assert relerr(out, ref) < tol * max(1, np.sqrt(k))


# name: (matrix builder, symbolic options, oracle extra, apply tolerance).
# Apply tolerance: both sides are FP64 with the same sketch stream.  Where
# every sketch is well conditioned (3D cases) only rounding separates them.
# In 2D the separator blocks' singular values fall below u^(1/3) of the
# largest within the rank, so the reference method's own basis (Householder
# QR of the cubed sketch) is rounding noise in those directions; the device
# path resolves them (orthonormalized subspace iteration), and the two agree
# to the reference's noise floor (~1e-6 of ||L21||).
CASES = {
    "p2d48_exact": (lambda: R.poisson2d(48), dict(max_dense_block=10**9), {}, 1e-12),
    "p2d64_c31": (lambda: R.poisson2d(64), dict(max_dense_block=31), {}, 1e-4),
    "p3d16_c63": (lambda: R.poisson3d(16, 1.0, 1234), dict(max_dense_block=63), {}, 1e-9),
    "p3d20_hodlr": (lambda: R.poisson3d(20, 1.0, 1234),
                    dict(max_dense_block=63, max_diagonal_block=50,
                         explicit_block_threshold=120), dict(explicit_threshold=120), 1e-8),
    "el6_c30": (lambda: R.elasticity3d(6, 1.0, 0.3, 2),
                dict(max_dense_block=30, max_diagonal_block=40), {}, 1e-8),
}

# Every rank at min(rows, cols): each range finder captures its block
# exactly and the factor is the exact Cholesky factor of P A P^T.
FULL = dict(rank_constant=1e4, diagonal_rank_constant=1e4)
FULL_CASES = {
    "p3d12_hodlr": (lambda: R.poisson3d(12, 1.0, 3),
                    dict(max_dense_block=20, max_diagonal_block=40)),
    "p3d12_explicit": (lambda: R.poisson3d(12, 1.0, 3),
                       dict(max_dense_block=20, max_diagonal_block=40,
                            explicit_block_threshold=80)),
    "el6": (lambda: R.elasticity3d(6, 1.0, 0.3, 2),
            dict(max_dense_block=30, max_diagonal_block=40)),
}


def _factor(A, xyz, **opts):
    M = R.FactorPreconditioner()
    mdb = opts.pop("max_dense_block")
    mdiag = opts.pop("max_diagonal_block", 400)
    M.initialize(A, xyz, mdb, mdiag, **opts)
    return M


@pytest.mark.parametrize("name", list(CASES))
def test_apply_matches_oracle_fp64(name):
    build, sopts, okw, tol = CASES[name]
    A, xyz = build()
    sym = symbolic(A, xyz, **sopts)
    O = oracle_for(A, sym, **okw)
    M = _factor(A, xyz, **dict(sopts))
    assert np.array_equal(M.permutation(), sym["perm"])
    b = R.uniform_rhs(A._nrow, 5678)
    x = np.empty_like(b)
    M.solve(x, b)
    xo = O.apply(b)
    assert relerr(x, xo) < tol, relerr(x, xo)


@pytest.mark.parametrize("name", list(CASES))
def test_pcg_iterations_match_oracle(name):
    build, sopts, okw, _ = CASES[name]
    A, xyz = build()
    sym = symbolic(A, xyz, **sopts)
    O = oracle_for(A, sym, **okw)
    M = _factor(A, xyz, **dict(sopts))
    b = R.uniform_rhs(A._nrow, 5678)
    x = np.zeros_like(b)
    res = R.PCG_SOLVER(A, 1e-8, 500).solvePCG(x, b, M)
    xo, ro = O.pcg(b, 1e-8, 500)
    assert abs(res["iterations"] - ro["iterations"]) <= 1, (res, ro)
    assert res["true_relres"] < 1e-7
    assert relerr(A.multiply(x), b) < 1e-7
    if "exact" in name:
        assert res["iterations"] == 1


def test_fp32_factor_converges():
    A, xyz = R.poisson3d(20, 1.0, 1234)
    M = _factor(A, xyz, max_dense_block=63, factor_precision=R.RSC_FP32)
    M64 = _factor(A, xyz, max_dense_block=63)
    b = R.uniform_rhs(A._nrow, 5678)
    x = np.zeros_like(b)
    x64 = np.zeros_like(b)
    r32 = R.PCG_SOLVER(A, 1e-8, 500).solvePCG(x, b, M)
    r64 = R.PCG_SOLVER(A, 1e-8, 500).solvePCG(x64, b, M64)
    assert r32["true_relres"] < 1e-7
    assert r32["iterations"] <= r64["iterations"] + 3


def test_block_pcg_is_batch_invariant():
    """Several right-hand sides run on the matrix-core task bodies, one on
    the scalar ones.  A column's result does not depend on the other columns
    or its position (bitwise), and agrees with the single solve to the PCG
    tolerance."""
    A, xyz = R.poisson3d(16, 1.0, 1234)
    M = _factor(A, xyz, max_dense_block=63)
    n = A._nrow
    B = R.uniform_rhs(4 * n, 5678).reshape(4, n).T.copy()
    X = np.zeros_like(B)
    res = R.PCG_SOLVER(A, 1e-8, 500).solvePCG(X, B, M)
    perm = [2, 0, 3, 1]
    Xp = np.zeros_like(B)
    resp = R.PCG_SOLVER(A, 1e-8, 500).solvePCG(Xp, B[:, perm].copy(), M)
    two = np.zeros((n, 2))
    res2 = R.PCG_SOLVER(A, 1e-8, 500).solvePCG(two, B[:, [3, 1]].copy(), M)
    for q, j in enumerate(perm):
        assert np.array_equal(Xp[:, q], X[:, j])
        assert resp[q]["iterations"] == res[j]["iterations"]
    assert np.array_equal(two[:, 0], X[:, 3]) and np.array_equal(two[:, 1], X[:, 1])
    assert res2[0]["true_relres"] == res[3]["true_relres"]
    for j in range(4):
        x = np.zeros(n)
        r1 = R.PCG_SOLVER(A, 1e-8, 500).solvePCG(x, B[:, j].copy(), M)
        assert abs(res[j]["iterations"] - r1["iterations"]) <= 1
        assert relerr(X[:, j], x) < 1e-6
        assert res[j]["true_relres"] < 1e-7


@pytest.mark.parametrize("env", [dict(RSC_APPLY_LEVELS="1"), dict(RSC_SUBTREE_ELEMS="1"),
                                 dict(RSC_SUBTREE_ELEMS="1000000"), dict(RSC_DF_QUEUE="0"),
                                 dict(RSC_MS_SUBTREE_TASKS="3")])
def test_apply_schedules_agree(monkeypatch, env):
    """The level-scheduled launches, the dataflow launch with no or with
    large subtree tasks, ticket dispatch instead of the ready queue, and the
    multi-right-hand-side subtrees merged into three tasks compute the same
    apply (one and three right-hand sides) as the default dataflow schedule,
    to rounding."""
    build, sopts, okw, tol = CASES["p3d20_hodlr"]
    A, xyz = build()
    n = A._nrow
    B = R.uniform_rhs(3 * n, 77).reshape(3, n).T.copy()
    ref = {}
    M = _factor(A, xyz, **dict(sopts))
    for k in (1, 3):
        X = np.empty((n, k))
        M.solve(X, B[:, :k].copy())
        ref[k] = X
    del M
    for key, val in env.items():
        monkeypatch.setenv(key, val)
    M = _factor(A, xyz, **dict(sopts))
    for k in (1, 3):
        X = np.empty((n, k))
        M.solve(X, B[:, :k].copy())
        assert relerr(X, ref[k]) < 1e-12


def test_multi_rhs_apply_matches_oracle():
    """The matrix-core apply bodies (two and seventeen right-hand sides: one
    full group of sixteen plus a partial one) against the FP64 oracle."""
    build, sopts, okw, tol = CASES["p3d20_hodlr"]
    A, xyz = build()
    sym = symbolic(A, xyz, **sopts)
    O = oracle_for(A, sym, **okw)
    M = _factor(A, xyz, **dict(sopts))
    n = A._nrow
    for k in (2, 17):
        B = R.uniform_rhs(k * n, 99).reshape(k, n).T.copy()
        X = np.empty_like(B)
        M.solve(X, B)
        for j in (0, k - 1):
            assert relerr(X[:, j], O.apply(B[:, j])) < tol


@pytest.mark.parametrize("prec,tol", [(R.RSC_FP64, 1e-11), (R.RSC_FP32, 1e-4)])
def test_level_batched_factor_matches_node_by_node(monkeypatch, prec, tol):
    """The compressed supernodes of a level are emitted as shared launches;
    emitting them one supernode at a time must give the same factor up to
    the summation order of split-k GEMM shares."""
    A, xyz = R.poisson3d(20, 1.0, 1234)
    opts = dict(max_dense_block=63, max_diagonal_block=50, explicit_block_threshold=120,
                factor_precision=prec)
    sym = symbolic(A, xyz, max_dense_block=63)
    comp_levels = sym["level"][sym["compressed"].astype(bool)]
    assert np.max(np.bincount(comp_levels)) > 1  # some level holds several
    b = R.uniform_rhs(A._nrow, 5678)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("RSC_FACTOR_BATCH", mode)
        M = _factor(A, xyz, **opts)
        x = np.empty_like(b)
        M.solve(x, b)
        out[mode] = x
    assert relerr(out["1"], out["0"]) < tol


@pytest.mark.parametrize("name", ["p3d16_c63", "p3d20_hodlr", "el6_c30"])
def test_bf16_sketch_factor_converges(name):
    """FP32 factor with BF16 range-finder products (FP32 accumulation): the
    apply stays within 1e-2 of the FP64 oracle and PCG reaches 1e-8 within
    a few iterations of the FP64 factor's count."""
    build, sopts, okw, tol = CASES[name]
    A, xyz = build()
    sym = symbolic(A, xyz, **sopts)
    O = oracle_for(A, sym, **okw)
    b = R.uniform_rhs(A._nrow, 5678)
    M64 = _factor(A, xyz, **dict(sopts))
    x = np.zeros_like(b)
    r64 = R.PCG_SOLVER(A, 1e-8, 500).solvePCG(x, b, M64)
    M = _factor(A, xyz, factor_precision=R.RSC_FP32, sketch_precision=R.RSC_BF16,
                **dict(sopts))
    y = np.empty_like(b)
    M.solve(y, b)
    assert np.all(np.isfinite(y))
    assert relerr(y, O.apply(b)) < 1e-2
    x = np.zeros_like(b)
    res = R.PCG_SOLVER(A, 1e-8, 500).solvePCG(x, b, M)
    assert res["true_relres"] < 1e-7
    assert res["iterations"] <= r64["iterations"] + max(3, r64["iterations"] // 5)


@pytest.mark.parametrize("name", list(CASES))
def test_fp32_factor_apply_close_to_oracle(name):
    """FP32 factor (FP64 vectors): the apply stays within FP32 factor accuracy
    of the FP64 oracle and PCG still reaches 1e-8."""
    build, sopts, okw, tol = CASES[name]
    A, xyz = build()
    sym = symbolic(A, xyz, **sopts)
    O = oracle_for(A, sym, **okw)
    M = _factor(A, xyz, factor_precision=R.RSC_FP32, **dict(sopts))
    b = R.uniform_rhs(A._nrow, 5678)
    x = np.empty_like(b)
    M.solve(x, b)
    assert np.all(np.isfinite(x))
    assert relerr(x, O.apply(b)) < max(1e-3, tol)
    x = np.zeros_like(b)
    res = R.PCG_SOLVER(A, 1e-8, 500).solvePCG(x, b, M)
    assert res["true_relres"] < 1e-7


@pytest.mark.parametrize("name", list(FULL_CASES))
def test_full_rank_compression_is_exact(name):
    import scipy.sparse.linalg as spla
    build, sopts = FULL_CASES[name]
    A, xyz = build()
    M = _factor(A, xyz, **dict(sopts, **FULL))
    assert M.info()["ncompressed"] > 0
    b = R.uniform_rhs(A._nrow, 99)
    x = np.empty_like(b)
    M.solve(x, b)
    e = spla.spsolve(A.to_scipy().tocsc(), b)
    assert relerr(x, e) < 1e-9, relerr(x, e)


def test_apply_wait_timeout_is_reported_and_device_recovers(monkeypatch):
    """A dependency wait that exceeds its bound (forced here with a zero
    bound) aborts the dataflow launch: every workgroup drains, the apply and
    the PCG solve report a device error, and the next factor works."""
    A, xyz = R.poisson3d(16, 1.0, 1234)
    b = R.uniform_rhs(A._nrow, 5678)
    monkeypatch.setenv("RSC_DF_TIMEOUT_US", "0")
    M = _factor(A, xyz, max_dense_block=63)
    y = np.empty_like(b)
    with pytest.raises(_lib.RscError) as e:
        M.solve(y, b)
    assert e.value.status == _lib.RSC_DEVICE_ERROR
    assert "timed out" in str(e.value)
    with pytest.raises(_lib.RscError):
        R.PCG_SOLVER(A, 1e-8, 50).solvePCG(np.zeros_like(b), b, M)
    del M
    monkeypatch.delenv("RSC_DF_TIMEOUT_US")
    M2 = _factor(A, xyz, max_dense_block=63)
    x = np.zeros_like(b)
    res = R.PCG_SOLVER(A, 1e-8, 500).solvePCG(x, b, M2)
    assert res["true_relres"] < 1e-7
