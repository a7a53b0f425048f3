"""GPU parity at the benchmark configurations and for every apply variant:
the HIP path (through the C ABI) against the FP64 CPU oracle on the same
ordering and sketch stream, or -- where the oracle does not finish in
seconds -- against the device's own independent schedules and the true
residual.

Tolerances (stated per test): FP64 factor against the FP64 oracle to
rounding (3D) or the reference method's noise floor (2D, 1e-4: the
reference's Householder QR of the cubed sketch resolves nothing below
u^(1/3) of the block's largest singular value, the device's orthonormalized
subspace iteration does); FP32 factor 1e-3 and BF16 sketches 1e-2 of the
FP64 oracle; PCG iteration counts within 10% (and at least +-1) of the
oracle's at the same ranks, the north-star parity bar."""
import ctypes as C
import os

import numpy as np
import pytest

import rank_structured_cholesky_amd as R
from rank_structured_cholesky_amd import _lib
from helpers import oracle_for, relerr, symbolic

pytestmark = pytest.mark.gpu


def _factor(A, xyz, **opts):
    M = R.FactorPreconditioner()
    mdb = opts.pop("max_dense_block")
    mdiag = opts.pop("max_diagonal_block", 400)
    M.initialize(A, xyz, mdb, mdiag, **opts)
    return M


def _iters_close(a, b):
    return abs(a - b) <= max(1, int(0.1 * b))


def _sym_and_oracle(A, xyz, sopts, **okw):
    sym = symbolic(A, xyz, **sopts)
    kw = dict(okw)
    for key in ("max_rank", "oversampling"):
        if key in sopts:
            kw[key] = sopts[key]
    if "explicit_block_threshold" in sopts:
        kw["explicit_threshold"] = sopts["explicit_block_threshold"]
    return sym, oracle_for(A, sym, **kw)


# ---- FP32-factor matrix-core apply (k_apply_df<float, true>) -------------

VARIANT_CASES = {
    "p3d20_hodlr": (lambda: R.poisson3d(20, 1.0, 1234),
                    dict(max_dense_block=63, max_diagonal_block=50,
                         explicit_block_threshold=120)),
    "el6_c30": (lambda: R.elasticity3d(6, 1.0, 0.3, 2),
                dict(max_dense_block=30, max_diagonal_block=40)),
}


@pytest.mark.parametrize("name", list(VARIANT_CASES))
@pytest.mark.parametrize("sketch,tol", [(R.RSC_FP32, 1e-3), (R.RSC_BF16, 1e-2)])
def test_fp32_factor_multi_rhs_apply(name, sketch, tol):
    """FP32 factor (FP32 or BF16 range-finder products) on the matrix-core
    multi-right-hand-side bodies: 2, 16 and 17 columns (one partial group of
    sixteen, one full, one full plus one) against the FP64 oracle, and each
    column bit-identical whatever else shares the apply."""
    build, sopts = VARIANT_CASES[name]
    A, xyz = build()
    sym, O = _sym_and_oracle(A, xyz, sopts)
    M = _factor(A, xyz, factor_precision=R.RSC_FP32, sketch_precision=sketch, **dict(sopts))
    n = A._nrow
    B = R.uniform_rhs(17 * n, 4242).reshape(17, n).T.copy()
    ref = np.stack([O.apply(B[:, j]) for j in range(17)], axis=1)
    out = {}
    for k in (2, 16, 17):
        X = np.empty((n, k))
        M.solve(X, B[:, :k].copy())
        assert np.all(np.isfinite(X))
        for j in range(k):
            assert relerr(X[:, j], ref[:, j]) < tol, (k, j, relerr(X[:, j], ref[:, j]))
        out[k] = X
    # batch invariance of the matrix-core bodies: column j's bits do not
    # depend on how many columns share the apply
    assert np.array_equal(out[2], out[16][:, :2])
    assert np.array_equal(out[16], out[17][:, :16])
    # and a column's bits do not depend on its position
    Bp = B[:, [5, 0, 16, 1]].copy()
    Xp = np.empty((n, 4))
    M.solve(Xp, Bp)
    for q, j in enumerate([5, 0, 16, 1]):
        assert np.array_equal(Xp[:, q], out[17][:, j])


def test_mixed_factor_block_pcg_sigma2():
    """The cfg3 form at 20^3: sigma = 2 log-normal Poisson, FP32 factor with
    BF16 sketches (oversampling 10), FP64 block PCG on 16 right-hand sides;
    per column against the FP64 oracle's PCG at the same ranks."""
    A, xyz = R.poisson3d(20, 2.0, 1234)
    sopts = dict(max_dense_block=63, oversampling=10)
    sym, O = _sym_and_oracle(A, xyz, sopts)
    M = _factor(A, xyz, factor_precision=R.RSC_FP32, sketch_precision=R.RSC_BF16,
                **dict(sopts))
    n = A._nrow
    B = R.uniform_rhs(16 * n, 5678).reshape(16, n).T.copy()
    X = np.zeros_like(B)
    res = R.PCG_SOLVER(A, 1e-8, 1000).solvePCG(X, B, M)
    for j in range(16):
        xo, ro = O.pcg(B[:, j], 1e-8, 1000)
        assert res[j]["recurrence_relres"] <= 1e-8
        assert res[j]["true_relres"] < 2e-8, res[j]
        assert _iters_close(res[j]["iterations"], ro["iterations"]), (j, res[j], ro)
        assert relerr(A.multiply(X[:, j]), B[:, j]) < 2e-8


# ---- cfg1 at its full size ------------------------------------------------

@pytest.mark.parametrize("max_rank", [32, 0])
def test_poisson2d_256_full_size(max_rank):
    """cfg1 at full size: 2D 5-point Poisson 256^2 (n = 65,536), separators
    of at least 128 compressed, ranks CAPPED(32) or REFERENCE; one uniform
    right-hand side.  Apply within the 2D noise floor of the oracle's,
    iterations +-1, true residual at the PCG tolerance."""
    A, xyz = R.poisson2d(256)
    sopts = dict(max_dense_block=127, max_rank=max_rank)
    sym, O = _sym_and_oracle(A, xyz, sopts)
    assert sym["compressed"].sum() > 0
    M = _factor(A, xyz, **dict(sopts))
    b = R.uniform_rhs(A._nrow, 5678)
    y = np.empty_like(b)
    M.solve(y, b)
    assert relerr(y, O.apply(b)) < 1e-4
    x = np.zeros_like(b)
    res = R.PCG_SOLVER(A, 1e-8, 1000).solvePCG(x, b, M)
    xo, ro = O.pcg(b, 1e-8, 1000)
    assert abs(res["iterations"] - ro["iterations"]) <= 1, (res, ro)
    assert res["true_relres"] < 1e-8 * 1.5


# ---- cfg2 family ---------------------------------------------------------

def test_poisson3d_32_cfg2_options_against_oracle():
    """cfg2's options (separators of at least 256 compressed, CAPPED(64),
    sigma 1) at 32^3 (n = 32,768), where the oracle finishes in seconds:
    apply to 1e-9, iterations +-1."""
    A, xyz = R.poisson3d(32, 1.0, 1234)
    sopts = dict(max_dense_block=255, max_rank=64)
    sym, O = _sym_and_oracle(A, xyz, sopts)
    assert sym["compressed"].sum() > 0
    M = _factor(A, xyz, **dict(sopts))
    b = R.uniform_rhs(A._nrow, 5678)
    y = np.empty_like(b)
    M.solve(y, b)
    assert relerr(y, O.apply(b)) < 1e-9
    x = np.zeros_like(b)
    res = R.PCG_SOLVER(A, 1e-8, 1000).solvePCG(x, b, M)
    xo, ro = O.pcg(b, 1e-8, 1000)
    assert abs(res["iterations"] - ro["iterations"]) <= 1, (res, ro)
    assert res["true_relres"] < 2e-8


def test_poisson3d_64_full_size_schedules(monkeypatch):
    """cfg2 at full size (64^3, n = 262,144; separators of at least 256
    compressed, CAPPED(64)).  The oracle does not finish in seconds here, so
    the dataflow apply -- the ~10^5-task persistent launch and its
    cross-workgroup protocol -- is checked against the level-scheduled
    launches (one kernel per level and phase, ordered by the stream) to
    1e-12, the sixteen-column matrix-core apply column by column against
    the single-column apply to 1e-12, and PCG by its true residual."""
    A, xyz = R.poisson3d(64, 1.0, 1234)
    sopts = dict(max_dense_block=255, max_rank=64)
    n = A._nrow
    B = R.uniform_rhs(16 * n, 5678).reshape(16, n).T.copy()
    M = _factor(A, xyz, **dict(sopts))
    y1 = np.empty(n)
    M.solve(y1, B[:, 0].copy())
    Y16 = np.empty((n, 16))
    M.solve(Y16, B)
    for j in (0, 7, 15):
        yj = np.empty(n)
        M.solve(yj, B[:, j].copy())
        assert relerr(Y16[:, j], yj) < 1e-12, j
    x = np.zeros(n)
    res = R.PCG_SOLVER(A, 1e-8, 1000).solvePCG(x, B[:, 0].copy(), M)
    assert res["recurrence_relres"] <= 1e-8
    assert res["true_relres"] < 2e-8
    assert relerr(A.multiply(x), B[:, 0]) < 2e-8
    # the same apply twice gives the same bits (no race in the dataflow
    # launch leaves a trace in the result)
    y1b = np.empty(n)
    M.solve(y1b, B[:, 0].copy())
    assert np.array_equal(y1, y1b)
    del M
    monkeypatch.setenv("RSC_APPLY_LEVELS", "1")
    ML = _factor(A, xyz, **dict(sopts))
    z1 = np.empty(n)
    ML.solve(z1, B[:, 0].copy())
    assert relerr(z1, y1) < 1e-12
    Z16 = np.empty((n, 16))
    ML.solve(Z16, B)
    assert relerr(Z16, Y16) < 1e-12


# ---- cfg4 family: elasticity, rank 96, 64 right-hand sides ----------------

def test_elasticity_rank96_mixed_block_pcg_64rhs():
    """cfg4's form at 10^3 elements (n = 3,630): trilinear-hex elasticity,
    CAPPED(96) ranks, FP32 factor with BF16 sketches, FP64 block PCG on 64
    uniform right-hand sides; each column against the FP64 oracle's PCG."""
    A, xyz = R.elasticity3d(10, 1.0, 0.3, 1234)
    sopts = dict(max_dense_block=100, max_rank=96)
    sym, O = _sym_and_oracle(A, xyz, sopts)
    assert sym["compressed"].sum() > 0
    M = _factor(A, xyz, factor_precision=R.RSC_FP32, sketch_precision=R.RSC_BF16,
                **dict(sopts))
    n = A._nrow
    B = R.uniform_rhs(64 * n, 97).reshape(64, n).T.copy()
    X = np.zeros_like(B)
    res = R.PCG_SOLVER(A, 1e-8, 1000).solvePCG(X, B, M)
    for j in range(64):
        xo, ro = O.pcg(B[:, j], 1e-8, 1000)
        assert res[j]["recurrence_relres"] <= 1e-8
        assert _iters_close(res[j]["iterations"], ro["iterations"]), (j, res[j], ro)
        assert relerr(A.multiply(X[:, j]), B[:, j]) < 2e-8, j
    Y = np.empty((n, 64))
    M.solve(Y, B)
    for j in (0, 31, 63):
        assert relerr(Y[:, j], O.apply(B[:, j])) < 1e-2


# ---- the error against an FP64 direct solve --------------------------------

def test_exact_factor_is_a_direct_solver():
    """With no compression the factor is the exact Cholesky factor, so one
    apply is an FP64 direct solve: it must equal scipy's sparse direct solve
    to rounding times the condition number, and the compressed factor's PCG
    solution must lie within cond(A) x its true relative residual of it."""
    import scipy.sparse.linalg as spla
    A, xyz = R.poisson3d(16, 2.0, 1234)
    S = A.to_scipy().tocsc()
    b = R.uniform_rhs(A._nrow, 5678)
    E = _factor(A, xyz, max_dense_block=10 ** 9)
    assert E.info()["ncompressed"] == 0
    xd = np.empty_like(b)
    E.solve(xd, b)
    xs = spla.spsolve(S, b)
    lmax = spla.eigsh(S, k=1, which="LA", return_eigenvectors=False)[0]
    lmin = spla.eigsh(S, k=1, sigma=0, which="LM", return_eigenvectors=False)[0]
    kappa = lmax / lmin
    assert relerr(xd, xs) < 100 * kappa * 2.2e-16
    M = _factor(A, xyz, max_dense_block=63)
    x = np.zeros_like(b)
    res = R.PCG_SOLVER(A, 1e-8, 1000).solvePCG(x, b, M)
    assert res["true_relres"] < 2e-8
    assert relerr(x, xd) <= 1.01 * kappa * res["true_relres"] + 100 * kappa * 2.2e-16


# ---- refactor with new coefficients ----------------------------------------

@pytest.mark.parametrize("prec", [R.RSC_FP64, R.RSC_FP32])
def test_refactor_new_coefficients_matches_oracle(prec):
    """Same pattern, new values (coefficient seed 99 instead of 1234): the
    refactored preconditioner equals the oracle built on the new values, and
    equals a fresh factorization of them."""
    A, xyz = R.poisson3d(20, 1.0, 1234)
    A2, _ = R.poisson3d(20, 1.0, 99)
    assert np.array_equal(A._p, A2._p) and np.array_equal(A._i, A2._i)
    assert not np.array_equal(A._x, A2._x)
    sopts = dict(max_dense_block=63, max_diagonal_block=50, explicit_block_threshold=120)
    sym2, O2 = _sym_and_oracle(A2, xyz, sopts)
    M = _factor(A, xyz, factor_precision=prec, **dict(sopts))
    M.clearFactor()
    mdb = sopts["max_dense_block"]
    rest = {k: v for k, v in sopts.items() if k not in ("max_dense_block",
                                                        "max_diagonal_block")}
    M.initialize(A2, xyz, mdb, sopts["max_diagonal_block"], factor_precision=prec, **rest)
    assert np.array_equal(M.permutation(), sym2["perm"])
    b = R.uniform_rhs(A._nrow, 5678)
    y = np.empty_like(b)
    M.solve(y, b)
    tol = 1e-8 if prec == R.RSC_FP64 else 1e-3
    assert relerr(y, O2.apply(b)) < tol
    F = _factor(A2, xyz, factor_precision=prec, **dict(sopts))
    yf = np.empty_like(b)
    F.solve(yf, b)
    assert np.array_equal(y, yf)
    x = np.zeros_like(b)
    res = R.PCG_SOLVER(A2, 1e-8, 500).solvePCG(x, b, M)
    xo, ro = O2.pcg(b, 1e-8, 500)
    assert abs(res["iterations"] - ro["iterations"]) <= 1
    assert res["true_relres"] < 2e-8


@pytest.mark.parametrize("prec", [R.RSC_FP64, R.RSC_FP32])
def test_refactor_graph_replays_match_eager(prec):
    """Refactorizations after the first replay the numeric program as one
    captured graph.  Five refactorizations in a row, alternating two
    coefficient sets, with solves (eager work) between the replays: every
    replayed factor applies bitwise like a fresh, eagerly launched
    factorization of the same values, and like the eager launches of the
    same handle (RSC_FACTOR_GRAPH=0)."""
    A, xyz = R.poisson3d(18, 1.5, 1234)
    A2, _ = R.poisson3d(18, 1.5, 99)
    sopts = dict(max_dense_block=63, max_diagonal_block=50, explicit_block_threshold=120,
                 factor_precision=prec)
    b = R.uniform_rhs(A._nrow, 5678)

    def applied(M):
        y = np.empty_like(b)
        M.solve(y, b)
        return y

    want = {0: applied(_factor(A, xyz, **sopts)), 1: applied(_factor(A2, xyz, **sopts))}
    M = _factor(A, xyz, **sopts)
    assert np.array_equal(applied(M), want[0])
    mats = (A, A2)
    for rep in range(5):
        which = (rep + 1) % 2
        _lib.check(_lib.lib().rsc_factor_refactor(M.handle, _lib.dptr(mats[which]._x)))
        assert np.array_equal(applied(M), want[which]), rep
    old = os.environ.get("RSC_FACTOR_GRAPH")
    os.environ["RSC_FACTOR_GRAPH"] = "0"
    try:
        _lib.check(_lib.lib().rsc_factor_refactor(M.handle, _lib.dptr(A2._x)))
        assert np.array_equal(applied(M), want[1])
    finally:
        if old is None:
            del os.environ["RSC_FACTOR_GRAPH"]
        else:
            os.environ["RSC_FACTOR_GRAPH"] = old


# ---- indefinite-diagonal retry -----------------------------------------------

def _shifted(A, shift):
    """A - shift I on the same pattern (every diagonal entry is stored)."""
    x = A._x.copy()
    for j in range(A._nrow):
        lo, hi = A._p[j], A._p[j + 1]
        d = lo + int(np.nonzero(A._i[lo:hi] == j)[0][0])
        x[d] -= shift
    return R.SparseColumnMatrix(A._nrow, A._ncol, A._p, A._i, x)


@pytest.mark.parametrize("max_attempts", [3, 10])
def test_indefinite_retry_schedule_matches_oracle(max_attempts):
    """SparseSolver::factorNumeric's retry (SparseSolver.cpp:174-218) on a
    symmetric indefinite matrix (A - 3 I): every attempt meets an indefinite
    diagonal block, the diagonal rank constant grows by 1.25 per attempt,
    exactly max_attempts factorizations run, and the status is
    RSC_INDEFINITE -- the same attempts and final constant as the oracle's
    retry on the same ordering.  Afterwards the same handle refactors an SPD
    matrix with the raised constant kept (the reference keeps
    _diagonalRankConstant)."""
    from oracle.oracle import OracleIndefinite
    A, xyz = R.poisson3d(12, 1.0, 1234)
    Ai = _shifted(A, 3.0)
    sopts = dict(max_dense_block=20, max_diagonal_block=40, max_attempts=max_attempts)
    sym = symbolic(Ai, xyz, **sopts)
    with pytest.raises(OracleIndefinite) as oe:
        oracle_for(Ai, sym, max_attempts=max_attempts)
    assert oe.value.attempts == max_attempts
    # create on the SPD matrix (one attempt), refactor with indefinite values
    M = _factor(A, xyz, **dict(sopts))
    assert M.info()["attempts"] == 1
    assert M.info()["final_diagonal_rank_constant"] == 0.5
    h = M.handle
    st = _lib.lib().rsc_factor_refactor(h, _lib.dptr(Ai._x))
    assert st == _lib.RSC_INDEFINITE
    inf = M.info()
    assert inf["attempts"] == max_attempts == oe.value.attempts
    assert inf["final_diagonal_rank_constant"] == pytest.approx(0.5 * 1.25 ** (max_attempts - 1))
    assert inf["final_diagonal_rank_constant"] == pytest.approx(oe.value.final_constant)
    # refactor with SPD values: one attempt at the kept (raised) constant,
    # and the result equals the oracle run at that constant
    _lib.check(_lib.lib().rsc_factor_refactor(h, _lib.dptr(A._x)))
    inf = M.info()
    assert inf["attempts"] == 1
    assert inf["final_diagonal_rank_constant"] == pytest.approx(oe.value.final_constant)
    O = oracle_for(A, sym, diagonal_rank_constant=oe.value.final_constant)
    b = R.uniform_rhs(A._nrow, 5678)
    y = np.empty_like(b)
    M.solve(y, b)
    # separators of 20-40 columns: blocks whose singular values fall below
    # u^(1/3) within the rank, as in 2D (the reference method's noise floor)
    assert relerr(y, O.apply(b)) < 1e-6
    # a fresh create on the indefinite matrix fails the same way
    M2 = R.FactorPreconditioner()
    with pytest.raises(_lib.RscError) as e:
        M2.initialize(Ai, xyz, 20, 40, max_attempts=max_attempts)
    assert e.value.status == _lib.RSC_INDEFINITE


# ---- reference entry points ---------------------------------------------------

def test_pcg_pysolve_and_solve_match_oracle():
    """pcg_pysolve (pcg_pysolver.cpp:83-101: no positions, maxDenseBlock 200,
    maxDiagonalBlock 400, PCG to 1e-6 from zero) and solve(A, b)
    (rschol.pyx:38-49) against the oracle on the same ordering: the same
    iteration count (+-1) and a solution within the 1e-6 tolerance."""
    A, _ = R.poisson3d(24, 1.0, 1234)
    b = R.uniform_rhs(A._nrow, 5678)
    sym = symbolic(A, None, max_dense_block=200, max_diagonal_block=400,
                   reorder_separators=0)
    assert sym["compressed"].sum() > 0
    O = oracle_for(A, sym)
    xo, ro = O.pcg(b, 1e-6, 10000)
    x = R.pcg_pysolve(A._nrow, A._p, A._i, A._x, b)
    assert relerr(A.multiply(x), b) < 1e-6 * 1.5
    xs = R.solve(A, b)
    assert np.array_equal(xs, x)
    M = R.FactorPreconditioner()
    M.initialize(A, None, 200, 400, reorderSeparators=False)
    assert np.array_equal(M.permutation(), sym["perm"])
    S = R.PCG_SOLVER(A, 1e-6, 10000)
    x2 = np.zeros_like(b)
    S.solvePCG(x2, b, M)
    assert abs(S.last["iterations"] - ro["iterations"]) <= 1
    assert np.array_equal(x2, x)


def test_compressed_supernode_without_sketch_draws():
    """Regression: a compressed root separator that is one diagonal leaf with
    no off-diagonal rows draws no sketch entries; the factor must not expect
    a sketch stream (it threw before)."""
    A, xyz = R.poisson3d(10, 1.0, 1234)
    sym = symbolic(A, xyz, max_dense_block=60)
    assert sym["compressed"].sum() == 1 and sym["rng_count"] == 0
    O = oracle_for(A, sym)
    M = _factor(A, xyz, max_dense_block=60)
    b = R.uniform_rhs(A._nrow, 5678)
    y = np.empty_like(b)
    M.solve(y, b)
    assert relerr(y, O.apply(b)) < 1e-12


@pytest.mark.parametrize("dconst", [0.5, 3.725290298461914])
def test_full_rank_diagonal_sketches_stay_spd(dconst):
    """Regression: at a diagonal rank constant that asks for full rank (the
    tenth attempt of the retry, 0.5 * 1.25^9) the HODLR sketches are
    rank-deficient to rounding; the thin QR drops the dependent directions
    and the factor stays finite, SPD and equal to the oracle's (whose
    Householder QR completes them with arbitrary orthonormal directions,
    which V = L21 U maps to zero)."""
    A, xyz = R.poisson3d(12, 1.0, 1234)
    sopts = dict(max_dense_block=20, max_diagonal_block=40)
    sym = symbolic(A, xyz, **sopts)
    O = oracle_for(A, sym, diagonal_rank_constant=dconst)
    M = _factor(A, xyz, diagonal_rank_constant=dconst, max_attempts=1, **dict(sopts))
    b = R.uniform_rhs(A._nrow, 5678)
    y = np.empty_like(b)
    M.solve(y, b)
    assert np.all(np.isfinite(y))
    assert relerr(y, O.apply(b)) < 1e-6
    x = np.zeros_like(b)
    res = R.PCG_SOLVER(A, 1e-8, 500).solvePCG(x, b, M)
    assert res["true_relres"] < 2e-8
