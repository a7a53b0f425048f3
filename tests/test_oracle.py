"""Pins the CPU oracle (no GPU): the sketch generator against the MT19937
known answer and an independent MT19937 (numpy's RandomState), the
Box-Muller step against its formula (MERSENNETWISTER.h:158-177), the factor
against exact sparse solves (no compression, and compression at full rank,
where every low-rank block is exact), SPD-ness of the compressed
preconditioner, truncation error against the SVD optimum, and PCG against
scipy's conjugate gradients with the same preconditioner."""
import ctypes as C
import json
import os

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse.linalg as spla

import rank_structured_cholesky_amd as R
from rank_structured_cholesky_amd import _lib
from helpers import oracle_for, relerr, symbolic
from oracle.oracle import gaussian, mt_output

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_mt19937_known_answer():
    # 10000th output of the reference generator for seed 5489
    assert mt_output(5489, 10000) == 4123659995
    rs = np.random.RandomState(5489)
    assert int(rs.randint(0, 2 ** 32, size=10000, dtype=np.uint64)[-1]) == 4123659995


def test_mt19937_seed152_against_independent_generator_and_golden():
    gold = json.load(open(os.path.join(GOLD, "gaussian_seed152.json")))
    rs = np.random.RandomState(152)  # init_genrand seeding, as MERSENNETWISTER.h:285-297
    raw = rs.randint(0, 2 ** 32, size=64, dtype=np.uint64).tolist()
    assert raw == gold["mt_outputs"]
    assert [mt_output(152, k + 1) for k in range(64)] == raw


def box_muller(raw):
    """randNorm: r from randDblExc, phi from randExc, consecutive draws."""
    a = np.asarray(raw[0::2], dtype=np.float64)
    b = np.asarray(raw[1::2], dtype=np.float64)
    u1 = (a + 0.5) * (1.0 / 4294967296.0)
    r = np.sqrt(-2.0 * np.log(1.0 - u1))
    phi = 2.0 * np.pi * (b * (1.0 / 4294967296.0))
    return r * np.cos(phi)


def test_box_muller_matches_formula_and_device_stream():
    gold = json.load(open(os.path.join(GOLD, "gaussian_seed152.json")))
    rs = np.random.RandomState(152)
    raw = rs.randint(0, 2 ** 32, size=4096, dtype=np.uint64)
    want = box_muller(raw)
    got = gaussian(152, 2048)
    assert np.max(np.abs(got - want)) < 1e-13
    assert np.array_equal(got[:64], np.asarray(gold["normals"]))
    dev = np.empty(2048)
    _lib.check(_lib.lib().rsc_gaussian_stream(152, 2048, _lib.dptr(dev)))
    assert np.array_equal(dev, got)
    # sanity of the distribution over a longer stream
    g = gaussian(7, 200000)
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1.0) < 0.01


def test_threaded_stream_matches_oracle_across_chunks():
    # the library draws raw pairs on one thread into a ring of 4 x threads
    # chunks of 2^16 entries and transforms them on the others; a stream long
    # enough to wrap the ring at 16 threads, with a ragged tail, must give the
    # oracle's single-threaded stream bit for bit
    n = 5 * (1 << 20) + 12345
    dev = np.empty(n)
    _lib.check(_lib.lib().rsc_gaussian_stream(152, n, _lib.dptr(dev)))
    assert np.array_equal(dev, gaussian(152, n))
    small = np.empty(5)
    _lib.check(_lib.lib().rsc_gaussian_stream(152, 5, _lib.dptr(small)))
    assert np.array_equal(small, dev[:5])


def exact_solve(A, b):
    return spla.spsolve(A.to_scipy().tocsc(), b)


@pytest.mark.parametrize("build", [lambda: R.poisson3d(8, 1.0, 3), lambda: R.poisson2d(20),
                                   lambda: R.elasticity3d(3, 1.0, 0.3, 1)])
def test_uncompressed_factor_is_exact_cholesky(build):
    A, xyz = build()
    n = A._nrow
    sym = symbolic(A, xyz, max_dense_block=n + 1)
    assert not sym["compressed"].any()
    o = oracle_for(A, sym)
    b = np.random.default_rng(0).standard_normal(n)
    assert relerr(o.apply(b), exact_solve(A, b)) < 1e-12
    x, info = o.pcg(b, 1e-10, 50)
    assert info["iterations"] <= 1 and info["true_relres"] < 1e-10


FULL = dict(rank_constant=1e4, diagonal_rank_constant=1e4)


@pytest.mark.parametrize("name,build,mdb,mdiag,thr", [
    ("p3d12_hodlr", lambda: R.poisson3d(12, 1.0, 3), 20, 40, 0),
    ("p3d12_explicit", lambda: R.poisson3d(12, 1.0, 3), 20, 40, 80),
    ("p2d40", lambda: R.poisson2d(40), 10, 12, 0),
    ("el6", lambda: R.elasticity3d(6, 1.0, 0.3, 2), 30, 40, 0),
])
def test_full_rank_compression_is_exact(name, build, mdb, mdiag, thr):
    """With every rank at min(rows, cols) the range finders capture each
    block exactly, so the rank-structured factor is the exact Cholesky
    factor: this checks the sketch, QR, low-rank updates and the HODLR
    solves without the truncation."""
    A, xyz = build()
    n = A._nrow
    sym = symbolic(A, xyz, max_dense_block=mdb, max_diagonal_block=mdiag,
                   explicit_block_threshold=thr, **FULL)
    assert sym["compressed"].any()
    assert any(len(b) > 2 for b, c in zip(sym["diag_blocks"], sym["compressed"]) if c)
    o = oracle_for(A, sym, explicit_threshold=thr, **FULL)
    b = np.random.default_rng(1).standard_normal(n)
    assert relerr(o.apply(b), exact_solve(A, b)) < 1e-9


def _dense_M(o, n):
    M = np.empty((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        M[:, j] = o.apply(e)
    return M


def test_compressed_preconditioner_is_spd_and_close():
    A, xyz = R.poisson3d(10, 1.0, 4)
    n = A._nrow
    sym = symbolic(A, xyz, max_dense_block=20, max_diagonal_block=64)
    assert sym["compressed"].any()
    o = oracle_for(A, sym)
    M = _dense_M(o, n)
    assert np.linalg.norm(M - M.T) <= 1e-12 * np.linalg.norm(M)
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > 0
    Ad = A.to_scipy().toarray()
    ev = np.linalg.eigvals(M @ Ad).real
    assert ev.min() > 0.3 and ev.max() < 3.0  # a useful preconditioner


def test_truncation_error_near_svd_optimum():
    """Lowest compressed supernodes (only dense descendants, exact diagonal):
    the randomized range finder with one power iteration stays within a
    small factor of the best rank-(r - oversampling) approximation."""
    A, xyz = R.poisson3d(16, 1.0, 6)
    n = A._nrow
    sym = symbolic(A, xyz, max_dense_block=63, diagonal_rank_constant=1e4)
    o = oracle_for(A, sym, diagonal_rank_constant=1e4)
    P = sym["perm"]
    Ad = A.to_scipy().toarray()[np.ix_(P, P)]
    L = np.linalg.cholesky(Ad)
    first, comp, par = sym["first"], sym["compressed"], sym["parent"]
    has_comp_desc = np.zeros(len(comp), bool)
    for s in range(len(comp)):
        a = par[s]
        while a >= 0:
            if comp[s]:
                has_comp_desc[a] = True
            a = par[a]
    checked = 0
    for s in np.nonzero(comp & ~has_comp_desc)[0]:
        a, b = first[s], first[s + 1]
        rows = sym["rows"][sym["rows_ptr"][s]:sym["rows_ptr"][s + 1]]
        if len(rows) == 0:
            continue
        L21 = L[np.ix_(rows, np.arange(a, b))]
        W = o.offdiag_dense(s, b - a)
        r = o.rank(s)
        sv = np.linalg.svd(L21, compute_uv=False)
        k = max(r - 5, 1)
        opt = np.sqrt(np.sum(sv[k:] ** 2))
        err = np.linalg.norm(L21 - W)
        assert err <= 3.0 * opt + 1e-12 * sv[0], (s, err, opt)
        checked += 1
    assert checked >= 2


def test_pcg_matches_scipy_cg_with_same_preconditioner():
    A, xyz = R.poisson3d(14, 1.0, 2)
    n = A._nrow
    sym = symbolic(A, xyz, max_dense_block=40)
    o = oracle_for(A, sym)
    b = R.uniform_rhs(n, 3)
    x, info = o.pcg(b, 1e-8, 200)
    its = []
    Msp = spla.LinearOperator((n, n), matvec=o.apply)
    xs, code = spla.cg(A.to_scipy().tocsr(), b, rtol=1e-8, atol=0.0, M=Msp, maxiter=200,
                       callback=lambda xk: its.append(1))
    assert code == 0
    assert abs(info["iterations"] - len(its)) <= 1
    assert relerr(x, xs) < 1e-6
    assert info["true_relres"] < 1e-7
    # the recurrence residual is what stops the loop (PCG_SOLVER.cpp:226-363)
    assert info["recurrence_relres"] <= 1e-8


def test_unpreconditioned_spmv():
    A, xyz = R.poisson3d(6, 1.0, 2)
    sym = symbolic(A, xyz, max_dense_block=10 ** 6)
    o = oracle_for(A, sym)
    x = np.random.default_rng(5).standard_normal(A._nrow)
    assert relerr(o.spmv(x), A.to_scipy() @ x) < 1e-14


def test_oracle_retry_schedule_on_indefinite_matrix():
    """The oracle's restatement of the indefinite-diagonal retry
    (SparseSolver.cpp:174-218): on A - 3 I (symmetric indefinite: a pivot
    fails at every rank) exactly max_attempts factorizations run and the
    diagonal rank constant of the last one is 0.5 * 1.25^(attempts-1); on
    the SPD matrix one attempt at 0.5."""
    from oracle.oracle import OracleIndefinite
    A, xyz = R.poisson3d(10, 1.0, 1234)
    x = A._x.copy()
    for j in range(A._nrow):
        lo, hi = A._p[j], A._p[j + 1]
        x[lo + int(np.nonzero(A._i[lo:hi] == j)[0][0])] -= 3.0
    Ai = R.SparseColumnMatrix(A._nrow, A._ncol, A._p, A._i, x)
    sym = symbolic(A, xyz, max_dense_block=20, max_diagonal_block=40)
    for attempts in (1, 4, 10):
        with pytest.raises(OracleIndefinite) as e:
            oracle_for(Ai, sym, max_attempts=attempts)
        assert e.value.attempts == attempts
        assert e.value.final_constant == pytest.approx(0.5 * 1.25 ** (attempts - 1))
    O = oracle_for(A, sym)
    assert O.attempts == 1 and O.final_constant == 0.5
