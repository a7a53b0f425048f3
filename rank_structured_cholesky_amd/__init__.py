"""MI355X-native rank-structured supernodal Cholesky preconditioner + PCG.

Python mirror of the reference's operator interface (names, argument meaning
and error behaviour): ``FactorPreconditioner`` (src/rschol/solver/
FactorPreconditioner.h:60-139), ``PCG_SOLVER`` (src/rschol/linearalgebra/
PCG_SOLVER.h:49-122), ``SparseColumnMatrix`` (SPARSE_MATRIX.h:187-317) and
``solve(A, b)`` (src/python/rschol.pyx:38-49).  All numeric work runs in
librsc.so (gfx950 HIP kernels); importing works without a GPU, factoring
needs one and raises when the library or the device is missing.
"""
import ctypes as C
import gzip
import struct

import numpy as np

from . import _lib
from ._lib import (RSC_BF16, RSC_FP32, RSC_FP64, RscError, as_f64, as_i32, check,
                   default_options, dptr, iptr, lib)

# src/rschol/TYPES.h:134-165
CHOLESKY, LDL = 0, 1
EXTENDED_VARIABLE, IN_PLACE = 0, 1
DECOMPOSE_STANDARD, DECOMPOSE_TRANSPOSE = 0, 1
DECOMPOSE_SCHUR, DECOMPOSE_FACTOR = 0, 1

__all__ = ["SparseColumnMatrix", "FactorPreconditioner", "PCG_SOLVER", "solve",
           "read_reference_ordering", "write_reference_ordering", "nested_dissection",
           "poisson2d", "poisson3d", "elasticity3d", "uniform_rhs", "pcg_pysolve",
           "CHOLESKY", "LDL", "EXTENDED_VARIABLE", "IN_PLACE", "DECOMPOSE_STANDARD",
           "DECOMPOSE_TRANSPOSE", "DECOMPOSE_SCHUR", "DECOMPOSE_FACTOR",
           "RSC_FP64", "RSC_FP32", "RSC_BF16", "RscError", "default_options"]


class SparseColumnMatrix:
    """Full symmetric CSC matrix (_nrow, _ncol, _p, _i, _x)."""

    def __init__(self, nrow=0, ncol=0, p=None, i=None, x=None):
        self._nrow = int(nrow)
        self._ncol = int(ncol)
        self._p = as_i32(p if p is not None else np.zeros(ncol + 1))
        self._i = as_i32(i if i is not None else np.zeros(0))
        self._x = as_f64(x if x is not None else np.zeros(len(self._i)))

    @property
    def _nzmax(self):
        return int(self._p[-1]) if len(self._p) else 0

    @classmethod
    def from_scipy(cls, A):
        A = A.tocsc()
        A.sort_indices()
        return cls(A.shape[0], A.shape[1], A.indptr, A.indices, A.data)

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csc_matrix((self._x, self._i, self._p), shape=(self._nrow, self._ncol))

    def multiply(self, x):
        """y = A x on the host (reference SPARSE_MATRIX.cpp:353-369)."""
        return self.to_scipy() @ np.asarray(x, dtype=np.float64)

    # .bcsm: uint64 nnz, uint64 nrow, uint64 ncol, int32 p[ncol+1],
    # int32 i[nnz], float64 x[nnz] (SPARSE_MATRIX.cpp:1557-1718)
    def writeToBinary(self, path):
        opener = gzip.open if path.endswith(".gz") else open
        with opener(path, "wb") as f:
            f.write(struct.pack("<QQQ", self._nzmax, self._nrow, self._ncol))
            f.write(self._p.astype("<i4").tobytes())
            f.write(self._i[: self._nzmax].astype("<i4").tobytes())
            f.write(self._x[: self._nzmax].astype("<f8").tobytes())

    writeToBinaryGZ = writeToBinary

    @classmethod
    def readFromBinary(cls, path):
        opener = gzip.open if path.endswith(".gz") else open
        with opener(path, "rb") as f:
            nnz, nrow, ncol = struct.unpack("<QQQ", f.read(24))
            p = np.frombuffer(f.read(4 * (ncol + 1)), dtype="<i4")
            i = np.frombuffer(f.read(4 * nnz), dtype="<i4")
            x = np.frombuffer(f.read(8 * nnz), dtype="<f8")
        return cls(nrow, ncol, p, i, x)

    readFromBinaryGZ = readFromBinary


def read_vector(path):
    """.vector: int32 n, float64 x[n] (matlab/write_vector.m:6-7)."""
    with open(path, "rb") as f:
        (n,) = struct.unpack("<i", f.read(4))
        return np.frombuffer(f.read(8 * n), dtype="<f8").copy()


def write_vector(path, x):
    x = as_f64(x)
    with open(path, "wb") as f:
        f.write(struct.pack("<i", len(x)))
        f.write(x.astype("<f8").tobytes())


def read_positions(path):
    """meshPositions.dat: int32 count, count x 3 float64 (STLUtil.h:105-118)."""
    with open(path, "rb") as f:
        (n,) = struct.unpack("<i", f.read(4))
        return np.frombuffer(f.read(24 * n), dtype="<f8").reshape(n, 3).copy()


def write_positions(path, xyz):
    xyz = as_f64(xyz).reshape(-1, 3)
    with open(path, "wb") as f:
        f.write(struct.pack("<i", xyz.shape[0]))
        f.write(xyz.astype("<f8").tobytes())


# The reference's ordering dumps (readVector format: int32 count, then the
# records): SparseSolver_permutation.vector (SparseSolver.cpp:136, the final
# permutation, perm[new] = old), separatorParents.vector and
# separators.vector (NestedDissection.cpp:394-395; one {size, first, last,
# numChildren} int32 record per component, CHOLMOD_Environment.h:233-240).
_SEP_RECORD = np.dtype([("size", "<i4"), ("first", "<i4"), ("last", "<i4"),
                        ("numChildren", "<i4")])


def _read_records(path, dtype):
    with open(path, "rb") as f:
        (n,) = struct.unpack("<i", f.read(4))
        data = f.read(n * np.dtype(dtype).itemsize)
    if len(data) != n * np.dtype(dtype).itemsize:
        raise ValueError("truncated file " + path)
    return np.frombuffer(data, dtype=dtype).copy()


def _write_records(path, a):
    with open(path, "wb") as f:
        f.write(struct.pack("<i", len(a)))
        f.write(a.tobytes())


def read_reference_ordering(directory="."):
    """A reference run's ordering dumps -> (perm, cparent, cmember), the
    nested-dissection triple of FactorPreconditioner.initialize(ordering=)
    and rsc_factor_create_with_ordering.  cmember is rebuilt from the
    components' column ranges.  The dumped permutation already carries the
    reference's separator sub-ordering: factor it with
    reorderSeparators=False."""
    import os
    perm = _read_records(os.path.join(directory, "SparseSolver_permutation.vector"), "<i4")
    cparent = _read_records(os.path.join(directory, "separatorParents.vector"), "<i4")
    seps = _read_records(os.path.join(directory, "separators.vector"), _SEP_RECORD)
    n, nc = len(perm), len(cparent)
    if len(seps) != nc:
        raise ValueError("separator and parent counts differ")
    cmember = np.full(n, -1, np.int32)
    nxt = 0
    for c, s in enumerate(seps):
        if (s["first"] != nxt or s["last"] - s["first"] + 1 != s["size"] or s["last"] >= n
                or not -1 <= cparent[c] < nc):
            raise ValueError("separator ranges do not tile the permutation")
        cols = perm[s["first"]:s["last"] + 1]
        if np.any(cmember[cols] != -1):
            raise ValueError("not a permutation")
        cmember[cols] = c
        nxt = s["last"] + 1
    if nxt != n or np.any(cmember < 0):
        raise ValueError("separator ranges do not cover the permutation")
    return perm, cparent, cmember


def write_reference_ordering(directory, perm, cparent, cmember):
    """Writes (perm, cparent, cmember) as the reference's three dumps."""
    import os
    perm, cparent, cmember = as_i32(perm), as_i32(cparent), as_i32(cmember)
    comp = cmember[perm]
    if np.any(np.diff(comp) < 0):
        raise ValueError("components are not contiguous in perm")
    nc = len(cparent)
    seps = np.zeros(nc, _SEP_RECORD)
    counts = np.bincount(comp, minlength=nc)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    seps["size"] = counts
    seps["first"] = starts
    seps["last"] = starts + counts - 1
    seps["numChildren"] = np.bincount(cparent[cparent >= 0], minlength=nc)
    _write_records(os.path.join(directory, "SparseSolver_permutation.vector"), perm)
    _write_records(os.path.join(directory, "separatorParents.vector"), cparent)
    _write_records(os.path.join(directory, "separators.vector"), seps)


def nested_dissection(A, positions=None, nd_small=200):
    """This library's nested dissection: (perm, cparent, cmember) in
    CHOLMOD's format (CHOLMOD_Environment.cpp:326-432)."""
    n = A._nrow
    perm = np.empty(n, np.int32)
    cparent = np.empty(n, np.int32)
    cmember = np.empty(n, np.int32)
    nc = C.c_int()
    pos = None if positions is None else as_f64(np.asarray(positions).reshape(-1, 3))
    check(lib().rsc_nested_dissection(n, iptr(A._p), iptr(A._i), dptr(pos), int(nd_small),
                                      iptr(perm), C.byref(nc), iptr(cparent), iptr(cmember)))
    return perm, cparent[:nc.value].copy(), cmember


def _matrix_from_handle(h):
    L = lib()
    n = C.c_int()
    nnz = C.c_int64()
    check(L.rsc_matrix_info(h, C.byref(n), C.byref(nnz)))
    p = np.empty(n.value + 1, np.int32)
    i = np.empty(nnz.value, np.int32)
    x = np.empty(nnz.value, np.float64)
    xyz = np.empty(3 * n.value, np.float64)
    check(L.rsc_matrix_export(h, iptr(p), iptr(i), dptr(x), dptr(xyz)))
    L.rsc_matrix_destroy(h)
    return SparseColumnMatrix(n.value, n.value, p, i, x), xyz.reshape(-1, 3)


def poisson2d(N):
    """5-point Poisson on an N x N grid (Dirichlet eliminated)."""
    h = C.c_void_p()
    check(lib().rsc_matrix_poisson2d(N, C.byref(h)))
    return _matrix_from_handle(h)


def poisson3d(N, sigma=1.0, seed=1234):
    """7-point variable-coefficient Poisson on N^3 cells, log-normal faces."""
    h = C.c_void_p()
    check(lib().rsc_matrix_poisson3d(N, float(sigma), int(seed), C.byref(h)))
    return _matrix_from_handle(h)


def elasticity3d(NE, sigma=1.0, nu=0.3, seed=1234):
    """Trilinear-hex elasticity on NE^3 elements, x=0 face clamped."""
    h = C.c_void_p()
    check(lib().rsc_matrix_elasticity3d(NE, float(sigma), float(nu), int(seed),
                                        C.byref(h)))
    return _matrix_from_handle(h)


def uniform_rhs(count, seed=5678):
    out = np.empty(int(count), np.float64)
    check(lib().rsc_uniform_rhs(int(count), int(seed), dptr(out)))
    return out


class FactorPreconditioner:
    """Rank-structured factor preconditioner on the current HIP device."""

    def __init__(self):
        self._h = None
        self._matrix = None
        self._useTranspose = False
        self._symbolicFactorNeeded = True
        self._numericFactorNeeded = True
        self._positions = None
        self._options = None

    def __del__(self):
        # at interpreter shutdown the module's globals may already be gone;
        # the process then releases the device memory itself
        try:
            self.clearSystem()
        except (AttributeError, TypeError, ImportError):
            pass

    def clearSystem(self):
        if self._h:
            lib().rsc_factor_destroy(self._h)
        self._h = None
        self._symbolicFactorNeeded = True
        self._numericFactorNeeded = True

    def clearFactor(self):
        if not self._h:
            self._symbolicFactorNeeded = True
        self._numericFactorNeeded = True

    def setUseTranspose(self, useTranspose):
        self._useTranspose = bool(useTranspose)

    def initialize(self, matrix, meshPositions=None, maxDenseBlock=200,
                   maxDiagonalBlock=400, extFactorType=CHOLESKY,
                   compressionType=IN_PLACE, blockType=DECOMPOSE_TRANSPOSE,
                   decompType=DECOMPOSE_FACTOR, reorderSeparators=True, ordering=None,
                   **options):
        """Symbolic + numeric factorization (FactorPreconditioner.cpp:70-131).

        ordering: optional (perm, cparent, cmember) nested dissection in
        CHOLMOD's format (e.g. read_reference_ordering()) used instead of
        this library's own.  Extra keyword options set fields of rsc_options
        (e.g. factor_precision=RSC_FP32, max_rank=64,
        explicit_block_threshold=...)."""
        if not isinstance(matrix, SparseColumnMatrix):
            matrix = SparseColumnMatrix.from_scipy(matrix)
        o = default_options(max_dense_block=int(maxDenseBlock),
                            max_diagonal_block=int(maxDiagonalBlock),
                            extended_factor_type=int(extFactorType),
                            compression_type=int(compressionType),
                            decomposition_block=int(blockType),
                            decomposition_type=int(decompType),
                            reorder_separators=int(bool(reorderSeparators)), **options)
        pos = None
        if meshPositions is not None and len(meshPositions):
            pos = as_f64(np.asarray(meshPositions).reshape(-1, 3))
        if ordering is not None:
            perm, cparent, cmember = (as_i32(a) for a in ordering)
            if len(perm) != matrix._nrow or len(cmember) != matrix._nrow:
                raise ValueError("ordering size differs from the matrix")
            self.clearSystem()
            h = C.c_void_p()
            check(lib().rsc_factor_create_with_ordering(
                matrix._nrow, iptr(matrix._p), iptr(matrix._i), dptr(matrix._x), dptr(pos),
                iptr(perm), len(cparent), iptr(cparent), iptr(cmember), C.byref(o),
                C.byref(h)))
            self._h = h
        elif (self._h and not self._symbolicFactorNeeded and self._numericFactorNeeded
                and self._matrix is not None and self._matrix._nzmax == matrix._nzmax
                and np.array_equal(self._matrix._p, matrix._p)
                and np.array_equal(self._matrix._i, matrix._i)):
            check(lib().rsc_factor_refactor(self._h, dptr(matrix._x)))
        elif self._symbolicFactorNeeded or self._numericFactorNeeded or not self._h:
            self.clearSystem()
            h = C.c_void_p()
            check(lib().rsc_factor_create(matrix._nrow, iptr(matrix._p), iptr(matrix._i),
                                          dptr(matrix._x), dptr(pos), C.byref(o),
                                          C.byref(h)))
            self._h = h
        self._matrix = matrix
        self._positions = pos
        self._options = o
        self._symbolicFactorNeeded = False
        self._numericFactorNeeded = False

    def init(self, *args):
        """The factor is built by initialize(), as in the reference."""

    def solve(self, x, b):
        """x = M^{-1} b (host arrays; b may hold several columns)."""
        b = np.asarray(b, dtype=np.float64)
        B = np.asfortranarray(b.reshape(b.shape[0], -1))
        X = np.empty_like(B, order="F")
        check(lib().rsc_apply_host(self._h, B.shape[1], B.ctypes.data_as(_lib._DP),
                                   B.shape[0], X.ctypes.data_as(_lib._DP), X.shape[0]))
        x[...] = X.reshape(b.shape)
        return x

    def release_workspace(self):
        """Free the numeric factorization's workspace; solves keep working,
        and a later refactorization allocates it again."""
        check(lib().rsc_factor_release_workspace(self._h))

    def info(self):
        inf = _lib.FactorInfo()
        check(lib().rsc_factor_get_info(self._h, C.byref(inf)))
        return {k: getattr(inf, k) for k, _ in inf._fields_}

    def permutation(self):
        p = np.empty(self._matrix._nrow, np.int32)
        check(lib().rsc_factor_get_permutation(self._h, iptr(p)))
        return p

    @property
    def handle(self):
        return self._h


class PCG_SOLVER:
    """PCG with a FactorPreconditioner (PCG_SOLVER.cpp:226-363), on the device."""

    def __init__(self, A, eps=1e-4, iterations=1000):
        if not isinstance(A, SparseColumnMatrix):
            A = SparseColumnMatrix.from_scipy(A)
        self._A = A
        self._eps = float(eps)
        self._iterations = int(iterations)
        self._totalSolves = 0
        self._totalIterations = 0
        self._totalResidual = 0.0
        self.last = None

    def maxIterations(self):
        return self._iterations

    def eps(self):
        return self._eps

    def meanIterations(self):
        return self._totalIterations / max(1, self._totalSolves)

    def meanResidual(self):
        return self._totalResidual / max(1, self._totalSolves)

    def solvePCG(self, x, b, M):
        """Solve A x = b from the initial guess in x (one or more columns)."""
        if not isinstance(M, FactorPreconditioner) or M.handle is None:
            raise TypeError("M must be an initialized FactorPreconditioner")
        Mm, A = M._matrix, self._A
        if Mm is not A and not (Mm._nrow == A._nrow and np.array_equal(Mm._p, A._p)
                                and np.array_equal(Mm._i, A._i)
                                and np.array_equal(Mm._x, A._x)):
            # the device recurrence runs on the factor's copy of A
            raise ValueError("the preconditioner was initialized on a different matrix")
        b = np.asarray(b, dtype=np.float64)
        k = 1 if b.ndim == 1 else b.shape[1]
        B = np.asfortranarray(b.reshape(b.shape[0], k))
        X = np.asfortranarray(np.asarray(x, dtype=np.float64).reshape(b.shape[0], k))
        res = (_lib.PcgResult * k)()
        st = lib().rsc_block_pcg_solve_host(M.handle, k, B.ctypes.data_as(_lib._DP),
                                            B.shape[0], X.ctypes.data_as(_lib._DP),
                                            X.shape[0], self._eps, self._iterations, res)
        if st not in (_lib.RSC_OK, _lib.RSC_NOT_CONVERGED):
            check(st)
        x[...] = X.reshape(x.shape)
        out = [dict(iterations=r.iterations, recurrence_relres=r.recurrence_relres,
                    true_relres=r.true_relres, seconds=r.seconds) for r in res]
        for r in out:
            self._totalSolves += 1
            self._totalIterations += r["iterations"]
            self._totalResidual += r["recurrence_relres"]
        self.last = out if k > 1 else out[0]
        return self.last


class RequestServer:
    """Continuous-batching PCG server over factors resident in HBM (the C
    ABI rsc_server_*).  Right-hand sides and solutions are DEVICE buffers
    given by address (e.g. a torch tensor's data_ptr()); each must stay
    alive until wait() returns for its ticket.  The factors must not be used
    elsewhere while the server exists."""

    def __init__(self, factors, eps=1e-8, max_iters=10000, slots=64, memory_bytes=0,
                 release_workspace=True):
        """slots: most column slots of one factor; memory_bytes: device
        memory for the slots of all factors (0: half of the free memory),
        divided among the factors by their bytes per column;
        release_workspace: free the factors' numeric workspaces (they are
        not refactorized while the server exists)."""
        self._factors = list(factors)  # keep the handles alive
        hs = (C.c_void_p * len(self._factors))(*[M.handle for M in self._factors])
        o = _lib.ServerOptions(float(eps), int(max_iters), int(slots), int(memory_bytes),
                               int(bool(release_workspace)), 0)
        h = C.c_void_p()
        check(lib().rsc_server_create(hs, len(self._factors), C.byref(o), C.byref(h)))
        self._h = h

    def submit(self, factor, nrhs, B_ptr, ldb, X_ptr, ldx):
        t = C.c_int64()
        check(lib().rsc_server_submit(self._h, int(factor), int(nrhs), C.c_void_p(B_ptr),
                                      int(ldb), C.c_void_p(X_ptr), int(ldx), C.byref(t)))
        return t.value

    def wait(self, ticket, nrhs):
        res = (_lib.PcgResult * nrhs)()
        st = lib().rsc_server_wait(self._h, int(ticket), res)
        if st not in (_lib.RSC_OK, _lib.RSC_NOT_CONVERGED):
            check(st)
        return [dict(iterations=r.iterations, recurrence_relres=r.recurrence_relres,
                     true_relres=r.true_relres, seconds=r.seconds) for r in res]

    def stats(self):
        steps, busy = C.c_int64(), C.c_int()
        check(lib().rsc_server_stats(self._h, C.byref(steps), C.byref(busy)))
        return dict(steps=steps.value, max_busy=busy.value)

    def slots(self, factor):
        """Column slots the memory budget gave factor `factor`."""
        k = C.c_int()
        check(lib().rsc_server_slots(self._h, int(factor), C.byref(k)))
        return k.value

    def close(self):
        if self._h:
            lib().rsc_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def solve(A, b, positions=None, eps=1e-6, iterations=10000, **options):
    """rschol.solve(A, b) counterpart (rschol.pyx:38-49): factor with
    maxDenseBlock=200 / maxDiagonalBlock=400 and run PCG from zero."""
    M = FactorPreconditioner()
    M.initialize(A, positions, 200, 400, reorderSeparators=positions is not None,
                 **options)
    x = np.zeros_like(np.asarray(b, dtype=np.float64))
    PCG_SOLVER(A, eps, iterations).solvePCG(x, b, M)
    return x


def pcg_pysolve(nrow, jc, ir, pr, rhs):
    """The reference's C entry point pcg_pysolve (pcg_pysolver.h:35-36)."""
    jc, ir, pr, rhs = as_i32(jc), as_i32(ir), as_f64(pr), as_f64(rhs)
    sol = np.zeros(nrow, np.float64)
    lib().pcg_pysolve(int(nrow), iptr(jc), iptr(ir), dptr(pr), dptr(rhs), dptr(sol))
    return sol
