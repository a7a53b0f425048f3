"""ORACLE -- test infrastructure only (tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py).  ctypes wrapper of librsc_oracle.so, the FP64
CPU restatement of the reference's numeric factor, solves and PCG (see the
header of rsc_oracle.cpp for the reference file:line each part follows)."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "librsc_oracle.so")


class OracleOpts(C.Structure):
    _fields_ = [("rank_constant", C.c_double), ("diagonal_rank_constant", C.c_double),
                ("oversampling", C.c_int), ("min_rank", C.c_int), ("max_rank", C.c_int),
                ("power_iterations", C.c_int), ("explicit_threshold", C.c_int),
                ("seed", C.c_uint32), ("max_attempts", C.c_int),
                ("retry_alpha", C.c_double)]


_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)
_lib = None


def lib(path=None):
    """The oracle library (path: another build of the same source, e.g. the
    -march=native build of the CPU baseline; loaded in its place)."""
    global _lib
    if path is not None and _lib is not None and getattr(_lib, "_path", None) != path:
        raise RuntimeError("another oracle build is already loaded")
    if _lib is None:
        if path is None and not os.path.exists(LIB):
            from . import build as _b
            _b.build(verbose=False)
        L = C.CDLL(path or LIB)
        L._path = path
        L.oracle_factor.restype = C.c_int
        L.oracle_factor.argtypes = [C.c_int, _IP, _IP, _DP, _IP, C.c_int, _IP, _IP, _IP,
                                    _IP, C.POINTER(OracleOpts), C.POINTER(C.c_void_p),
                                    _IP, _DP]
        L.oracle_attempts.restype = C.c_int
        L.oracle_attempts.argtypes = [C.c_void_p]
        L.oracle_apply.restype = None
        L.oracle_apply.argtypes = [C.c_void_p, _DP, _DP]
        L.oracle_spmv.restype = None
        L.oracle_spmv.argtypes = [C.c_void_p, _DP, _DP]
        L.oracle_pcg.restype = C.c_int
        L.oracle_pcg.argtypes = [C.c_void_p, _DP, _DP, C.c_double, C.c_int, _DP, _DP]
        L.oracle_offdiag.restype = C.c_int
        L.oracle_offdiag.argtypes = [C.c_void_p, C.c_int, _DP, C.c_int]
        L.oracle_rank.restype = C.c_int
        L.oracle_rank.argtypes = [C.c_void_p, C.c_int]
        L.oracle_destroy.restype = None
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_gaussian.restype = None
        L.oracle_gaussian.argtypes = [C.c_uint32, C.c_int64, _DP]
        L.oracle_block_pcg.restype = None
        L.oracle_block_pcg.argtypes = [C.c_void_p, C.c_int, _DP, _DP, C.c_double, C.c_int,
                                       _IP, _DP, _DP]
        L.oracle_set_threads.restype = None
        L.oracle_set_threads.argtypes = [C.c_int]
        L.oracle_max_threads.restype = C.c_int
        L.oracle_max_threads.argtypes = []
        L.oracle_mt_output.restype = C.c_uint32
        L.oracle_mt_output.argtypes = [C.c_uint32, C.c_int64]
        _lib = L
    return _lib


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_IP)


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_DP)


class OracleIndefinite(RuntimeError):
    """Every attempt of the retry met an indefinite diagonal block."""

    def __init__(self, attempts, final_constant):
        super().__init__("oracle factorization indefinite after %d attempts" % attempts)
        self.attempts = attempts
        self.final_constant = final_constant


class OracleFactor:
    """FP64 CPU factor for a given ordering (perm, supernodes, diag blocks)."""

    def __init__(self, A, perm, first, compressed, diag_blocks, rank_constant=0.4,
                 diagonal_rank_constant=0.5, oversampling=5, min_rank=16, max_rank=0,
                 power_iterations=1, explicit_threshold=4000, seed=152,
                 max_attempts=10, retry_alpha=1.25):
        self.n = A._nrow
        ptr = [0]
        offs = []
        for b in diag_blocks:
            offs.extend(list(b))
            ptr.append(len(offs))
        o = OracleOpts(rank_constant, diagonal_rank_constant, oversampling, min_rank,
                       max_rank, power_iterations, explicit_threshold, seed, max_attempts,
                       retry_alpha)
        keep = [_i(A._p), _i(A._i), _d(A._x), _i(perm), _i(first), _i(compressed),
                _i(ptr), _i(offs if offs else [0])]
        self._keep = keep
        h = C.c_void_p()
        att = C.c_int()
        fc = C.c_double()
        st = lib().oracle_factor(self.n, keep[0][1], keep[1][1], keep[2][1], keep[3][1],
                                 len(first) - 1, keep[4][1], keep[5][1], keep[6][1],
                                 keep[7][1], C.byref(o), C.byref(h), C.byref(att),
                                 C.byref(fc))
        self.final_constant = fc.value
        if st != 0:
            raise OracleIndefinite(att.value, fc.value)
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_destroy(self._h)
            self._h = None

    @property
    def attempts(self):
        return lib().oracle_attempts(self._h)

    def apply(self, b):
        b, bp = _d(b)
        x = np.empty(self.n)
        lib().oracle_apply(self._h, bp, x.ctypes.data_as(_DP))
        return x

    def spmv(self, x):
        x, xp = _d(x)
        y = np.empty(self.n)
        lib().oracle_spmv(self._h, xp, y.ctypes.data_as(_DP))
        return y

    def pcg(self, b, eps, iters, x0=None):
        b, bp = _d(b)
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        rr = C.c_double()
        tr = C.c_double()
        it = lib().oracle_pcg(self._h, bp, x.ctypes.data_as(_DP), eps, iters,
                              C.byref(rr), C.byref(tr))
        return x, dict(iterations=it, recurrence_relres=rr.value, true_relres=tr.value)

    def block_pcg(self, B, eps, iters):
        """k independent PCG solves from zero (columns of B, n x k)."""
        B = np.asfortranarray(B, dtype=np.float64)
        k = B.shape[1]
        X = np.zeros_like(B, order="F")
        it = np.zeros(k, np.int32)
        rr = np.zeros(k)
        tr = np.zeros(k)
        lib().oracle_block_pcg(self._h, k, B.ctypes.data_as(_DP), X.ctypes.data_as(_DP),
                               eps, iters, it.ctypes.data_as(_IP), rr.ctypes.data_as(_DP),
                               tr.ctypes.data_as(_DP))
        return X, dict(iterations=it, recurrence_relres=rr, true_relres=tr)

    def offdiag(self, s):
        R = lib().oracle_offdiag(self._h, s, None, 0)
        return R

    def offdiag_dense(self, s, ncols):
        R = lib().oracle_offdiag(self._h, s, None, 0)
        out = np.zeros((ncols, R)).T.copy(order="F") if R else np.zeros((0, ncols))
        buf = np.zeros(R * ncols)
        lib().oracle_offdiag(self._h, s, buf.ctypes.data_as(_DP), R * ncols)
        return buf.reshape((ncols, R)).T if R else out

    def rank(self, s):
        return lib().oracle_rank(self._h, s)


def gaussian(seed, count):
    out = np.empty(int(count))
    lib().oracle_gaussian(seed, int(count), out.ctypes.data_as(_DP))
    return out


def set_threads(n):
    lib().oracle_set_threads(int(n))


def max_threads():
    return lib().oracle_max_threads()


def mt_output(seed, index):
    return lib().oracle_mt_output(seed, int(index))
