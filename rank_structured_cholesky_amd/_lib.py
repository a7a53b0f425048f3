"""ctypes binding of librsc.so (the C ABI declared in include/rsc.h)."""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librsc.so")

RSC_OK, RSC_INDEFINITE, RSC_OOM, RSC_BAD_INPUT, RSC_UNSUPPORTED_MODE, \
    RSC_DEVICE_ERROR, RSC_NOT_CONVERGED = range(7)
RSC_FP64, RSC_FP32, RSC_BF16 = 0, 1, 2


class Options(C.Structure):
    _fields_ = [
        ("max_dense_block", C.c_int), ("max_diagonal_block", C.c_int),
        ("extended_factor_type", C.c_int), ("compression_type", C.c_int),
        ("decomposition_block", C.c_int), ("decomposition_type", C.c_int),
        ("reorder_separators", C.c_int), ("explicit_block_threshold", C.c_int),
        ("rank_constant", C.c_double), ("diagonal_rank_constant", C.c_double),
        ("oversampling", C.c_int), ("min_rank", C.c_int), ("max_rank", C.c_int),
        ("power_iterations", C.c_int), ("seed", C.c_uint32),
        ("max_attempts", C.c_int), ("retry_alpha", C.c_double),
        ("nd_small", C.c_int), ("factor_precision", C.c_int),
        ("sketch_precision", C.c_int), ("verbose", C.c_int),
    ]


class PcgResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("recurrence_relres", C.c_double),
                ("true_relres", C.c_double), ("seconds", C.c_double)]


class ServerOptions(C.Structure):
    _fields_ = [("eps", C.c_double), ("max_iters", C.c_int), ("slots", C.c_int),
                ("memory_bytes", C.c_int64), ("release_workspace", C.c_int),
                ("reserved", C.c_int)]


class FactorInfo(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("nsuper", C.c_int), ("ncompressed", C.c_int),
        ("nlevels", C.c_int), ("attempts", C.c_int),
        ("final_diagonal_rank_constant", C.c_double),
        ("factor_bytes", C.c_int64), ("workspace_bytes", C.c_int64),
        ("symbolic_seconds", C.c_double), ("numeric_seconds", C.c_double),
        ("factor_flops", C.c_double), ("apply_bytes", C.c_double),
        ("spmv_bytes", C.c_double), ("pivot_replacements", C.c_int),
    ]


_P = C.c_void_p
_I = C.c_int
_I64 = C.c_int64
_D = C.c_double
_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)
_I64P = C.POINTER(C.c_int64)

# name -> (restype, argtypes); every symbol of include/rsc.h
SIGNATURES = {
    "rsc_options_default": (_I, [C.POINTER(Options)]),
    "rsc_factor_create": (_I, [_I, _IP, _IP, _DP, _DP, C.POINTER(Options), C.POINTER(_P)]),
    "rsc_factor_create_with_ordering": (_I, [_I, _IP, _IP, _DP, _DP, _IP, _I, _IP, _IP,
                                             C.POINTER(Options), C.POINTER(_P)]),
    "rsc_factor_refactor": (_I, [_P, _DP]),
    "rsc_apply": (_I, [_P, _I, _P, _I, _P, _I, _P]),
    "rsc_pcg_solve": (_I, [_P, _P, _P, _D, _I, C.POINTER(PcgResult), _P]),
    "rsc_block_pcg_solve": (_I, [_P, _I, _P, _I, _P, _I, _D, _I, C.POINTER(PcgResult), _P]),
    "rsc_apply_host": (_I, [_P, _I, _DP, _I, _DP, _I]),
    "rsc_pcg_solve_host": (_I, [_P, _DP, _DP, _D, _I, C.POINTER(PcgResult)]),
    "rsc_block_pcg_solve_host": (_I, [_P, _I, _DP, _I, _DP, _I, _D, _I,
                                      C.POINTER(PcgResult)]),
    "rsc_apply_status": (_I, [_P, _P]),
    "rsc_apply_noalloc": (_I, [_P, _I, _P, _I, _P, _I, _P]),
    "rsc_spmv": (_I, [_P, _P, _P, _P]),
    "rsc_factor_get_info": (_I, [_P, C.POINTER(FactorInfo)]),
    "rsc_factor_get_permutation": (_I, [_P, _IP]),
    "rsc_factor_get_supernodes": (_I, [_P, _IP, _IP]),
    "rsc_factor_destroy": (_I, [_P]),
    "rsc_last_error": (C.c_char_p, []),
    "rsc_symbolic_create": (_I, [_I, _IP, _IP, _DP, C.POINTER(Options), C.POINTER(_P)]),
    "rsc_symbolic_create_with_ordering": (_I, [_I, _IP, _IP, _DP, _IP, _I, _IP, _IP,
                                               C.POINTER(Options), C.POINTER(_P)]),
    "rsc_symbolic_sizes": (_I, [_P, _IP, _I64P, _IP, _I64P, _I64P]),
    "rsc_symbolic_get": (_I, [_P, _IP, _IP, _IP, _IP, _I64P, _IP, _IP, _IP]),
    "rsc_symbolic_diag_blocks": (_I, [_P, _I, _IP, _IP]),
    "rsc_symbolic_destroy": (None, [_P]),
    "rsc_nested_dissection": (_I, [_I, _IP, _IP, _DP, _I, _IP, _IP, _IP, _IP]),
    "rsc_etree_colcounts": (_I, [_I, _IP, _IP, _IP, _IP, _IP]),
    "rsc_relaxed_supernodes": (_I, [_I, _IP, _IP, _I, _IP, _IP, _IP, _IP, _IP, _I, _IP, _IP,
                                    _IP]),
    "rsc_separator_tree": (_I, [_I, _DP, _I, _IP, _IP, _IP]),
    "rsc_gaussian_stream": (_I, [C.c_uint32, _I64, _DP]),
    "rsc_block_rank": (_I, [_I, _I, _D, _I, _I, _I]),
    "rsc_matrix_poisson2d": (_I, [_I, C.POINTER(_P)]),
    "rsc_matrix_poisson3d": (_I, [_I, _D, C.c_uint64, C.POINTER(_P)]),
    "rsc_matrix_elasticity3d": (_I, [_I, _D, _D, C.c_uint64, C.POINTER(_P)]),
    "rsc_matrix_info": (_I, [_P, _IP, _I64P]),
    "rsc_matrix_export": (_I, [_P, _IP, _IP, _DP, _DP]),
    "rsc_matrix_destroy": (None, [_P]),
    "rsc_uniform_rhs": (_I, [_I64, C.c_uint64, _DP]),
    "rsc_selftest_gemm": (_I, [_I, _I, _I, _I, _I, _I, _D, _DP, _I, _DP, _I, _D, _DP, _I]),
    "pcg_pysolve": (None, [_I, _IP, _IP, _DP, _DP, _DP]),
    "rsc_server_options_default": (_I, [C.POINTER(ServerOptions)]),
    "rsc_server_slots": (_I, [_P, _I, C.POINTER(_I)]),
    "rsc_factor_release_workspace": (_I, [_P]),
    "rsc_server_create": (_I, [C.POINTER(_P), _I, C.POINTER(ServerOptions), C.POINTER(_P)]),
    "rsc_server_submit": (_I, [_P, _I, _I, _P, _I, _P, _I, _I64P]),
    "rsc_server_wait": (_I, [_P, _I64, C.POINTER(PcgResult)]),
    "rsc_server_stats": (_I, [_P, _I64P, _IP]),
    "rsc_server_destroy": (None, [_P]),
}

_lib = None


class RscError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"rsc status {status}: {msg}")
        self.status = status


def lib():
    """Load librsc.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} not built: run `python -m rank_structured_cholesky_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != RSC_OK:
        raise RscError(status, lib().rsc_last_error().decode())
    return status


def iptr(a):
    return a.ctypes.data_as(_IP) if a is not None else None


def dptr(a):
    return a.ctypes.data_as(_DP) if a is not None else None


def i64ptr(a):
    return a.ctypes.data_as(_I64P) if a is not None else None


def default_options(**kw):
    o = Options()
    check(lib().rsc_options_default(C.byref(o)))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError(k)
        setattr(o, k, v)
    return o


def as_i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def as_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
