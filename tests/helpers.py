"""Shared test helpers: symbolic analysis through the C ABI and the oracle
factor built on the same ordering."""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import rank_structured_cholesky_amd as R  # noqa: E402
from rank_structured_cholesky_amd import _lib  # noqa: E402


def symbolic(A, xyz=None, ordering=None, **opts):
    """Run rsc_symbolic_create and return its arrays."""
    L = _lib.lib()
    o = _lib.default_options(**opts)
    h = C.c_void_p()
    pos = None if xyz is None else _lib.as_f64(np.asarray(xyz).reshape(-1, 3))
    if ordering is None:
        _lib.check(L.rsc_symbolic_create(A._nrow, _lib.iptr(A._p), _lib.iptr(A._i),
                                         _lib.dptr(pos), C.byref(o), C.byref(h)))
    else:
        perm, cparent, cmember = (_lib.as_i32(a) for a in ordering)
        _lib.check(L.rsc_symbolic_create_with_ordering(
            A._nrow, _lib.iptr(A._p), _lib.iptr(A._i), _lib.dptr(pos), _lib.iptr(perm),
            len(cparent), _lib.iptr(cparent), _lib.iptr(cmember), C.byref(o),
            C.byref(h)))
    try:
        ns = C.c_int()
        tr = C.c_int64()
        nl = C.c_int()
        fe = C.c_int64()
        rc = C.c_int64()
        L.rsc_symbolic_sizes(h, C.byref(ns), C.byref(tr), C.byref(nl), C.byref(fe),
                             C.byref(rc))
        n = A._nrow
        s = ns.value
        out = dict(perm=np.empty(n, np.int32), first=np.empty(s + 1, np.int32),
                   compressed=np.empty(s, np.int32), parent=np.empty(s, np.int32),
                   rows_ptr=np.empty(s + 1, np.int64),
                   rows=np.empty(max(tr.value, 1), np.int32),
                   rank=np.empty(s, np.int32), level=np.empty(s, np.int32))
        _lib.check(L.rsc_symbolic_get(h, _lib.iptr(out["perm"]), _lib.iptr(out["first"]),
                                      _lib.iptr(out["compressed"]),
                                      _lib.iptr(out["parent"]),
                                      _lib.i64ptr(out["rows_ptr"]),
                                      _lib.iptr(out["rows"]), _lib.iptr(out["rank"]),
                                      _lib.iptr(out["level"])))
        blocks = []
        for j in range(s):
            cnt = C.c_int()
            L.rsc_symbolic_diag_blocks(h, j, C.byref(cnt), None)
            if cnt.value:
                offs = np.empty(cnt.value + 1, np.int32)
                L.rsc_symbolic_diag_blocks(h, j, C.byref(cnt), _lib.iptr(offs))
                blocks.append(offs)
            else:
                blocks.append(np.zeros(0, np.int32))
        out["diag_blocks"] = blocks
        out["nlevels"] = nl.value
        out["factor_elems"] = fe.value
        out["rng_count"] = rc.value
        return out
    finally:
        L.rsc_symbolic_destroy(h)


def oracle_for(A, sym, **kw):
    """Oracle factor on the product's ordering.  Diagonal blocks are passed
    only for compressed supernodes (the reference splits only those that are
    compressed or IN_PLACE; uncompressed ones are dense)."""
    from oracle.oracle import OracleFactor
    blocks = [b if c else np.zeros(0, np.int32)
              for b, c in zip(sym["diag_blocks"], sym["compressed"])]
    return OracleFactor(A, sym["perm"], sym["first"], sym["compressed"], blocks, **kw)


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    d = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (d if d > 0 else 1.0)
