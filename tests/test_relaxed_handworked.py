"""The relaxed-supernode rule (NestedDissection.cpp:1266-1667, CHOLMOD's
relaxed amalgamation) on hand-worked cases, fixture
tests/golden/relaxed_supernodes_handworked.json.  Every expected value there
was derived by hand, as written below; none was produced by the library.

chain7: the path graph 0-1-...-6 (tridiagonal matrix) with the dissection
  components {0} {2} | {1} | {4} {6} | {5} | {3} in postorder:
  perm (new -> old) = 0 2 1 4 6 5 3; cparent = 2 2 6 5 5 6 -1.
  New edges: (0,2) (1,2) (1,6) (3,6) (3,5) (4,5).  Eliminating in order:
  col 0 rows {2}; col 1 rows {2,6}; col 2 gets 6 from col 1 -> {6};
  col 3 rows {5,6}; col 4 rows {5}; col 5 gets 6 from col 3 -> {6}; col 6
  none.  So etree parent = 2 2 6 5 5 6 -1 and column counts (diagonal
  included) = 2 3 2 3 2 2 1; columns 2, 5, 6 have two etree children.
  mdb = 0: components of size > 0 with children, {1}->col 2, {5}->col 5 and
  the root col 6, are compressed whole supernodes; cols 0, 1 and 3, 4 do
  not chain (parent[0] = 2 != 1, parent[3] = 5 != 4): seven supernodes,
  compressed 0 0 1 0 0 1 1.
  mdb = 1: nothing compressed; no column chains with its predecessor
  (cols 2, 6 have two children, col 5 has count 2 != 2 + 1, the others
  parent mismatches), so seven fundamental supernodes.  Amalgamation from
  the top: 5+6 (2 columns <= 4), 4+{5,6} (3), 3+{4,5,6} (4) merge
  outright; 2 + {3..6}: 5 columns, counts 2 and 4, new zeros 1 (4 + 1 - 2)
  = 3, entries 5*6/2 + 5 (4 - 4) = 15, ratio 0.2 < 0.8 -> merge (zeros 3,
  count 5); 1 + {2..6}: 6 columns, counts 3 and 5, new zeros 1 (5+1-3) = 3,
  ratio (3+3) / 21 = 0.286 -> merge (zeros 6, count 6); 0 + {1..6}: 7
  columns, counts 2 and 6, new 1 (6+1-2) = 5, ratio 11/28 = 0.393 ->
  merge.  One supernode.

two40_*: an etree chain of 40 columns whose column counts break between
  X = 0..19 and Y = 20..39 (count(j) = 40 - j on Y; on X count(19) = c and
  count(j) = c + 19 - j), so there are two fundamental supernodes.  With
  20 + 20 columns (above 16, at most 48) they merge iff the zero ratio is
  below 0.1: new zeros 20 (20 + 20 - count(0)) over 40*41/2 + 40 (20 - 20)
  = 820 entries.  c = 16: count(0) = 35, 100/820 = 0.122 -> two
  supernodes; c = 17: count(0) = 36, 80/820 = 0.098 -> one.  Split into
  two components (Y the parent of X, Y with one child): with mdb 10 Y is
  compressed and never merged (two supernodes, the second compressed); with
  mdb 30 nothing is compressed and the c = 17 merge happens.

two12_small: X = 0..5, Y = 6..11, count(5) = 4, count(0) = 9, Y's first
  count 6: 12 columns (<= 16), new zeros 6 (6 + 6 - 9) = 18 over 78: 0.23 <
  0.8 -> one supernode.
two6_zeros: X = 0..2, Y = 3..5, count(2) = 2, count(0) = 4, Y's first 3: 6
  columns (> 4), new zeros 3 (3 + 3 - 4) = 6 over 21: 0.286 -> one.
two_children: X1 = 0..1 and X2 = 2..3 both children of Y = 4..5
  (parent = 1 4 3 4 5 -1, counts 3 2 3 2 2 1): X2 + Y have 4 columns ->
  merge; X1's parent Y now stands as X2 = its next supernode: 6 columns,
  counts 3 and 4, new zeros 2 (4 + 2 - 3) = 6 over 21 -> merge: one.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import rank_structured_cholesky_amd as R
from rank_structured_cholesky_amd import _lib
from helpers import symbolic

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                    "relaxed_supernodes_handworked.json")
CASES = json.load(open(GOLD))


def _tridiagonal(n):
    rows, cols, vals = [], [], []
    for j in range(n):
        for i, v in ((j - 1, -1.0), (j, 2.0), (j + 1, -1.0)):
            if 0 <= i < n:
                rows.append(i)
                cols.append(j)
                vals.append(v)
    import scipy.sparse as sp
    return R.SparseColumnMatrix.from_scipy(sp.csc_matrix((vals, (rows, cols)), shape=(n, n)))


def test_chain7_etree_and_partition():
    c = CASES["chain7"]
    A = _tridiagonal(7)
    perm = np.array(c["perm"], np.int32)
    par = np.empty(7, np.int32)
    cnt = np.empty(7, np.int32)
    _lib.check(_lib.lib().rsc_etree_colcounts(7, _lib.iptr(A._p), _lib.iptr(A._i),
                                              _lib.iptr(perm), _lib.iptr(par),
                                              _lib.iptr(cnt)))
    assert par.tolist() == c["etree_parent"]
    assert cnt.tolist() == c["colcount"]
    ordering = (perm, np.array(c["cparent"], np.int32), np.array(c["cmember"], np.int32))
    for key, mdb in (("mdb0", 0), ("mdb1", 1)):
        s = symbolic(A, None, ordering=ordering, max_dense_block=mdb, reorder_separators=0)
        assert s["perm"].tolist() == c["perm"]
        assert s["first"].tolist() == c[key]["first"], key
        assert s["compressed"].tolist() == c[key]["compressed"], key


def _etree_case(case):
    n = case["n"]
    if case.get("chain"):
        h = n // 2
        parent = [j + 1 for j in range(n - 1)] + [-1]
        count = [case["colcount_x_last"] + (h - 1 - j) for j in range(h)] + \
                [n - j for j in range(h, n)]
    else:
        parent, count = case["parent"], case["colcount"]
    return n, parent, count


@pytest.mark.parametrize("case", CASES["etrees"], ids=[c["name"] for c in CASES["etrees"]])
def test_handworked_etrees(case):
    n, parent, count = _etree_case(case)
    comps = case["comps"]
    first_c = np.array([a for a, b, k, p in comps], np.int32)
    last_c = np.array([b for a, b, k, p in comps], np.int32)
    kids = np.array([k for a, b, k, p in comps], np.int32)
    cpar = np.array([p for a, b, k, p in comps], np.int32)
    size = last_c - first_c + 1
    first = np.empty(n + 1, np.int32)
    comp = np.empty(n, np.int32)
    ns = C.c_int()
    _lib.check(_lib.lib().rsc_relaxed_supernodes(
        n, _lib.iptr(np.array(parent, np.int32)), _lib.iptr(np.array(count, np.int32)),
        len(comps), _lib.iptr(size), _lib.iptr(first_c), _lib.iptr(last_c), _lib.iptr(kids),
        _lib.iptr(cpar), case["mdb"], _lib.iptr(first), _lib.iptr(comp), C.byref(ns)))
    assert first[: ns.value + 1].tolist() == case["first"]
    assert comp[: ns.value].tolist() == case["compressed"]


def test_relaxed_supernodes_rejects_bad_components():
    n = 4
    parent = np.array([1, 2, 3, -1], np.int32)
    count = np.array([4, 3, 2, 1], np.int32)
    first = np.empty(n + 1, np.int32)
    comp = np.empty(n, np.int32)
    ns = C.c_int()
    z = np.zeros(1, np.int32)
    # one component that does not cover the columns
    st = _lib.lib().rsc_relaxed_supernodes(
        n, _lib.iptr(parent), _lib.iptr(count), 1, _lib.iptr(np.array([3], np.int32)),
        _lib.iptr(z), _lib.iptr(np.array([2], np.int32)), _lib.iptr(z),
        _lib.iptr(np.array([-1], np.int32)), 10, _lib.iptr(first), _lib.iptr(comp),
        C.byref(ns))
    assert st == _lib.RSC_BAD_INPUT
