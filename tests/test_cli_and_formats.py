"""The pcg_solver driver (reference src/pcg_solver.cpp:50-122: same file
contract) and the reference's ordering dumps (SparseSolver.cpp:136,
NestedDissection.cpp:394-395, STLUtil.h:105-138 record format).

CPU: the driver builds and fails cleanly on missing files; the dump writer
and reader round-trip; a dumped ordering fed back through
rsc_symbolic_create_with_ordering gives the same supernode partition and
diagonal blocks.  GPU: the driver solves a package-written system (with
and without positions) and matches the Python path; a factor built from
ingested dumps equals the factor of the original ordering bit for bit, and
one built from a different ordering equals the oracle on that ordering."""
import os
import struct
import subprocess

import numpy as np
import pytest

import rank_structured_cholesky_amd as R
from rank_structured_cholesky_amd import build as B
from helpers import oracle_for, relerr, symbolic


def _write_case(tmp_path, A, b, xyz=None, name="case"):
    base = str(tmp_path / name)
    A.writeToBinaryGZ(base + "_system.bcsm.gz")
    R.write_vector(base + "_rhs.vector", b)
    if xyz is not None:
        R.write_positions(base + "_meshPositions.dat", xyz)
    return base


def _run_cli(*args, timeout=300):
    if not os.path.exists(B.CLI):
        pytest.fail(f"{B.CLI} is not built (python -m rank_structured_cholesky_amd.build)")
    r = subprocess.run([B.CLI, *args], capture_output=True, text=True, timeout=timeout)
    return r.returncode, r.stdout, r.stderr


def test_cli_usage_and_missing_files(tmp_path):
    rc, out, err = _run_cli()
    assert rc == 1 and "Usage" in err
    rc, out, err = _run_cli(str(tmp_path / "nothing"))
    assert rc == 2 and "cannot open" in err
    # a truncated matrix file is rejected before any device work
    base = str(tmp_path / "bad")
    with open(base + "_system.bcsm.gz", "wb") as f:
        f.write(struct.pack("<QQQ", 10, 4, 4))
    rc, out, err = _run_cli(base)
    assert rc == 2 and "truncated" in err


def test_reference_ordering_dump_roundtrip(tmp_path):
    A, xyz = R.poisson3d(12, 1.0, 1234)
    perm, cparent, cmember = R.nested_dissection(A, xyz, 60)
    R.write_reference_ordering(str(tmp_path), perm, cparent, cmember)
    # the record layout the reference writes: int32 count + 16-byte records
    raw = open(tmp_path / "separators.vector", "rb").read()
    (nc,) = struct.unpack("<i", raw[:4])
    assert nc == len(cparent) and len(raw) == 4 + 16 * nc
    size0, first0, last0, kids0 = struct.unpack("<4i", raw[4:20])
    assert first0 == 0 and last0 == size0 - 1
    p2, c2, m2 = R.read_reference_ordering(str(tmp_path))
    assert np.array_equal(p2, perm) and np.array_equal(c2, cparent)
    assert np.array_equal(m2, cmember)
    # a symbolic analysis on the dumped ordering gives the partition of the
    # library's own run (the separator sub-ordering is already in perm)
    opts = dict(max_dense_block=40, max_diagonal_block=50, nd_small=60)
    own = symbolic(A, xyz, **opts)
    ing = symbolic(A, None, ordering=(own["perm"], cparent, cmember),
                   **dict(opts, reorder_separators=0))
    assert np.array_equal(ing["perm"], own["perm"])
    assert np.array_equal(ing["first"], own["first"])
    assert np.array_equal(ing["compressed"], own["compressed"])
    assert own["compressed"].sum() > 0
    for a, b in zip(ing["diag_blocks"], own["diag_blocks"]):
        assert np.array_equal(a, b)


def test_reference_ordering_rejects_bad_dumps(tmp_path):
    A, xyz = R.poisson3d(6, 1.0, 1)
    perm, cparent, cmember = R.nested_dissection(A, xyz, 20)
    R.write_reference_ordering(str(tmp_path), perm, cparent, cmember)
    p = tmp_path / "SparseSolver_permutation.vector"
    data = bytearray(open(p, "rb").read())
    data[4:8] = data[8:12]  # a repeated index: not a permutation
    open(p, "wb").write(bytes(data))
    with pytest.raises(ValueError):
        R.read_reference_ordering(str(tmp_path))


@pytest.mark.gpu
@pytest.mark.parametrize("with_positions", [True, False])
def test_cli_solves_package_written_system(tmp_path, with_positions):
    A, xyz = R.poisson3d(24, 1.0, 1234)
    b = R.uniform_rhs(A._nrow, 5678)
    base = _write_case(tmp_path, A, b, xyz if with_positions else None)
    rc, out, err = _run_cli(base)
    print(out, err)
    assert rc == 0, out + err
    line = [l for l in out.splitlines() if l.startswith("PCG iterations")][0]
    its = int(line.split("=")[1].split(",")[0])
    assert "PCG solution took" in out
    if not with_positions:
        assert "no system degree of freedom positions" in err
    x = R.read_vector(base + "_solution.vector")
    assert relerr(A.multiply(x), b) < 1.5e-6
    # the same driver settings through the Python mirror
    M = R.FactorPreconditioner()
    M.initialize(A, xyz if with_positions else None, 200, 400,
                 reorderSeparators=with_positions)
    S = R.PCG_SOLVER(A, 1e-6, 10000)
    x2 = np.zeros_like(b)
    S.solvePCG(x2, b, M)
    assert its == S.last["iterations"]
    assert np.array_equal(x, x2)


@pytest.mark.gpu
def test_factor_from_ingested_dumps(tmp_path):
    """Dumps of this library's ordering, read back, give the identical
    factor; dumps of another ordering (graph dissection, leaves of 100)
    give the oracle's factor on that ordering; and the driver accepts the
    dump directory as its second argument."""
    A, xyz = R.poisson3d(20, 1.0, 1234)
    b = R.uniform_rhs(A._nrow, 5678)
    M = R.FactorPreconditioner()
    M.initialize(A, xyz, 63, 400)
    own = symbolic(A, xyz, max_dense_block=63)
    # the triple of the factor's own run: final permutation + its components
    _, cparent, cmember = R.nested_dissection(A, xyz, 200)
    R.write_reference_ordering(str(tmp_path), own["perm"], cparent, cmember)
    ordering = R.read_reference_ordering(str(tmp_path))
    Mi = R.FactorPreconditioner()
    Mi.initialize(A, None, 63, 400, reorderSeparators=False, ordering=ordering)
    assert np.array_equal(Mi.permutation(), M.permutation())
    y, yi = np.empty_like(b), np.empty_like(b)
    M.solve(y, b)
    Mi.solve(yi, b)
    assert np.array_equal(y, yi)
    # a different ordering
    other = tmp_path / "other"
    other.mkdir()
    po, co, mo = R.nested_dissection(A, None, 100)
    R.write_reference_ordering(str(other), po, co, mo)
    ordering2 = R.read_reference_ordering(str(other))
    M2 = R.FactorPreconditioner()
    M2.initialize(A, None, 63, 400, reorderSeparators=False, ordering=ordering2)
    sym2 = symbolic(A, None, ordering=ordering2, max_dense_block=63, reorder_separators=0)
    assert np.array_equal(M2.permutation(), sym2["perm"])
    assert sym2["compressed"].sum() > 0
    O = oracle_for(A, sym2)
    y2 = np.empty_like(b)
    M2.solve(y2, b)
    assert relerr(y2, O.apply(b)) < 1e-9
    # the driver with the dump directory
    base = _write_case(tmp_path, A, b, xyz)
    rc, out, err = _run_cli(base, str(other))
    assert rc == 0, out + err
    x = R.read_vector(base + "_solution.vector")
    assert relerr(A.multiply(x), b) < 1.5e-6
