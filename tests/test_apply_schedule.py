"""The apply's task lists on the CPU (tests/cpp/emulate_apply.cpp): the
dataflow list run in ticket order, on threads that take tickets, and on
threads that pop a ready queue gives the result of the level-scheduled list (bitwise, or to rounding when the dataflow list splits
long rows into partial sums), every wait is satisfied when its task starts,
and every buffer access is in range; with and without subtree tasks, in
level order and in list-schedule ticket order."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "emulate_apply")


@pytest.fixture(scope="module", autouse=True)
def built():
    import importlib.util
    spec = importlib.util.spec_from_file_location("rsc_cpp_test_build",
                                                  os.path.join(HERE, "cpp", "build.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    mod.build()


@pytest.mark.parametrize("args", [
    "p3d 12 20 40 1 8",      # compressed separators, HODLR diagonals
    "p3d 16 63 64 2 8",      # two right-hand sides
    "p2d 40 10 16 1 8",      # 2D, many compressed levels
    "el 5 30 40 1 8",        # elasticity (3 dof per node)
    "p2d 48 1000000 400 1 8",  # no compression
    "p3d 24 63 128 3 8",     # split column tasks (partial sums), fused tv reductions
    "p3d 32 63 64 1 8",      # many fused tv reductions (separators over 512 columns)
    "p3d 20 63 50 1 8 0",    # no subtree tasks
    "p3d 20 63 50 1 8 100000000",  # subtree tasks as large as allowed
    "p3d 36 2000 400 2 8",   # uncompressed separators over 1024 columns: split rows
    "p2d 1100 1000000 4000 1 8 0",  # 1100-column top separator, no subtree tasks
    "p3d 20 63 50 1 8 0 0",  # level order (no list schedule)
    "p3d 20 63 50 16 8 0 3",  # list schedule on three workgroups, 16 RHS costs
    "p2d 40 10 16 1 8 1000 1",  # list schedule on one workgroup
    "p3d 32 2000 400 1 8 16384 1024 20000",  # wide products split finer
    "p3d 24 63 128 2 8 16384 4096 5000",  # the same with compressed separators
    # the multi-right-hand-side list with its sixteen-column subtree tasks
    "p3d 20 63 50 16 8 16384 1024 0 multi",
    "p3d 24 63 128 17 8 16384 1024 0 multi",  # two passes of right-hand sides
    "p2d 40 10 16 3 8 16384 64 0 multi",
    "el 5 30 40 2 8 16384 1024 0 multi",
    "p3d 16 63 64 1 8 16384 0 0 multi",  # level order
    "p3d 24 63 64 16 8 16384 1024 0 multi 7",  # subtrees merged into seven tasks
    "el 6 30 40 2 8 16384 1024 0 multi 3",  # three tasks of several trees each
])
def test_dataflow_schedule_matches_level_schedule(args):
    r = subprocess.run([BIN] + args.split(), capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok")
