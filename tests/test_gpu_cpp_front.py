"""The TNL-named C++ front (tnl-dev_amd/TNL) used the way TNL user code is written: the
SpMV unit tests of SparseMatrixVectorProductTest.hpp for every segments type, conversion,
capacity errors, CG / GMRES through LinearSolver::setMatrix / solve, and the hipcc-compiled traversal
program."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin", "test_front")


def test_cpp_front_program(dev):
    assert os.path.exists(BIN), "tests/cpp/bin/test_front is not built (run __graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0
    assert "ALL PASSED" in r.stdout


def test_cpp_front_traversals_monitor_and_rebinding(dev):
    """User code compiled with hipcc: reduceAllRows / forAllElements / forAllRows with device lambdas,
    getCompressedRowLengths of SparseMatrixTest.hpp:241-297 and the per-row maximum of
    SegmentsTest.hpp:101-150 in every format, a solver that follows its matrix through setElement
    and operator= after setMatrix, the solver monitor and the residual history file."""
    binary = os.path.join(ROOT, "tests", "cpp", "bin", "test_traverse")
    assert os.path.exists(binary), "tests/cpp/bin/test_traverse is not built (run __graft_entry__.build())"
    r = subprocess.run([binary], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0
    assert "ALL PASSED" in r.stdout
