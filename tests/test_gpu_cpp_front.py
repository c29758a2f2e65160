"""The TNL-named C++ front (tnl-dev_amd/TNL) used the way TNL user code is written: the
SpMV unit tests of SparseMatrixVectorProductTest.hpp for every segments type, conversion,
capacity errors, and CG / GMRES through LinearSolver::setMatrix / solve."""
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
