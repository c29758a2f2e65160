"""The padded kernels' row distribution (tnl-dev_amd/csrc/rowdist.hpp, the code the kernels run)
checked on the host for every launch shape the library can make: each row of the range taken once,
every load inside the storage.  No GPU."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_padded_row_distribution_covers_every_row_once_and_stays_in_bounds():
    binary = os.path.join(ROOT, "tests", "cpp", "bin", "test_rowdist")
    assert os.path.exists(binary), "tests/cpp/bin/test_rowdist is not built (run __graft_entry__.build())"
    r = subprocess.run([binary], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:])
    assert r.returncode == 0 and "ALL PASSED" in r.stdout
