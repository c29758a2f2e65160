import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tnl-dev_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X)")


def fixture_checksums():
    """(file, expected, actual) for every fixture listed in tests/golden/SHA256SUMS."""
    import hashlib
    g = os.path.join(ROOT, "tests", "golden")
    out = []
    with open(os.path.join(g, "SHA256SUMS")) as f:
        for line in f:
            want, name = line.split()
            with open(os.path.join(g, name), "rb") as h:
                out.append((name, want, hashlib.sha256(h.read()).hexdigest()))
    return out


def pytest_sessionstart(session):
    # a damaged fixture fails here, naming the file and both checksums, before any test reads it
    bad = [(n, w, a) for n, w, a in fixture_checksums() if w != a]
    if bad:
        raise pytest.UsageError("golden fixtures differ from tests/golden/SHA256SUMS: "
                                + "; ".join(f"{n} expected {w} got {a}" for n, w, a in bad))


@pytest.fixture(scope="session")
def tnl():
    from __graft_entry__ import load_package
    return load_package()


class LoadedNpz(dict):
    """The arrays of an .npz file read into memory; `files` lists the names as NpzFile does."""

    @property
    def files(self):
        return list(self.keys())


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    g = os.path.join(ROOT, "tests", "golden")
    out = {}
    for k in ("segments", "adaptive", "solvers"):
        # read every array now (np.load is lazy), so that a read error names the file at once
        with np.load(os.path.join(g, f"{k}.npz")) as z:
            out[k] = LoadedNpz((name, z[name]) for name in z.files)
    return out


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")
