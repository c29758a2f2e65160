"""The autotuner on the GPU: it times every requested format / launch configuration, returns the
fastest, and the matrix it returns computes the same product as the oracle."""
import numpy as np
import pytest
import torch

import pyoracle as O
import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_autotune_returns_fastest_and_correct(tnl, dev, dtype):
    rp, c, v = W.irregular(20000, seed=11)
    n = len(rp) - 1
    A, best, allt = tnl.autotune(rp, c, v, dtype=dtype, repetitions=5)
    assert {t["format"] for t in allt} == set(tnl.FORMATS)
    assert best["seconds"] == min(t["seconds"] for t in allt)
    assert tnl.FORMAT_NAMES[A.info.format] == best["format"]
    x = W.uniform_vector(n, 3)
    vt = torch.float64 if dtype == "f64" else torch.float32
    xd = torch.tensor(x, device=dev).to(vt)
    yd = torch.zeros(n, dtype=vt, device=dev)
    A.vectorProduct(xd, yd)
    ref = O.Segments("CSR", np.diff(rp).astype(np.int32)).spmv(v, c, x)
    mag = np.bincount(np.repeat(np.arange(n), np.diff(rp)), weights=np.abs(v * x[c]), minlength=n)
    tol = 1e-13 if dtype == "f64" else 2e-6
    assert np.all(np.abs(yd.cpu().numpy().astype(np.float64) - ref) <= tol * mag + 1e-300)


def test_autotune_subset(tnl):
    rp, c, v = W.laplacian_3d(20)
    A, best, allt = tnl.autotune(rp, c, v, formats=["Ellpack", "SlicedEllpack"], repetitions=3)
    assert {t["format"] for t in allt} == {"Ellpack", "SlicedEllpack"}
    assert len(allt) == 8
    with pytest.raises(tnl.TnlHipError):
        tnl.autotune(rp, c, v, formats=[])
