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
    # Ellpack: 2 x 2 launch shapes; SlicedEllpack: 3 slice sizes x 2 x 2
    assert len(allt) == 16
    assert sorted({t["slice_size"] for t in allt if t["format"] == "SlicedEllpack"}) == [16, 32, 64]
    with pytest.raises(tnl.TnlHipError):
        tnl.autotune(rp, c, v, formats=[])


@pytest.mark.parametrize("fmt", ["SlicedEllpack", "CSRLight"])
def test_autotune_returns_the_best_candidates_matrix(tnl, dev, fmt):
    """The returned matrix is configured as the best candidate (slice size, lanes) and computes
    the same bits as a matrix built and configured that way from the start."""
    rp, c, v = W.irregular(30000, seed=5)
    n = len(rp) - 1
    A, best, allt = tnl.autotune(rp, c, v, formats=[fmt], repetitions=3)
    B = tnl.SparseMatrix.from_csr(rp, c, v, format=fmt, dtype="f64")
    if fmt == "SlicedEllpack":
        assert A.metadata("sliceSize")[0] == best["slice_size"]
        B.set_param("slice_size", best["slice_size"])
        assert np.array_equal(A.metadata("sliceOffsets"), B.metadata("sliceOffsets"))
    else:
        assert A.info.lanes_per_row == (best["lanes"] or B.info.lanes_per_row)
        if best["lanes"]:
            B.set_param("lanes_per_row", best["lanes"])
    B.set_param("blocks_per_cu", best["blocks_per_cu"] or 16)
    B.set_param("pair_rows", 1 if best["pair_rows"] > 0 else 0)
    x = torch.tensor(W.uniform_vector(n, 8), device=dev)
    ya = torch.zeros(n, dtype=torch.float64, device=dev)
    yb = torch.zeros(n, dtype=torch.float64, device=dev)
    A.vectorProduct(x, ya)
    B.vectorProduct(x, yb)
    assert torch.equal(ya, yb)


@pytest.mark.parametrize("S", [1, 8, 16, 64, 128, 256])
def test_slice_sizes_follow_the_oracle(tnl, dev, S):
    """SlicedEllpack with SliceSize S: metadata as the oracle builds it for S, the product of every
    row in TNL's Host order (bit for bit); files keep TNL's 32 only."""
    rp, c, v = W.irregular(5003, 7.0, 1.0, 200, seed=S)
    n = len(rp) - 1
    A = tnl.SparseMatrix.from_csr(rp, c, v, format="SlicedEllpack", dtype="f64")
    A.set_param("slice_size", S)
    caps = np.diff(rp).astype(np.int32)
    seg = O.Segments("SlicedEllpack", caps, slice_size=S)
    assert np.array_equal(A.metadata("sliceOffsets"), seg.get("sliceOffsets"))
    assert np.array_equal(A.metadata("sliceSegmentSizes"), seg.get("sliceSegmentSizes"))
    x = W.uniform_vector(n, 2)
    y = torch.zeros(n, dtype=torch.float64, device=dev)
    A.vectorProduct(torch.tensor(x, device=dev), y)
    ov, oc = seg.convert(rp, c, v)
    assert np.array_equal(y.cpu().numpy(), seg.spmv(ov, oc, x))
    ref = O.Segments("CSR", caps).spmv(v, c, x)
    assert np.array_equal(y.cpu().numpy(), ref)
    with pytest.raises(tnl.TnlHipError):
        A.set_param("slice_size", 48)
    if S != 32:
        with pytest.raises(tnl.TnlHipError):
            A.save("/tmp/sliced_nondefault.tnl")
