"""Products with several right-hand sides (tnlhip_spmv_multi) on the GPU: column j of the block
result equals the oracle's one-vector product (TNL's Host order) of column j, bit for bit in FP64,
in every format; keepers, row ranges and the refusals follow tnlhip_spmv."""
import numpy as np
import pytest
import torch

import pyoracle as O
import workloads as W

pytestmark = pytest.mark.gpu

FORMATS = ["CSRScalar", "CSRAdaptive", "Ellpack", "SlicedEllpack", "ChunkedEllpack", "BiEllpack"]


def oracle_columns(rp, c, v, X):
    seg = O.Segments("CSR", np.diff(rp))
    ov, oc = seg.convert(rp, c, v)
    return np.stack([seg.spmv(ov, oc, np.ascontiguousarray(X[:, j])) for j in range(X.shape[1])], axis=1)


@pytest.mark.parametrize("fmt", FORMATS)
@pytest.mark.parametrize("k", [1, 3, 16, 40])
def test_block_product_equals_oracle_per_column(tnl, dev, fmt, k):
    rp, c, v = W.irregular(3001, 9.0, 1.0, 300, seed=k)
    n = len(rp) - 1
    X = np.random.default_rng(k).uniform(-1, 1, (n, k))
    want = oracle_columns(rp, c, v, X)
    A = tnl.SparseMatrix.from_csr(rp, c, v, format=fmt, dtype="f64")
    Y = torch.full((n, k), np.nan, dtype=torch.float64, device=dev)
    A.vectorProductMulti(torch.tensor(X, device=dev), Y)
    torch.cuda.synchronize()
    assert np.array_equal(Y.cpu().numpy(), want), fmt


@pytest.mark.parametrize("fmt", ["CSRAdaptive", "SlicedEllpack"])
def test_block_product_keepers_and_row_range(tnl, dev, fmt):
    rp, c, v = W.laplacian_3d(14)
    n = len(rp) - 1
    k = 5
    rng = np.random.default_rng(3)
    X = rng.uniform(-1, 1, (n, k))
    Y0 = rng.uniform(-1, 1, (n, k))
    AX = oracle_columns(rp, c, v, X)
    A = tnl.SparseMatrix.from_csr(rp, c, v, format=fmt, dtype="f64")
    Xd = torch.tensor(X, device=dev)
    for alpha, beta in ((1.0, 0.0), (-2.0, 0.0), (1.0, 0.5), (0.25, -1.5)):
        Y = torch.tensor(Y0, device=dev)
        A.vectorProductMulti(Xd, Y, alpha, beta, 100, 2000)
        got = Y.cpu().numpy()
        want = Y0.copy()
        want[100:2000] = beta * Y0[100:2000] + alpha * AX[100:2000] if beta else alpha * AX[100:2000]
        np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-13)
        assert np.array_equal(got[:100], Y0[:100]) and np.array_equal(got[2000:], Y0[2000:])


@pytest.mark.parametrize("dt", ["f32", "bf16"])
def test_block_product_low_precision(tnl, dev, dt):
    rp, c, v = W.convection_diffusion_3d(16, 10.0)
    n = len(rp) - 1
    k = 8
    X = np.random.default_rng(5).uniform(-1, 1, (n, k)).astype(np.float32)
    A = tnl.SparseMatrix.from_csr(rp, c, v, format="ChunkedEllpack", dtype=dt)
    Y = torch.zeros((n, k), dtype=torch.float32, device=dev)
    A.vectorProductMulti(torch.tensor(X, device=dev), Y)
    y1 = torch.zeros(n, dtype=torch.float32, device=dev)
    got = Y.cpu().numpy()
    for j in range(k):
        # the one-vector product of the same matrix on column j: same order, same bits
        A.vectorProduct(torch.tensor(np.ascontiguousarray(X[:, j]), device=dev), y1)
        ref = y1.cpu().numpy()
        scale = np.abs(v).max() * 7
        assert np.max(np.abs(got[:, j] - ref)) <= 1e-6 * scale


def test_block_product_refusals(tnl, dev):
    rp, c, v = W.laplacian_3d(6)
    n = len(rp) - 1
    A = tnl.SparseMatrix.from_csr(rp, c, v, format="CSRScalar", dtype="f64")
    for k in (65, 100):
        with pytest.raises(tnl.TnlHipError):
            A.vectorProductMulti(torch.zeros((n, k), dtype=torch.float64, device=dev),
                                 torch.zeros((n, k), dtype=torch.float64, device=dev))
    rows = np.repeat(np.arange(n), np.diff(rp))
    keep = c <= rows
    lrp = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=n))]).astype(np.int32)
    S = tnl.SparseMatrix.symmetric_from_csr(lrp, c[keep], v[keep], format="CSRScalar")
    with pytest.raises(tnl.TnlHipError):
        S.vectorProductMulti(torch.zeros((n, 2), dtype=torch.float64, device=dev),
                             torch.zeros((n, 2), dtype=torch.float64, device=dev))


@pytest.mark.parametrize("fmt", FORMATS)
def test_block_product_without_columns(tnl, dev, fmt):
    """A matrix with no columns (every row empty) and an empty X: only the keepers act, and X,
    which holds no entry, is never read."""
    n, k = 70, 5
    rp = np.zeros(n + 1, np.int32)
    A = tnl.SparseMatrix.from_csr(rp, np.zeros(0, np.int32), np.zeros(0), columns=0, format=fmt, dtype="f64")
    Y0 = np.random.default_rng(4).uniform(-1, 1, (n, k))
    X = torch.zeros((0, k), dtype=torch.float64, device=dev)
    for alpha, beta in ((1.0, 0.0), (2.0, 0.5)):
        Y = torch.tensor(Y0, device=dev)
        A.vectorProductMulti(X, Y, alpha, beta)
        assert np.array_equal(Y.cpu().numpy(), beta * Y0), (fmt, alpha, beta)
        y = torch.tensor(Y0[:, 0].copy(), device=dev)
        A.vectorProduct(torch.zeros(0, dtype=torch.float64, device=dev), y, alpha, beta)
        assert np.array_equal(y.cpu().numpy(), beta * Y0[:, 0]), (fmt, alpha, beta)


def test_binary_product_without_columns(tnl, dev):
    n = 45
    rp = np.zeros(n + 1, np.int32)
    B = tnl.SparseMatrix.from_csr(rp, np.zeros(0, np.int32), None, columns=0, format="SlicedEllpack", dtype="binary")
    y0 = torch.arange(n, dtype=torch.int32, device=dev)
    for alpha, beta in ((1, 0), (3, 2)):
        y = y0.clone()
        B.vectorProduct(torch.zeros(0, dtype=torch.int32, device=dev), y, alpha, beta)
        assert torch.equal(y, beta * y0)


def test_block_product_unaligned_blocks_take_the_scalar_lanes(tnl, dev):
    """X and Y starting 8 bytes past a 16-byte boundary (FP64, k even) take the one-column-per-lane
    kernel instead of the 16-byte lanes; both give the oracle's bits."""
    rp, c, v = W.irregular(2501, 8.0, 1.0, 120, seed=21)
    n, k = len(rp) - 1, 6
    X = np.random.default_rng(21).uniform(-1, 1, (n, k))
    want = oracle_columns(rp, c, v, X)
    A = tnl.SparseMatrix.from_csr(rp, c, v, format="SlicedEllpack", dtype="f64")
    Xbig = torch.zeros(n * k + 1, dtype=torch.float64, device=dev)
    Xbig[1:] = torch.tensor(X.ravel(), device=dev)
    Ybig = torch.full((n * k + 1,), np.nan, dtype=torch.float64, device=dev)
    Xu, Yu = Xbig[1:].view(n, k), Ybig[1:].view(n, k)
    assert Xu.data_ptr() % 16 == 8 and Yu.data_ptr() % 16 == 8
    A.vectorProductMulti(Xu, Yu)
    Ya = torch.full((n, k), np.nan, dtype=torch.float64, device=dev)
    A.vectorProductMulti(torch.tensor(X, device=dev), Ya)
    torch.cuda.synchronize()
    assert np.array_equal(Yu.cpu().numpy(), want)
    assert np.array_equal(Ya.cpu().numpy(), want)
