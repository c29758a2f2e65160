"""Parity at the sizes the benchmark figures are quoted on, through the C ABI, against the oracle
(TNL's Host traversal order, pinned to TNL's own results) -- the check TNL's SpMV benchmark makes
of every format against host CSR (Benchmarks/SpMV/SpmvBenchmarkResult.h:45-59):
  - the 160^3 7-point Laplacian with FP32 values in every format (A 1 exact; random x against the
    oracle's float32 product of the float32-rounded inputs: bit for bit for the one-lane-per-row
    kernels, within 4e-6 of sum |a_ij x_j| otherwise);
  - the 2 M-row lognormal irregular matrix (rows up to 2,495 entries) on CSR-Adaptive (plain and
    column-striped), ChunkedEllpack and BiEllpack in FP64 (within 1e-13 of sum |a_ij x_j|) and FP32,
    and Jacobi BiCGStab to 1e-10 with the true residual from the oracle's product;
  - FP64 refinement over an FP32 inner GMRES(30) on the 128^3 convection-diffusion matrix (Pe 10)
    with the true FP64 residual from the oracle's product."""
import numpy as np
import pytest
import torch

import pyoracle as O
import workloads as W

pytestmark = pytest.mark.gpu

FORMATS = ["CSRScalar", "CSRVector", "CSRLight", "CSRHybrid", "CSRAdaptive", "Ellpack", "SlicedEllpack",
           "ChunkedEllpack", "BiEllpack"]
KIND = {"CSRScalar": "CSR", "CSRVector": "CSR", "CSRLight": "CSR", "CSRHybrid": "CSR", "CSRAdaptive": "CSR",
        "Ellpack": "Ellpack", "SlicedEllpack": "SlicedEllpack", "ChunkedEllpack": "ChunkedEllpack",
        "BiEllpack": "BiEllpack"}
EXACT = {"CSRScalar", "Ellpack", "SlicedEllpack"}


def row_abs(rp, cols, vals, x):
    return np.add.reduceat(np.abs(vals) * np.abs(x[cols]), rp[:-1])


@pytest.fixture(scope="module")
def lap160():
    rp, c, v = W.laplacian_3d(160)
    n = len(rp) - 1
    x32 = W.uniform_vector(n, 3).astype(np.float32)
    seg = O.Segments("CSR", np.diff(rp))
    y32 = seg.spmv(v.astype(np.float32), c, x32, dtype=np.float32)
    return rp, c, v, x32, y32, row_abs(rp, c, v, x32.astype(np.float64))


@pytest.mark.parametrize("fmt", FORMATS)
def test_laplacian160_fp32_every_format(tnl, dev, lap160, fmt):
    rp, c, v, x32, y32, scale = lap160
    n = len(rp) - 1
    print(f"160^3 FP32 {fmt}", flush=True)
    A = tnl.SparseMatrix.from_csr(rp, c, v, format=fmt, dtype="f32")
    y = torch.zeros(n, dtype=torch.float32, device=dev)
    A.vectorProduct(torch.ones(n, dtype=torch.float32, device=dev), y)
    ones = y.cpu().numpy()
    want = np.add.reduceat(v, rp[:-1])
    np.testing.assert_array_equal(ones, want.astype(np.float32))
    A.vectorProduct(torch.tensor(x32, device=dev), y)
    got = y.cpu().numpy()
    if fmt in EXACT or (fmt in ("CSRVector", "CSRLight", "CSRHybrid")):  # short rows: the staged Host-order kernel
        np.testing.assert_array_equal(got, y32)
    else:
        err = np.abs(got.astype(np.float64) - y32.astype(np.float64))
        assert np.all(err <= 4e-6 * scale + 1e-30), f"{fmt}: worst {np.max(err / np.maximum(scale, 1e-30)):.3e}"


@pytest.fixture(scope="module")
def irregular2m():
    n = 2_000_000
    rp, c, v = W.irregular(n, 16.0, 1.2, 4096, seed=2024)
    x = W.uniform_vector(n, 5)
    y = O.Segments("CSR", np.diff(rp)).spmv(v, c, x)
    return rp, c, v, x, y, row_abs(rp, c, v, x)


@pytest.mark.parametrize("fmt,stripes", [("CSRAdaptive", 0), ("CSRAdaptive", 8), ("ChunkedEllpack", 0),
                                         ("BiEllpack", 0)])
def test_irregular_2m_fp64(tnl, dev, irregular2m, fmt, stripes):
    rp, c, v, x, y_ref, scale = irregular2m
    n = len(rp) - 1
    assert np.diff(rp).max() > 2000, "the long rows are the point of the case"
    print(f"irregular 2M FP64 {fmt} stripes={stripes}", flush=True)
    A = tnl.SparseMatrix.from_csr(rp, c, v, format=fmt, dtype="f64")
    if stripes:
        A.set_param("column_stripes", stripes)
    y = torch.zeros(n, dtype=torch.float64, device=dev)
    A.vectorProduct(torch.tensor(x, device=dev), y)
    err = np.abs(y.cpu().numpy() - y_ref)
    assert np.all(err <= 1e-13 * scale), f"{fmt}: worst {np.max(err / scale):.3e}"
    # alpha, beta and a row range through the same kernels
    y0 = W.uniform_vector(n, 8)
    yd = torch.tensor(y0, device=dev)
    A.vectorProduct(torch.tensor(x, device=dev), yd, -0.5, 2.0, 1000, n - 999)
    got = yd.cpu().numpy()
    want = y0.copy()
    want[1000:n - 999] = 2.0 * y0[1000:n - 999] - 0.5 * y_ref[1000:n - 999]
    err = np.abs(got - want)
    assert np.all(err <= 1e-13 * (0.5 * scale + 2.0 * np.abs(y0)) + 1e-300)


@pytest.mark.parametrize("fmt,stripes", [("CSRAdaptive", 0), ("CSRAdaptive", 4), ("ChunkedEllpack", 0),
                                         ("BiEllpack", 0)])
def test_irregular_2m_fp32(tnl, dev, irregular2m, fmt, stripes):
    rp, c, v, x, _, _ = irregular2m
    n = len(rp) - 1
    print(f"irregular 2M FP32 {fmt} stripes={stripes}", flush=True)
    x32 = x.astype(np.float32)
    v32 = v.astype(np.float32)
    y_ref = O.Segments("CSR", np.diff(rp)).spmv(v32, c, x32, dtype=np.float32).astype(np.float64)
    scale = row_abs(rp, c, v32.astype(np.float64), x32.astype(np.float64))
    A = tnl.SparseMatrix.from_csr(rp, c, v, format=fmt, dtype="f32")
    if stripes:
        A.set_param("column_stripes", stripes)
    y = torch.zeros(n, dtype=torch.float32, device=dev)
    A.vectorProduct(torch.tensor(x32, device=dev), y)
    err = np.abs(y.cpu().numpy().astype(np.float64) - y_ref)
    # the rows sum up to 2,496 float32 products in another order than the Host's
    assert np.all(err <= 2e-5 * scale + 1e-30), f"{fmt}: worst {np.max(err / scale):.3e}"


def test_irregular_2m_jacobi_bicgstab(tnl, dev, irregular2m):
    rp, c, v, _, _, _ = irregular2m
    n = len(rp) - 1
    b = W.uniform_vector(n, 6)
    A = tnl.SparseMatrix.from_csr(rp, c, v, format="CSRAdaptive", dtype="f64")
    s = tnl.BICGStab()
    s.setMatrix(A)
    s.setup(convergence_residue=1e-10, max_iterations=10000, preconditioner=1)
    xd = torch.zeros(n, dtype=torch.float64, device=dev)
    assert s.solve(torch.tensor(b, device=dev), xd)
    x = xd.cpu().numpy()
    r = b - O.Segments("CSR", np.diff(rp)).spmv(v, c, x)
    # the stopping rule is on the Jacobi-preconditioned residue; the true relative residual follows
    true = np.linalg.norm(r) / np.linalg.norm(b)
    print(f"BiCGStab: {s.getIterations()} iterations, preconditioned residue {s.getResidue():.3e}, true {true:.3e}")
    assert true < 1e-9, true


def test_convection_diffusion_128_refinement(tnl, dev):
    rp, c, v = W.convection_diffusion_3d(128, 10.0)
    n = len(rp) - 1
    b = W.uniform_vector(n, 7)
    A = tnl.SparseMatrix.from_csr(rp, c, v, format="CSRAdaptive", dtype="f64")
    xd = torch.zeros(n, dtype=torch.float64, device=dev)
    res = tnl.refine_solve(A, torch.tensor(b, device=dev), xd, tolerance=1e-10, inner_method="gmres",
                           inner_dtype="f32")
    x = xd.cpu().numpy()
    true = np.linalg.norm(b - O.Segments("CSR", np.diff(rp)).spmv(v, c, x)) / np.linalg.norm(b)
    print(f"refinement: {res}, true residual {true:.3e}")
    assert res["converged"] and true < 1e-10, (res, true)
    # the same with the degree-4 Jacobi polynomial over BF16 values preconditioning the FP32 GMRES
    xp = torch.zeros(n, dtype=torch.float64, device=dev)
    res = tnl.refine_solve(A, torch.tensor(b, device=dev), xp, tolerance=1e-10, inner_method="gmres",
                           inner_dtype="f32", poly_dtype="bf16", preconditioner=3, poly_degree=4)
    xpn = xp.cpu().numpy()
    true = np.linalg.norm(b - O.Segments("CSR", np.diff(rp)).spmv(v, c, xpn)) / np.linalg.norm(b)
    print(f"BF16-polynomial refinement: {res}, true residual {true:.3e}")
    assert res["converged"] and true < 1e-10, (res, true)
    assert np.linalg.norm(xpn - x) / np.linalg.norm(x) < 1e-7
