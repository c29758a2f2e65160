"""Mixed-precision iterative refinement (FP64 outer, FP32 / BF16 inner) on the GPU: the FP64
true residual reaches the target, and the solution agrees with the oracle's FP64 solve."""
import numpy as np
import pytest
import torch

import pyoracle as O
import workloads as W

pytestmark = pytest.mark.gpu


def spmv_host(rp, c, v, x):
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    return np.bincount(rows, weights=v * x[c], minlength=n)


@pytest.mark.parametrize("inner_dtype", ["f32", "bf16"])
@pytest.mark.parametrize("method,peclet", [("gmres", 10.0), ("gmres", 100.0), ("bicgstab", 10.0)])
def test_refinement_reaches_fp64_tolerance(tnl, dev, inner_dtype, method, peclet):
    rp, c, v = W.convection_diffusion_3d(16, peclet)
    n = len(rp) - 1
    b = W.uniform_vector(n, 5)
    A = tnl.SparseMatrix.from_csr(rp, c, v, format="CSRAdaptive", dtype="f64")
    bd = torch.tensor(b, device=dev)
    xd = torch.zeros(n, dtype=torch.float64, device=dev)
    r = tnl.refine_solve(A, bd, xd, tolerance=1e-10, inner_method=method, inner_dtype=inner_dtype,
                         convergence_residue=1e-4, max_iterations=20000, restarting_min=30, restarting_max=30,
                         preconditioner=1)
    x = xd.cpu().numpy()
    true_res = np.linalg.norm(b - spmv_host(rp, c, v, x)) / np.linalg.norm(b)
    print(method, inner_dtype, peclet, r["outer_iterations"], r["inner_iterations"], r["residue"], true_res)
    assert r["converged"] and r["residue"] < 1e-10
    assert true_res < 1.01e-10
    ref = O.solve("gmres", rp, c, v, b, convergence_residue=1e-12, max_iterations=20000, gmres_variant="MGSR",
                  restarting_min=30, restarting_max=30, preconditioner=1)
    assert np.linalg.norm(x - ref["x"]) / np.linalg.norm(ref["x"]) < 1e-7
    h = r["history"]
    assert all(h[i + 1] < h[i] for i in range(len(h) - 1))


def test_refinement_with_given_low_matrix(tnl, dev):
    rp, c, v = W.laplacian_3d(12)
    n = len(rp) - 1
    b = W.uniform_vector(n, 2)
    A = tnl.SparseMatrix.from_csr(rp, c, v, format="Ellpack", dtype="f64")
    A32 = A.convert("SlicedEllpack", "f32")
    xd = torch.zeros(n, dtype=torch.float64, device=dev)
    r = tnl.refine_solve(A, torch.tensor(b, device=dev), xd, A_low=A32, inner_method="cg", convergence_residue=1e-3,
                         max_iterations=5000)
    assert r["converged"]
    assert np.linalg.norm(b - spmv_host(rp, c, v, xd.cpu().numpy())) / np.linalg.norm(b) < 1.01e-10


@pytest.mark.parametrize("peclet", [1.0, 100.0])
def test_refinement_with_default_inner_target(tnl, dev, peclet):
    """The default parameters (FP32 Jacobi GMRES(30) CWY to 1e-2 per outer step) reach 1e-10 in
    FP64, each outer step cutting the residual at least tenfold."""
    rp, c, v = W.convection_diffusion_3d(16, peclet)
    n = len(rp) - 1
    b = W.uniform_vector(n, 8)
    A = tnl.SparseMatrix.from_csr(rp, c, v, format="SlicedEllpack", dtype="f64")
    xd = torch.zeros(n, dtype=torch.float64, device=dev)
    r = tnl.refine_solve(A, torch.tensor(b, device=dev), xd)
    print(peclet, r["outer_iterations"], r["inner_iterations"], r["history"])
    assert r["converged"] and r["residue"] < 1e-10
    assert np.linalg.norm(b - spmv_host(rp, c, v, xd.cpu().numpy())) / np.linalg.norm(b) < 1.01e-10
    h = r["history"]
    assert all(h[i + 1] < 0.1 * h[i] for i in range(len(h) - 1))


def test_refine_context_matches_the_one_shot_solve(tnl, dev):
    """tnlhip_refine_solve_in on one context, twice in a row (the second system smaller, then the
    first again), gives the same bits as tnlhip_refine_solve: the context only keeps buffers."""
    ctx = tnl.RefineContext()
    kw = dict(tolerance=1e-10, inner_method="gmres", inner_dtype="f32", convergence_residue=1e-2,
              max_iterations=3000, restarting_min=30, restarting_max=30, preconditioner=3, poly_degree=4)
    for N, pe in ((20, 10.0), (12, 100.0), (20, 10.0)):
        rp, c, v = W.convection_diffusion_3d(N, pe)
        n = len(rp) - 1
        A = tnl.SparseMatrix.from_csr(rp, c, v, format="CSRAdaptive", dtype="f64")
        A32 = tnl.SparseMatrix.from_csr(rp, c, v, format="CSRAdaptive", dtype="f32")
        b = torch.tensor(W.uniform_vector(n, N), device=dev)
        x1 = torch.zeros(n, dtype=torch.float64, device=dev)
        x2 = torch.zeros(n, dtype=torch.float64, device=dev)
        r1 = tnl.refine_solve(A, b, x1, A_low=A32, **kw)
        r2 = tnl.refine_solve(A, b, x2, A_low=A32, context=ctx, **kw)
        assert r1["converged"] and r2["converged"]
        assert (r1["outer_iterations"], r1["inner_iterations"]) == (r2["outer_iterations"], r2["inner_iterations"])
        assert torch.equal(x1, x2), (N, pe)
    ctx.close()
