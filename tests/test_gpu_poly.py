"""Jacobi-polynomial preconditioner (preconditioner 3, no TNL counterpart) on the GPU.

M^-1 s = z_k with z_0 = D^-1 s, z_{j+1} = z_j + D^-1 (s - P z_j): D the diagonal of the solved
matrix, P the preconditioner matrix (the solved matrix, or a BF16 copy of it).  The solvers that
use it are checked against plain numpy restatements of TNL's preconditioned CG (CG.hpp:17-107)
and left-preconditioned GMRES (GMRES.hpp:79-180, one restart cycle) with the same polynomial:
a dropped term, a sign error or the wrong diagonal changes the residue already after the first
step.  Degree 0 must give Jacobi's history.  With BF16 P inside FP64 refinement the solve reaches
the FP64 tolerance (the FP32 Krylov operator, not P, sets the accuracy)."""
import numpy as np
import pytest
import torch

import pyoracle as O
import workloads as W

pytestmark = pytest.mark.gpu


def csr_dot(rp, c, v, x):
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    return np.bincount(rows, weights=v * x[c], minlength=n)


def to_bf16(v):
    """FP64 -> FP32 -> BF16 (round to nearest even), as the device conversion does."""
    u = np.asarray(v, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def poly(rp, c, pv, d, k):
    def apply(s):
        z = s / d
        for _ in range(k):
            z = z + (s - csr_dot(rp, c, pv, z)) / d
        return z
    return apply


def pcg_history(rp, c, v, b, prec, iters):
    """TNL's CG.hpp:17-107 with z = M^-1 r; residue sqrt((r, z)) / |M^-1 b|."""
    x = np.zeros_like(b)
    normb = np.linalg.norm(prec(b))
    r = b - csr_dot(rp, c, v, x)
    z = prec(r)
    p = z.copy()
    s1 = r @ z
    hist = [np.sqrt(s1) / normb]
    for _ in range(iters):
        Ap = csr_dot(rp, c, v, p)
        alpha = s1 / (Ap @ p)
        x += alpha * p
        r -= alpha * Ap
        z = prec(r)
        s2 = r @ z
        beta = s2 / s1
        s1 = s2
        p = z + beta * p
        hist.append(np.sqrt(s1) / normb)
    return np.array(hist)


def gmres_history(rp, c, v, b, prec, steps):
    """Left-preconditioned GMRES from x0 = 0 within its first restart cycle: the residue after step
    j is |g_{j+1}| / |M^-1 b| (GMRES.hpp:79-180, Givens :722-763)."""
    w0 = prec(b)
    beta = np.linalg.norm(w0)
    V = [w0 / beta]
    H = np.zeros((steps + 1, steps))
    g = np.zeros(steps + 1)
    g[0] = beta
    cs = np.zeros(steps)
    sn = np.zeros(steps)
    hist = [1.0]
    for j in range(steps):
        w = prec(csr_dot(rp, c, v, V[j]))
        for i in range(j + 1):
            H[i, j] = V[i] @ w
            w = w - H[i, j] * V[i]
        H[j + 1, j] = np.linalg.norm(w)
        V.append(w / H[j + 1, j])
        for i in range(j):
            t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
            H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
            H[i, j] = t
        dd = np.hypot(H[j, j], H[j + 1, j])
        cs[j], sn[j] = H[j, j] / dd, H[j + 1, j] / dd
        g[j + 1] = -sn[j] * g[j]
        g[j] = cs[j] * g[j]
        hist.append(abs(g[j + 1]) / beta)
    return np.array(hist)


@pytest.mark.parametrize("degree", [0, 1, 3])
def test_polynomial_cg_follows_numpy_pcg(tnl, dev, degree):
    rp, c, v = W.laplacian_3d(10)
    n = len(rp) - 1
    b = W.uniform_vector(n, 3)
    d = v[np.concatenate([np.arange(rp[i], rp[i + 1])[c[rp[i]:rp[i + 1]] == i] for i in range(n)])]
    want = pcg_history(rp, c, v, b, poly(rp, c, v, d, degree), 15)
    A = tnl.SparseMatrix.from_csr(rp, c, v, format="SlicedEllpack", dtype="f64")
    s = tnl.CG()
    s.setMatrix(A)
    s.setup(convergence_residue=1e-30, max_iterations=15, preconditioner=3, poly_degree=degree)
    xd = torch.zeros(n, dtype=torch.float64, device=dev)
    s.solve(torch.tensor(b, device=dev), xd)
    got = s.getResidualHistory()
    print(degree, got[:5], want[:5])
    assert len(got) >= 16
    np.testing.assert_allclose(got[:16], want[:16], rtol=1e-9)
    if degree == 0:
        j = tnl.CG()
        j.setMatrix(A)
        j.setup(convergence_residue=1e-30, max_iterations=15, preconditioner=1)
        xj = torch.zeros(n, dtype=torch.float64, device=dev)
        j.solve(torch.tensor(b, device=dev), xj)
        np.testing.assert_allclose(got[:16], j.getResidualHistory()[:16], rtol=1e-12)


@pytest.mark.parametrize("variant", ["CWY", "MGS"])
def test_bf16_polynomial_gmres_follows_numpy(tnl, dev, variant):
    """FP32 GMRES on the FP32 operator, polynomial over the BF16 copy of the matrix."""
    rp, c, v = W.convection_diffusion_3d(12, 10.0)
    n = len(rp) - 1
    b = W.uniform_vector(n, 4)
    A64 = tnl.SparseMatrix.from_csr(rp, c, v, format="CSRAdaptive", dtype="f64")
    A32 = A64.convert("CSRAdaptive", "f32")
    P16 = A64.convert("CSRAdaptive", "bf16")
    v32 = np.asarray(v, np.float32).astype(np.float64)
    d32 = np.array([v32[k] for i in range(n) for k in range(rp[i], rp[i + 1]) if c[k] == i])
    want = gmres_history(rp, c, v32, b, poly(rp, c, to_bf16(v), d32, 3), 12)
    s = tnl.GMRES()
    s.setMatrix(A32)
    s.setPreconditionerMatrix(P16)
    s.setup(convergence_residue=1e-30, max_iterations=12, preconditioner=3, poly_degree=3, gmres_variant=variant,
            restarting_min=30, restarting_max=30)
    xd = torch.zeros(n, dtype=torch.float32, device=dev)
    s.solve(torch.tensor(b, dtype=torch.float32, device=dev), xd)
    # the history opens with the residue 1 of x0 = 0 (set once or, by the variants that recompute
    # it at the top of the restart cycle, twice), then one residue per step
    got = s.getResidualHistory()
    got = got[int(np.argmax(got < 0.999)) - 1:]
    print(variant, got[:6], want[:6])
    m = min(len(got), 13)
    assert m >= 10
    np.testing.assert_allclose(got[:m], want[:m], rtol=2e-4)


def test_preconditioner_matrix_checks(tnl, dev):
    rp, c, v = W.laplacian_3d(6)
    A32 = tnl.SparseMatrix.from_csr(rp, c, v, format="CSRScalar", dtype="f32")
    A64 = tnl.SparseMatrix.from_csr(rp, c, v, format="CSRScalar", dtype="f64")
    small = tnl.SparseMatrix.from_csr(*W.laplacian_3d(5), format="CSRScalar", dtype="f32")
    n = len(rp) - 1
    b = torch.ones(n, dtype=torch.float32, device=dev)
    for P in (A64, small):
        s = tnl.GMRES()
        s.setMatrix(A32)
        s.setPreconditionerMatrix(P)
        s.setup(preconditioner=3)
        with pytest.raises(tnl.TnlHipError):
            s.solve(b, torch.zeros(n, dtype=torch.float32, device=dev))
    s = tnl.GMRES()
    s.setMatrix(A32)
    s.setup(preconditioner=3, poly_degree=33)
    with pytest.raises(tnl.TnlHipError):
        s.solve(b, torch.zeros(n, dtype=torch.float32, device=dev))


@pytest.mark.parametrize("peclet", [1.0, 10.0, 100.0])
def test_bf16_polynomial_refinement_reaches_fp64_tolerance(tnl, dev, peclet):
    rp, c, v = W.convection_diffusion_3d(16, peclet)
    n = len(rp) - 1
    b = W.uniform_vector(n, 9)
    A = tnl.SparseMatrix.from_csr(rp, c, v, format="SlicedEllpack", dtype="f64")
    xd = torch.zeros(n, dtype=torch.float64, device=dev)
    r = tnl.refine_solve(A, torch.tensor(b, device=dev), xd, inner_dtype="f32", poly_dtype="bf16",
                         preconditioner=3, poly_degree=4)
    x = xd.cpu().numpy()
    true = np.linalg.norm(b - csr_dot(rp, c, v, x)) / np.linalg.norm(b)
    print(peclet, r, true)
    assert r["converged"] and true < 1.01e-10
    ref = O.solve("gmres", rp, c, v, b, convergence_residue=1e-12, max_iterations=20000, gmres_variant="MGSR",
                  restarting_min=30, restarting_max=30, preconditioner=1)
    assert np.linalg.norm(x - ref["x"]) / np.linalg.norm(ref["x"]) < 1e-7
