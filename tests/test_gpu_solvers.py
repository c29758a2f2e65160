"""Reductions, BLAS-1 and the Krylov solvers on the GPU through the C ABI, against the oracle
(single-thread TNL semantics, pinned to TNL's own residual histories) and the mathematics."""
import os
import numpy as np
import pytest
import torch

import pyoracle as O
import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.gpu


def t(a, dev, dt=torch.float64):
    return torch.tensor(np.asarray(a), dtype=dt, device=dev)


def test_dot_norm_closed_forms_and_determinism(tnl, dev):
    # reduceTest.h:91-128 closed forms; same bits on repeated calls
    for size in [1, 2, 127, 128, 255, 256, 257, 1000, 9377, 100000, 4096000]:
        ones = torch.ones(size, dtype=torch.float64, device=dev)
        assert tnl.dot(ones, ones) == size
        iota = torch.arange(size, dtype=torch.float64, device=dev)
        assert tnl.dot(iota, ones) == size * (size - 1) / 2
    rng = np.random.default_rng(1)
    a = rng.standard_normal(1 << 20)
    b = rng.standard_normal(1 << 20)
    ad, bd = t(a, dev), t(b, dev)
    d1 = tnl.dot(ad, bd)
    assert all(tnl.dot(ad, bd) == d1 for _ in range(5))
    assert abs(d1 - float(np.dot(a, b))) <= 1e-12 * float(np.abs(a) @ np.abs(b))
    assert abs(tnl.lpNorm2(ad) - np.linalg.norm(a)) <= 1e-13 * np.linalg.norm(a)
    a32 = t(a, dev, torch.float32)
    assert abs(tnl.dot(a32, a32) - float(np.dot(a.astype(np.float32).astype(np.float64),
                                                 a.astype(np.float32).astype(np.float64)))) < 1e-9 * len(a)


def test_axpby_multidot_gemv_scan(tnl, dev):
    rng = np.random.default_rng(2)
    n, k = 100003, 31
    ld = ((n + 31) // 32) * 32
    V = rng.standard_normal((k, ld))
    w = rng.standard_normal(n)
    Vd, wd = t(V.ravel(), dev), t(w, dev)
    got = tnl.multidot(Vd, ld, k, wd)
    want = V[:, :n] @ w
    assert np.all(np.abs(got - want) <= 1e-12 * (np.abs(V[:, :n]) @ np.abs(w)))
    c = rng.standard_normal(20)
    y = rng.standard_normal(n)
    yd = t(y, dev)
    tnl.gemv(-1.0, Vd, ld, 20, c, 0.5, yd)
    want = -(V[:20, :n].T @ c) + 0.5 * y
    np.testing.assert_allclose(yd.cpu().numpy(), want, rtol=1e-12, atol=1e-12)
    x = rng.standard_normal(n)
    yd = t(y, dev)
    tnl.axpby(2.0, t(x, dev), -3.0, yd)
    np.testing.assert_allclose(yd.cpu().numpy(), 2 * x - 3 * y, rtol=1e-15, atol=1e-15)
    caps = rng.integers(0, 50, 1234567).astype(np.int32)
    cd = torch.tensor(caps, device=dev)
    out = torch.zeros(len(caps) + 1, dtype=torch.int32, device=dev)
    tnl.exclusive_scan(cd, out)
    want = np.zeros(len(caps) + 1, np.int64)
    np.cumsum(caps, out=want[1:])
    np.testing.assert_array_equal(out.cpu().numpy(), want.astype(np.int32))


def _run_device_solver(tnl, dev, method, rp, c, v, b, fmt="CSRLight", dtype="f64", **params):
    A = tnl.SparseMatrix.from_csr(rp, c, v, format=fmt, dtype=dtype)
    s = tnl.getLinearSolver(method)
    s.setMatrix(A)
    s.setup(**params)
    bd = t(b, dev, A.vector_dtype)
    xd = torch.zeros(len(b), dtype=A.vector_dtype, device=dev)
    ok = s.solve(bd, xd)
    return ok, s.getIterations(), s.getResidue(), s.getResidualHistory(), xd.cpu().numpy().astype(np.float64)


def test_solvers_follow_tnl_histories(tnl, golden, dev):
    """Every fixture run (CG, BiCGStab, all GMRES variants, with and without Jacobi): same
    convergence flag, iteration count within 2 of the range TNL's own count spans over OpenMP
    thread counts, residual history within 1e-6 relative over the first iterations (the device
    sums in a different order), true residual at the tolerance."""
    d = golden["solvers"]
    runs = sorted({k.split("/")[0] for k in d.files if k.startswith("run")})
    errors = []
    for run in runs:
        try:
            _check_run(tnl, d, dev, run)
        except AssertionError as e:
            errors.append(f"{run}: {e}")
    assert not errors, "\n".join(errors)


def _check_run(tnl, d, dev, run):
    sysname = str(d[f"{run}/system"])
    method = str(d[f"{run}/method"])
    params = dict(eval(str(d[f"{run}/params"])))
    print(f"run {run} {sysname} {method} {params}", flush=True)
    rp, c, v, b = (d[f"{sysname}/{k}"] for k in ("row_ptr", "cols", "vals", "b"))
    ok, its, res, hist, x = _run_device_solver(tnl, dev, method, rp, c, v, b, **params)
    ref_its = int(d[f"{run}/iterations"][0])
    assert ok == bool(d[f"{run}/converged"][0]), run
    # within 2 (ell for BiCGStab(ell)) of TNL's count, or of the range TNL's own count spans over
    # OpenMP thread counts: BiCGStab on the 2D Laplacian moves by up to 10 iterations with the
    # reduction order of its dot products alone
    spread = d[f"{run}/iterations_threads"]
    slack = max(2, params.get("bicgstab_ell", 1))
    assert min(spread) - slack <= its <= max(spread) + slack, (run, its, ref_its, list(spread))
    h_ref = d[f"{run}/history"]
    m = min(len(hist), len(h_ref), 20)
    # residues near 1e-12 and below are exact residues b - A x at the rounding floor of A x
    # (about 1e-16 |A| |x| / |b|), hence the absolute 1e-14
    np.testing.assert_allclose(hist[:m], h_ref[:m], rtol=1e-6, atol=1e-14, err_msg=run)
    A = W.csr_to_dense(rp, c, v, len(b))
    true_res = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    tol = params.get("convergence_residue", 1e-6)
    assert true_res < 50 * tol, (run, true_res)


def test_cg_laplacian_2d_256_config(tnl, dev):
    """The CPU configuration: 256^2 5-point Laplacian, CG to 1e-8 from x0 = 0 (756 iterations on
    TNL's Host path for its seed); device and oracle agree on the iteration count."""
    rp, c, v = W.laplacian_2d(256)
    b = W.uniform_vector(len(rp) - 1, 42)
    ref = O.solve("cg", rp, c, v, b, convergence_residue=1e-8, max_iterations=5000)
    ok, its, res, hist, x = _run_device_solver(tnl, dev, "cg", rp, c, v, b, convergence_residue=1e-8,
                                               max_iterations=5000)
    assert ok and abs(its - ref["iterations"]) <= 3
    assert res < 1e-8
    np.testing.assert_allclose(hist[:50], ref["history"][:50], rtol=1e-8)


@pytest.mark.parametrize("fmt", ["CSRScalar", "CSRAdaptive", "Ellpack", "SlicedEllpack", "ChunkedEllpack", "BiEllpack"])
def test_solvers_independent_of_format(tnl, dev, fmt):
    rp, c, v = W.convection_diffusion_3d(10, 10.0)
    b = W.uniform_vector(len(rp) - 1, 3)
    ok, its, res, hist, x = _run_device_solver(tnl, dev, "bicgstab", rp, c, v, b, fmt=fmt, convergence_residue=1e-10,
                                               max_iterations=2000, preconditioner=1)
    ref = O.solve("bicgstab", rp, c, v, b, convergence_residue=1e-10, max_iterations=2000, preconditioner=1)
    assert ok and abs(its - ref["iterations"]) <= 2


def test_failure_paths(tnl, dev):
    """max_iterations reached -> solve() is false; NaN in b -> false; b = 0 -> immediate
    convergence with x = 0 (normb := 1, CG.hpp:32-33)."""
    rp, c, v = W.laplacian_2d(32)
    n = len(rp) - 1
    b = W.uniform_vector(n, 1)
    ok, its, *_ = _run_device_solver(tnl, dev, "cg", rp, c, v, b, convergence_residue=1e-12, max_iterations=5)
    assert not ok
    ref = O.solve("cg", rp, c, v, b, convergence_residue=1e-12, max_iterations=5)
    assert its == ref["iterations"]
    bn = b.copy()
    bn[7] = np.nan
    ok, *_ = _run_device_solver(tnl, dev, "gmres", rp, c, v, bn, convergence_residue=1e-8, max_iterations=100)
    assert not ok
    ok, its, res, hist, x = _run_device_solver(tnl, dev, "bicgstab", rp, c, v, np.zeros(n), convergence_residue=1e-8)
    assert ok and np.all(x == 0)


def test_float32_solve(tnl, dev):
    rp, c, v = W.laplacian_3d(16)
    b = W.uniform_vector(len(rp) - 1, 9)
    ok, its, res, hist, x = _run_device_solver(tnl, dev, "cg", rp, c, v, b, dtype="f32", convergence_residue=1e-5,
                                               max_iterations=1000)
    ref = O.solve("cg", rp, c, v, b, convergence_residue=1e-5, max_iterations=1000)
    assert ok and abs(its - ref["iterations"]) <= 3


@pytest.mark.parametrize("method,precond", [("cg", 0), ("cg", 1), ("bicgstab", 1), ("tfqmr", 0), ("tfqmr", 1),
                                            ("cg", 2), ("bicgstab", 2), ("bicgstabl", 1)])
def test_graph_replay_matches_direct_launches(tnl, dev, method, precond):
    """Iterations replayed from a captured hipGraph give the same bits as the same kernels
    launched one by one (long enough to go past the eager batches)."""
    rp, c, v = W.laplacian_3d(24) if method in ("cg", "tfqmr") else W.convection_diffusion_3d(20, 10.0)
    b = W.uniform_vector(len(rp) - 1, 4)
    runs = []
    for g in (0, 1):
        ok, its, res, hist, x = _run_device_solver(tnl, dev, method, rp, c, v, b, convergence_residue=1e-10,
                                                   max_iterations=3000, preconditioner=precond, use_graph=g)
        runs.append((ok, its, res, hist, x))
    assert runs[0][1] > 16  # past the eager chunks of 4 and 8 iterations
    assert runs[0][0] == runs[1][0] and runs[0][1] == runs[1][1]
    np.testing.assert_array_equal(runs[0][4], runs[1][4])
    np.testing.assert_array_equal(runs[0][3], runs[1][3])


@pytest.mark.parametrize("fmt", ["SlicedEllpack", "CSRAdaptive"])
def test_ilu0_wide_levels_follow_the_oracle(tnl, dev, fmt):
    """ILU(0) on the 64^3 Laplacian: the anti-diagonal planes hold up to 3072 rows, so the
    schedule mixes one-workgroup runs of small levels with device-wide levels.  CG follows the
    oracle's (TNL's) history and iteration count; BiCGStab, whose iteration count moves with the
    rounding of its dot products, follows the history of its first iterations."""
    rp, c, v = W.laplacian_3d(64)
    b = W.uniform_vector(len(rp) - 1, 5)
    kw = dict(convergence_residue=1e-8, max_iterations=500, preconditioner=2)
    for method, first in (("cg", 30), ("bicgstab", 10)):
        ref = O.solve(method, rp, c, v, b, **kw)
        ok, its, res, hist, x = _run_device_solver(tnl, dev, method, rp, c, v, b, fmt=fmt, **kw)
        print(method, its, ref["iterations"], hist[:first], ref["history"][:first], flush=True)
        assert ok and ref["converged"], method
        if method == "cg":
            assert abs(its - ref["iterations"]) <= 2, (its, ref["iterations"])
        np.testing.assert_allclose(hist[:first], ref["history"][:first], rtol=1e-6, err_msg=method)
        r = b - _csr_matvec(rp, c, v, x)
        assert np.linalg.norm(r) / np.linalg.norm(b) < 5e-8, method
    ok32, its32, res32, _, _ = _run_device_solver(tnl, dev, "bicgstab", rp, c, v, b, fmt=fmt, dtype="f32",
                                                  convergence_residue=1e-5, max_iterations=500, preconditioner=2)
    assert ok32 and its32 <= its + 5


def test_ilu0_refuses_rows_without_an_upper_entry(tnl, dev):
    # row 2 holds only a column below the diagonal: no U_22 to divide by
    rp = np.array([0, 1, 2, 3], np.int32)
    c = np.array([0, 1, 0], np.int32)
    v = np.array([2.0, 3.0, 1.0])
    with pytest.raises(tnl.TnlHipError, match="no entry on or above the diagonal"):
        _run_device_solver(tnl, dev, "cg", rp, c, v, np.ones(3), preconditioner=2)


def _csr_matvec(rp, c, v, x):
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    return np.bincount(rows, weights=v * x[c], minlength=len(rp) - 1)


def test_symmetric_matrix_solves(tnl, dev):
    """CG (plain and with Jacobi) on a SymmetricMatrix holding the lower triangle of the 3-D
    Laplacian follows the oracle's CG on the full matrix; ILU(0) refuses the symmetric type."""
    rp, c, v = W.laplacian_3d(16)
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    keep = c <= rows
    lrp = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=n))]).astype(np.int32)
    b = W.uniform_vector(n, 9)
    for pre in (0, 1):
        ref = O.solve("cg", rp, c, v, b, convergence_residue=1e-10, max_iterations=2000, preconditioner=pre)
        A = tnl.SparseMatrix.symmetric_from_csr(lrp, c[keep], v[keep], format="SlicedEllpack")
        s = tnl.getLinearSolver("cg")
        s.setMatrix(A)
        s.setup(convergence_residue=1e-10, max_iterations=2000, preconditioner=pre)
        x = torch.zeros(n, dtype=torch.float64, device=dev)
        assert s.solve(t(b, dev, torch.float64), x)
        assert abs(s.getIterations() - ref["iterations"]) <= 2, (pre, s.getIterations(), ref["iterations"])
        np.testing.assert_allclose(s.getResidualHistory()[:30], ref["history"][:30], rtol=1e-6)
    s = tnl.getLinearSolver("cg")
    s.setMatrix(A)
    s.setup(preconditioner=2)
    with pytest.raises(tnl.TnlHipError, match="symmetric"):
        s.solve(t(b, dev, torch.float64), torch.zeros(n, dtype=torch.float64, device=dev))


def _parse_history(text):
    """(iteration, residue) lines and the positions of the blank lines of a residual history file."""
    rows, blanks = [], []
    for line in text.split("\n"):
        if line == "":
            blanks.append(len(rows))
            continue
        it, r = line.split("\t")
        rows.append((int(it), float(r)))
    return rows, blanks


def test_residual_history_file_and_monitor_follow_tnl(tnl, dev, tmp_path):
    """setup( residual-history-file ) and setSolverMonitor: the file TNL writes
    (IterativeSolver.hpp:189-200: a blank line at iteration 0, then "<iteration>\\t<residue>"),
    compared with TNL's own file for the same solve -- iteration numbers equal line by line over the
    first lines, residues within the device's rounding -- and a monitor that last sees the final
    (iterations, residue)."""
    hist = np.load(os.path.join(ROOT, "tests", "golden", "history.npz"))
    sol = np.load(os.path.join(ROOT, "tests", "golden", "solvers.npz"))
    runs = sorted({k.split("/")[0] for k in hist.files})
    for run in runs:
        sysname, method = str(hist[f"{run}/system"]), str(hist[f"{run}/method"])
        params = dict(eval(str(hist[f"{run}/params"])))
        print(f"history {run} {sysname} {method} {params}", flush=True)
        rp, c, v, b = (sol[f"{sysname}/{k}"] for k in ("row_ptr", "cols", "vals", "b"))
        A = tnl.SparseMatrix.from_csr(rp, c, v, format="CSRScalar", dtype="f64")
        s = tnl.getLinearSolver(method)
        s.setMatrix(A)
        s.setup(**params)
        path = tmp_path / f"{run}.txt"
        s.setResidualHistoryFile(path)
        seen = []
        s.setSolverMonitor(lambda it, r: seen.append((it, r)))
        x = torch.zeros(len(b), dtype=torch.float64, device=dev)
        s.solve(torch.tensor(b, device=dev), x)
        s.setResidualHistoryFile(None)  # closes the file
        ours, ours_blank = _parse_history(path.read_text())
        ref, ref_blank = _parse_history(str(hist[f"{run}/text"]))
        assert ours_blank[0] == 0 and ref_blank[0] == 0, run
        m = min(len(ours), len(ref), 20)
        assert [i for i, _ in ours[:m]] == [i for i, _ in ref[:m]], run
        np.testing.assert_allclose([r for _, r in ours[:m]], [r for _, r in ref[:m]], rtol=2e-6, atol=1e-14,
                                   err_msg=run)
        assert [i for i, _ in ours] == list(s.getResidualHistoryIterations()), run
        np.testing.assert_allclose([r for _, r in ours], s.getResidualHistory(), rtol=1e-6, err_msg=run)
        assert seen and seen[-1] == (s.getIterations(), s.getResidue()), run
