"""The oracle pinned: against TNL's own results (golden fixtures written from the reference
build, and the reference build itself when present), TNL's unit-test golden values, closed
forms and brute force."""
import numpy as np
import pytest

import pyoracle as O
import tnl_cases
import workloads as W

KINDS = list(O.KINDS)


def _cases(golden):
    d = golden["segments"]
    return sorted({k.split("/")[0] for k in d.files})


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("org", ["col", "row"])
def test_metadata_matches_tnl_fixtures(golden, kind, org):
    d = golden["segments"]
    for case in _cases(golden):
        caps = np.diff(d[f"{case}/row_ptr"]).astype(np.int32)
        s = O.Segments(kind, caps, column_major=(org == "col"))
        for name in O.METADATA_NAMES[O.KINDS[kind]]:
            np.testing.assert_array_equal(s.get(name), d[f"{case}/{kind}/{org}/{name}"], err_msg=f"{case} {name}")
        assert s.storage_size() == int(d[f"{case}/{kind}/{org}/storage"][0])


@pytest.mark.parametrize("kind", KINDS)
def test_conversion_and_spmv_match_tnl_fixtures(golden, kind):
    d = golden["segments"]
    for case in _cases(golden):
        rp, c, v, x = (d[f"{case}/{k}"] for k in ("row_ptr", "cols", "vals", "x"))
        s = O.Segments(kind, np.diff(rp))
        vals, cols = s.convert(rp, c, v)
        key = f"{case}/{kind}/col/values"
        if key in d.files:
            used = len(d[key])
            np.testing.assert_array_equal(vals[:used], d[key])
            np.testing.assert_array_equal(cols[:used], d[f"{case}/{kind}/col/columns"])
            assert np.all(cols[used:] == -1)
        y = s.spmv(vals, cols, x)
        np.testing.assert_array_equal(y, d[f"{case}/{kind}/col/y"], err_msg=case)


def test_csr_adaptive_blocks_match_tnl_fixtures(golden):
    d = golden["adaptive"]
    s = golden["segments"]
    for key in d.files:
        case, vs = key.split("/")
        rp = s[f"{case}/row_ptr"]
        np.testing.assert_array_equal(O.csr_adaptive_blocks(rp, int(vs)), d[key], err_msg=key)


def _solver_runs(golden):
    d = golden["solvers"]
    runs = sorted({k.split("/")[0] for k in d.files if k.startswith("run")})
    for r in runs:
        sysname = str(d[f"{r}/system"])
        params = dict(eval(str(d[f"{r}/params"])))  # list of (key, value) pairs written by make_fixtures.py
        yield r, sysname, str(d[f"{r}/method"]), params


def test_solvers_match_tnl_histories(golden):
    """Single-thread TNL Host solves: same iterations, same residual history, same x (bit for
    bit under the no-contraction build of both)."""
    d = golden["solvers"]
    for run, sysname, method, params in _solver_runs(golden):
        rp, c, v, b = (d[f"{sysname}/{k}"] for k in ("row_ptr", "cols", "vals", "b"))
        r = O.solve(method, rp, c, v, b, **params)
        assert r["iterations"] == int(d[f"{run}/iterations"][0]), run
        assert r["converged"] == int(d[f"{run}/converged"][0]), run
        np.testing.assert_array_equal(r["history"], d[f"{run}/history"], err_msg=run)
        np.testing.assert_array_equal(r["x"], d[f"{run}/x"], err_msg=run)


@pytest.mark.parametrize("kind", KINDS)
def test_tnl_unit_test_golden_values(kind):
    for name, case in tnl_cases.all_cases().items():
        rp, cols, vals = tnl_cases.to_csr(case)
        s = O.Segments(kind, np.diff(rp))
        sv, sc = s.convert(rp, cols, vals)
        x = np.full(case["m"], case["x"])
        y = s.spmv(sv, sc, x)
        np.testing.assert_array_equal(y, np.asarray(case["y"], np.float64), err_msg=f"{kind} {name}")


@pytest.mark.parametrize("kind", KINDS)
def test_spmv_brute_force_and_keepers(kind):
    rng = np.random.default_rng(5)
    for n, m, dens in [(1, 1, 1.0), (7, 5, 0.5), (33, 40, 0.2), (100, 90, 0.05), (65, 65, 0.0)]:
        A = rng.standard_normal((n, m)) * (rng.random((n, m)) < dens)
        rp, c, v = W.dense_to_csr(A)
        s = O.Segments(kind, np.diff(rp))
        sv, sc = s.convert(rp, c, v)
        x = rng.standard_normal(m)
        y0 = rng.standard_normal(n)
        np.testing.assert_allclose(s.spmv(sv, sc, x), A @ x, rtol=1e-12, atol=1e-12)
        for alpha, beta in [(1.0, 0.0), (2.5, 0.0), (1.0, -0.5), (-1.0, 1.0), (0.75, 3.0)]:
            first, last = (1, n - 1) if n > 2 else (0, n)
            got = s.spmv(sv, sc, x, y0, alpha, beta, first, last)
            want = y0.copy()
            if not (kind == "BiEllpack" and s.storage_size() == 0):
                # TNL's Host BiEllpack returns before touching y when nothing is stored
                # (BiEllpackView.hpp:248-249); the oracle keeps that behaviour
                want[first:last] = alpha * (A @ x)[first:last] + beta * y0[first:last]
            np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_laplacian_row_sums_closed_form():
    """A * 1 of the Dirichlet Laplacians = number of missing neighbours per row."""
    for N, gen, dims in [(13, W.laplacian_2d, 2), (9, W.laplacian_3d, 3)]:
        rp, c, v = gen(N)
        n = len(rp) - 1
        s = O.Segments("CSR", np.diff(rp))
        y = s.spmv(v, c, np.ones(n))
        idx = np.arange(n)
        want = np.zeros(n)
        for d in range(dims):
            coord = (idx // N ** d) % N
            want += (coord == 0).astype(float) + (coord == N - 1).astype(float)
        np.testing.assert_array_equal(y, want)


def test_reductions_closed_forms():
    # reduceTest.h:91-128: sum of ones = size; sum of iota = n(n-1)/2
    for size in [1, 2, 127, 128, 255, 256, 257, 1000, 9377, 100000]:
        ones = np.ones(size)
        assert O.dot(ones, ones) == size
        iota = np.arange(size, dtype=np.float64)
        assert O.dot(iota, ones) == size * (size - 1) / 2


def test_cg_converges_in_distinct_eigenvalue_count():
    """CG on an SPD diagonal matrix with k distinct eigenvalues reaches the solution in k steps."""
    n, k = 300, 4
    diag = np.repeat(np.array([1.0, 2.0, 5.0, 9.0]), n // k)
    rp = np.arange(n + 1, dtype=np.int32)
    c = np.arange(n, dtype=np.int32)
    b = W.uniform_vector(n, 1)
    r = O.solve("cg", rp, c, diag, b, convergence_residue=1e-12, max_iterations=50)
    assert r["converged"] == 1
    assert r["iterations"] <= k + 1
    np.testing.assert_allclose(r["x"], b / diag, rtol=1e-10)


def test_jacobi_closed_form():
    """With the Diagonal preconditioner on a diagonal matrix, CG and GMRES solve in one step.
    BiCGStab breaks down there: s = 0 after the half step gives omega = 0/0 (TNL has no
    half-step convergence test, BICGStab.hpp:72-76), so its residue is NaN and solve() fails."""
    n = 200
    d = np.linspace(1.0, 3.0, n)
    rp = np.arange(n + 1, dtype=np.int32)
    c = np.arange(n, dtype=np.int32)
    b = W.uniform_vector(n, 2)
    for method in ("cg", "gmres"):
        r = O.solve(method, rp, c, d, b, convergence_residue=1e-12, max_iterations=20, preconditioner=1)
        assert r["converged"] == 1
        np.testing.assert_allclose(r["x"], b / d, rtol=1e-12, atol=1e-14)
    r = O.solve("bicgstab", rp, c, d, b, convergence_residue=1e-12, max_iterations=20, preconditioner=1)
    assert r["converged"] == 0 and np.isnan(r["residue"])


def test_gmres_residual_non_increasing_within_cycles():
    rp, c, v = W.convection_diffusion_3d(6, 10.0)
    b = W.uniform_vector(len(rp) - 1, 3)
    for variant in ("CGS", "MGS", "CWY"):
        r = O.solve("gmres", rp, c, v, b, convergence_residue=1e-10, max_iterations=500, gmres_variant=variant,
                    restarting_min=30, restarting_max=30)
        h = r["history"]
        assert r["converged"] == 1
        assert np.all(np.diff(h[1:31]) <= 1e-12 * h[0])


def test_true_residual_of_solutions():
    rp, c, v = W.laplacian_2d(20)
    n = len(rp) - 1
    A = W.csr_to_dense(rp, c, v, n)
    b = W.uniform_vector(n, 4)
    for method in ("cg", "bicgstab", "gmres"):
        r = O.solve(method, rp, c, v, b, convergence_residue=1e-10, max_iterations=5000)
        assert np.linalg.norm(b - A @ r["x"]) / np.linalg.norm(b) < 1e-9


def test_planted_mistake_is_caught():
    """A transposed operand in the product must fail the brute-force check used above."""
    rng = np.random.default_rng(9)
    A = rng.standard_normal((20, 20)) * (rng.random((20, 20)) < 0.3)
    rp, c, v = W.dense_to_csr(A.T)
    s = O.Segments("CSR", np.diff(rp))
    x = rng.standard_normal(20)
    assert not np.allclose(s.spmv(v, c, x), A @ x)


@pytest.mark.skipif(not O.ref_available(), reason="TNL reference build not present")
@pytest.mark.parametrize("kind", KINDS)
def test_oracle_matches_live_reference_build(kind):
    rp, c, v = W.irregular(5000, mean=12, sigma=1.4, max_len=2000, seed=21)
    caps = np.diff(rp)
    for cm in (True, False):
        a = O.Segments(kind, caps, cm)
        b = O.Segments(kind, caps, cm, backend="ref")
        for name in O.METADATA_NAMES[O.KINDS[kind]]:
            np.testing.assert_array_equal(a.get(name), b.get(name))
    x = W.uniform_vector(len(caps), 8)
    if kind != "Ellpack":
        vals, cols, y = O.ref_convert_spmv(kind, rp, c, v, len(caps), x)
        a = O.Segments(kind, caps)
        ov, oc = a.convert(rp, c, v)
        np.testing.assert_array_equal(ov, vals[:len(ov)])
        np.testing.assert_array_equal(a.spmv(ov, oc, x), y)


def _ilu_cases():
    return {"lap2d_12": W.laplacian_2d(12), "irr_500": W.irregular(500, mean=8, sigma=1.0, max_len=64, seed=5),
            "convdiff_5": W.convection_diffusion_3d(5, 10.0)}


def test_ilu0_factors_match_tnl():
    """The oracle's ILU(0) factors and solve equal TNL's ILU0 (Host) bit for bit."""
    if not O.ref_available():
        pytest.skip("TNL reference build not present")
    for name, (rp, c, v) in _ilu_cases().items():
        b = W.uniform_vector(len(rp) - 1, 1)
        Lo, Uo, xo = O.ilu0(rp, c, v, b)
        Lr, Ur, xr = O.ilu0(rp, c, v, b, backend="ref")
        for got, want in zip(Lo + Uo + (xo,), Lr + Ur + (xr,)):
            np.testing.assert_array_equal(got, want, err_msg=name)


def _dense_lu(L, U, n):
    Ld = np.eye(n)
    Ud = np.zeros((n, n))
    for (rp, c, v), D in ((L, Ld), (U, Ud)):
        for i in range(n):
            for k in range(rp[i], rp[i + 1]):
                D[i, c[k]] += v[k]
    return Ld, Ud


def test_ilu0_reproduces_a_on_its_pattern():
    """The defining property of ILU(0): (L U)_ij = a_ij wherever a_ij is stored, L unit lower
    and U upper triangular with A's sparsity; (L U) x = b for the solve."""
    for name, (rp, c, v) in _ilu_cases().items():
        n = len(rp) - 1
        b = W.uniform_vector(n, 2)
        L, U, x = O.ilu0(rp, c, v, b)
        Ld, Ud = _dense_lu(L, U, n)
        assert np.all(np.tril(Ud, -1) == 0) and np.all(np.triu(Ld, 1) == 0), name
        A = W.csr_to_dense(rp, c, v, n)
        LU = Ld @ Ud
        rows = np.repeat(np.arange(n), np.diff(rp))
        scale = np.abs(Ld) @ np.abs(Ud)
        err = np.abs(LU[rows, c] - A[rows, c]) / np.maximum(scale[rows, c], 1e-300)
        assert err.max() < 1e-13, (name, err.max())
        np.testing.assert_allclose(LU @ x, b, rtol=0, atol=1e-10 * np.abs(b).max(), err_msg=name)


def test_ilu0_is_the_exact_lu_without_fill():
    """A tridiagonal matrix has no fill-in: ILU(0) is its LU factorization (Thomas algorithm)."""
    n = 50
    rng = np.random.default_rng(3)
    lo, di, up = rng.uniform(-1, 1, n), rng.uniform(3, 4, n), rng.uniform(-1, 1, n)
    rp = [0]
    cols, vals = [], []
    for i in range(n):
        for j, val in ((i - 1, lo[i]), (i, di[i]), (i + 1, up[i])):
            if 0 <= j < n:
                cols.append(j)
                vals.append(val)
        rp.append(len(cols))
    rp, cols, vals = np.array(rp, np.int32), np.array(cols, np.int32), np.array(vals)
    b = rng.uniform(-1, 1, n)
    L, U, x = O.ilu0(rp, cols, vals, b)
    # Thomas: u_0 = d_0, l_i = a_i / u_{i-1}, u_i = d_i - l_i c_{i-1}
    u = np.zeros(n)
    l = np.zeros(n)
    u[0] = di[0]
    for i in range(1, n):
        l[i] = lo[i] / u[i - 1]
        u[i] = di[i] - l[i] * up[i - 1]
    np.testing.assert_allclose(L[2], l[1:], rtol=1e-15)
    np.testing.assert_allclose(U[2][U[0][:-1]], u, rtol=1e-15)
    np.testing.assert_allclose(x, np.linalg.solve(W.csr_to_dense(rp, cols, vals, n), b), rtol=1e-12)


def _lower(rp, c, v):
    """The lower triangle (diagonal included) of a CSR matrix."""
    keep = c <= np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    lens = np.add.reduceat(keep.astype(np.int64), rp[:-1]) * (np.diff(rp) > 0)
    return np.concatenate([[0], np.cumsum(lens)]).astype(np.int32), c[keep], v[keep]


def _symmetric_cases():
    out = {"lap3d_8": W.laplacian_3d(8)}
    rp, c, v = W.irregular(400, mean=10, sigma=1.0, max_len=60, seed=9)
    A = W.csr_to_dense(rp, c, v, 400)
    S = A + A.T
    rows, cols = np.nonzero(S)
    rp2 = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=400))]).astype(np.int32)
    out["irr_sym_400"] = (rp2, cols.astype(np.int32), S[rows, cols])
    return {k: _lower(*m) for k, m in out.items()}


def test_symmetric_product_matches_tnl():
    """The oracle's SymmetricMatrix product equals TNL's Host product (one thread) bit for bit,
    for the keeper variants and row ranges."""
    if not O.ref_available():
        pytest.skip("TNL reference build not present")
    for name, (rp, c, v) in _symmetric_cases().items():
        n = len(rp) - 1
        x = W.uniform_vector(n, 3)
        y0 = W.uniform_vector(n, 4)
        for alpha, beta, first, last in ((1.0, 0.0, 0, 0), (-1.0, 1.0, 0, 0), (0.5, 2.0, 0, 0), (1.0, 0.0, n // 3, n // 2)):
            got = O.symmetric_spmv(rp, c, v, x, y0, alpha, beta, first, last)
            want = O.symmetric_spmv(rp, c, v, x, y0, alpha, beta, first, last, backend="ref")
            np.testing.assert_array_equal(got, want, err_msg=f"{name} {alpha} {beta} {first} {last}")


def test_symmetric_product_is_the_full_product():
    """L + L^T - D applied to x, to rounding; beta scales y0 (NaN in y0 survives beta = 0, as
    TNL multiplies rather than clears)."""
    for name, (rp, c, v) in _symmetric_cases().items():
        n = len(rp) - 1
        L = W.csr_to_dense(rp, c, v, n)
        full = L + L.T - np.diag(np.diag(L))
        x = W.uniform_vector(n, 5)
        y0 = W.uniform_vector(n, 6)
        got = O.symmetric_spmv(rp, c, v, x, y0, 0.5, 2.0)
        scale = 0.5 * np.abs(full) @ np.abs(x) + 2.0 * np.abs(y0)
        assert np.all(np.abs(got - (0.5 * full @ x + 2.0 * y0)) <= 1e-14 * scale), name
    y0 = np.full(n, np.nan)
    assert np.all(np.isnan(O.symmetric_spmv(rp, c, v, x, y0, 1.0, 0.0)))


def test_symmetric_product_golden_values():
    """TNL's own expected products of its symmetric-matrix unit test."""
    for (rp, c, v), x, y in tnl_cases.symmetric_matrices():
        np.testing.assert_array_equal(O.symmetric_spmv(rp, c, v, x), np.array(y, np.float64))


def test_binary_product_golden_values_and_tnl():
    """Binary matrices (Real = bool, ComputeReal = int): TNL's unit-test values, and TNL's own
    product for random patterns, keepers and row ranges."""
    for (rp, c), xval, y in tnl_cases.binary_matrices():
        n = len(rp) - 1
        np.testing.assert_array_equal(O.binary_spmv(rp, c, np.full(n, xval)), np.array(y, np.int32))
    if not O.ref_available():
        return
    rng = np.random.default_rng(4)
    rp, c, _ = W.irregular(300, mean=8, sigma=1.0, max_len=40, seed=4)
    x = rng.integers(-1000, 1000, 300).astype(np.int32)
    y0 = rng.integers(-1000, 1000, 300).astype(np.int32)
    for alpha, beta, first, last in ((1, 0, 0, 0), (3, 0, 0, 0), (1, 2, 0, 0), (-2, 5, 50, 200)):
        np.testing.assert_array_equal(O.binary_spmv(rp, c, x, y0, alpha, beta, first, last),
                                      O.binary_spmv(rp, c, x, y0, alpha, beta, first, last, backend="ref"))


@pytest.mark.parametrize("S", [1, 8, 16, 64, 128, 256])
def test_sliced_slice_sizes(S):
    """SlicedEllpack< ..., SliceSize > for other slice sizes than 32: per slice the largest
    capacity, offsets the exclusive scan of SliceSize times it (SlicedEllpack.hpp:64-102), the
    oracle equal to TNL's own build in both organizations."""
    rng = np.random.default_rng(S)
    for caps in (rng.integers(0, 9, 1000), np.zeros(37, np.int64), rng.integers(1, 40, 2 * S + 3),
                 np.array([5], np.int64), np.array([], np.int64)):
        caps = caps.astype(np.int32)
        n = len(caps)
        nsl = -(-n // S)
        mx = np.array([caps[s * S:(s + 1) * S].max() for s in range(nsl)], np.int32)
        off = np.concatenate([[0], np.cumsum(mx.astype(np.int64) * S)]).astype(np.int32)
        for org in (True, False):
            seg = O.Segments("SlicedEllpack", caps, column_major=org, slice_size=S)
            assert np.array_equal(seg.get("sliceSegmentSizes"), mx)
            assert np.array_equal(seg.get("sliceOffsets"), off)
            assert seg.storage_size() == off[-1]
            if O.ref_available():
                ref = O.Segments("SlicedEllpack", caps, column_major=org, backend="ref", slice_size=S)
                for name in ("sliceOffsets", "sliceSegmentSizes", "sliced"):
                    assert np.array_equal(seg.get(name), ref.get(name)), (S, org, name)
                assert seg.storage_size() == ref.storage_size()
