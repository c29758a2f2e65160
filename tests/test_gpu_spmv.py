"""SpMV parity on the GPU through the C ABI: storage and metadata bit-exact against TNL's
(golden fixtures written from the reference build) and the oracle; products bit-exact for the
one-lane-per-row kernels (same summation order, no contraction) and within
|dy_i| <= 1e-13 * sum_j |a_ij x_j| for the cooperative kernels (FP64)."""
import os
import numpy as np
import pytest
import torch

import pyoracle as O
import tnl_cases
import workloads as W

pytestmark = pytest.mark.gpu

FORMATS = ["CSRScalar", "CSRVector", "CSRLight", "CSRHybrid", "CSRAdaptive", "Ellpack", "SlicedEllpack",
           "ChunkedEllpack", "BiEllpack"]
KIND = {"CSRScalar": "CSR", "CSRVector": "CSR", "CSRLight": "CSR", "CSRHybrid": "CSR", "CSRAdaptive": "CSR",
        "Ellpack": "Ellpack", "SlicedEllpack": "SlicedEllpack", "ChunkedEllpack": "ChunkedEllpack",
        "BiEllpack": "BiEllpack"}
# kernels that sum a row left to right on one lane, products and sums rounded separately
EXACT = {"CSRScalar", "Ellpack", "SlicedEllpack"}
FP64_TOL = 1e-13


def scale_of(rp, cols, vals, x):
    """sum_j |a_ij x_j| per row (the scale of the per-entry tolerance)."""
    n = len(rp) - 1
    ax = np.abs(vals) * np.abs(np.where(cols >= 0, x[np.clip(cols, 0, None)], 0.0))
    return np.add.reduceat(np.concatenate([ax, [0.0]]), np.minimum(rp[:-1], len(ax)))[:n] * (np.diff(rp) > 0)


def check_product(fmt, y, y_ref, scale, tol=FP64_TOL):
    if fmt in EXACT:
        np.testing.assert_array_equal(y, y_ref)
    else:
        err = np.abs(y - y_ref)
        bad = err > tol * scale + 1e-300
        assert not bad.any(), f"{fmt}: {bad.sum()} entries off, worst {np.max(err / np.maximum(scale, 1e-300)):.3e}"


def run_spmv(tnl, A, x, dev, y0=None, alpha=1.0, beta=0.0, first=0, last=0):
    dt = A.vector_dtype
    xd = torch.tensor(x, dtype=dt, device=dev)
    yd = torch.tensor(y0, dtype=dt, device=dev) if y0 is not None else torch.zeros(A.getRows(), dtype=dt, device=dev)
    A.vectorProduct(xd, yd, alpha, beta, first, last)
    torch.cuda.synchronize()
    return yd.cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("fmt", FORMATS)
def test_fixture_storage_metadata_and_product(tnl, golden, dev, fmt):
    d = golden["segments"]
    kind = KIND[fmt]
    cases = sorted({k.split("/")[0] for k in d.files})
    for case in cases:
        print(f"case {case} {fmt}", flush=True)
        rp, c, v, x = (d[f"{case}/{k}"] for k in ("row_ptr", "cols", "vals", "x"))
        ncols = int(d[f"{case}/ncols"][0])
        A = tnl.SparseMatrix.from_csr(rp, c, v, columns=ncols, format=fmt, dtype="f64")
        for name in O.METADATA_NAMES[O.KINDS[kind]]:
            np.testing.assert_array_equal(A.metadata(name), d[f"{case}/{kind}/col/{name}"], err_msg=name)
        assert A.getStorageSize() == int(d[f"{case}/{kind}/col/storage"][0])
        vals, cols = A.storage()
        key = f"{case}/{kind}/col/values"
        if key in d.files:
            np.testing.assert_array_equal(vals, d[key])
            np.testing.assert_array_equal(cols, d[f"{case}/{kind}/col/columns"])
        y = run_spmv(tnl, A, x, dev)
        check_product(fmt, y, d[f"{case}/{kind}/col/y"], scale_of(rp, c, v, x))


@pytest.mark.parametrize("fmt", FORMATS)
def test_tnl_unit_test_golden_values(tnl, dev, fmt):
    for name, case in tnl_cases.all_cases().items():
        rp, cols, vals = tnl_cases.to_csr(case)
        for dt in ("f64", "f32"):
            A = tnl.SparseMatrix.from_csr(rp, cols, vals, columns=case["m"], format=fmt, dtype=dt)
            y = run_spmv(tnl, A, np.full(case["m"], case["x"]), dev)
            np.testing.assert_array_equal(y, np.asarray(case["y"], np.float64), err_msg=f"{fmt} {dt} {name}")


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_csr_scalar_staged_and_walked_chunks_bit_exact(tnl, dev, dtype):
    """CSRScalar stages the elements of each 64-row chunk through LDS when they fit (512 FP64 or
    1024 FP32 elements) and walks the rows in global memory otherwise; both give the Host
    traversal's bits, also for chunks cut by a row range and for empty chunks."""
    rng = np.random.default_rng(21)
    n = 6000
    lens = rng.integers(0, 15, n)          # about 450 elements per chunk: most are staged
    lens[640:900] = 0                      # empty chunks and a chunk that starts empty
    lens[2048:2112] = 40                   # one chunk far over the LDS slice: walked
    lens[3000:3064] = rng.integers(12, 20, 64)  # over 512 elements: walked in FP64, staged in FP32
    rp = np.zeros(n + 1, np.int32)
    np.cumsum(lens, out=rp[1:])
    cols = np.concatenate([np.sort(rng.choice(n, k, replace=False)) for k in lens]).astype(np.int32)
    vals = rng.uniform(-1, 1, len(cols))
    x = rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, n)
    npdt = np.float64 if dtype == "f64" else np.float32
    seg = O.Segments("CSR", lens)
    sv, sc = seg.convert(rp, cols, vals)
    A = tnl.SparseMatrix.from_csr(rp, cols, vals, format="CSRScalar", dtype=dtype)
    for alpha, first, last in [(1.0, 0, 0), (2.5, 0, 0), (1.0, 17, 3001), (1.0, 2050, 2051), (-1.0, 700, n)]:
        print(f"CSRScalar {dtype} alpha={alpha} rows [{first}, {last})", flush=True)
        want = seg.spmv(sv.astype(npdt), sc, x.astype(npdt), y0.astype(npdt), alpha, 0.0, first, last, dtype=npdt)
        got = run_spmv(tnl, A, x.astype(npdt).astype(np.float64), dev, y0.astype(npdt).astype(np.float64),
                       alpha, 0.0, first, last)
        np.testing.assert_array_equal(got, want.astype(np.float64), err_msg=f"{dtype} {alpha} {first} {last}")


@pytest.mark.parametrize("fmt,lanes", [("CSRVector", 64), ("CSRVector", 4), ("CSRLight", 8), ("CSRLight", 2),
                                       ("CSRHybrid", 8), ("CSRHybrid", 16)])
def test_cooperative_csr_kernels_on_short_rows(tnl, dev, fmt, lanes):
    """Rows averaging at most 8 elements run CSRScalar's staged kernel in every CSR variant (the
    Host bits); lanes_per_row set explicitly keeps the cooperative kernel, which agrees with the
    oracle to 1e-13 of sum |a_ij x_j|, for every keeper and row range."""
    rng = np.random.default_rng(31)
    for rp, c, v in (W.laplacian_3d(20), W.irregular(5000, mean=5, sigma=0.8, max_len=200, seed=3)):
        n = len(rp) - 1
        x = rng.uniform(-1, 1, n)
        y0 = rng.uniform(-1, 1, n)
        seg = O.Segments("CSR", np.diff(rp))
        sv, sc = seg.convert(rp, c, v)
        scale = scale_of(rp, c, v, x)
        A = tnl.SparseMatrix.from_csr(rp, c, v, format=fmt, dtype="f64")
        exact = run_spmv(tnl, A, x, dev)
        np.testing.assert_array_equal(exact, seg.spmv(sv, sc, x))  # staged: the Host bits
        A.set_param("lanes_per_row", lanes)
        for alpha, beta, first, last in [(1.0, 0.0, 0, 0), (2.0, -0.5, 17, n - 3)]:
            print(f"{fmt} lanes={lanes} n={n} alpha={alpha} beta={beta} rows [{first}, {last})", flush=True)
            want = seg.spmv(sv, sc, x, y0, alpha, beta, first, last)
            got = run_spmv(tnl, A, x, dev, y0, alpha, beta, first, last)
            tol = FP64_TOL * (abs(alpha) * scale + abs(beta) * np.abs(y0)) + 1e-300
            assert np.all(np.abs(got - want) <= tol), (fmt, lanes, alpha, beta)


@pytest.mark.parametrize("fmt", FORMATS)
def test_keepers_row_ranges_and_value_types(tnl, dev, fmt):
    rng = np.random.default_rng(3)
    rp, c, v = W.irregular(4000, mean=10, sigma=1.3, max_len=1500, seed=8)
    n = len(rp) - 1
    x = rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, n)
    seg = O.Segments("CSR", np.diff(rp))
    sv, sc = seg.convert(rp, c, v)
    scale = scale_of(rp, c, v, x)
    A = tnl.SparseMatrix.from_csr(rp, c, v, format=fmt, dtype="f64")
    for alpha, beta, first, last in [(1.0, 0.0, 0, 0), (2.5, 0.0, 0, 0), (1.0, -0.5, 17, 3001), (-1.0, 1.0, 0, n),
                                     (0.75, 3.0, 1000, 1001)]:
        want = seg.spmv(sv, sc, x, y0, alpha, beta, first, last)
        got = run_spmv(tnl, A, x, dev, y0, alpha, beta, first, last)
        hi = n if last == 0 else last
        untouched = np.r_[0:first, hi:n]
        np.testing.assert_array_equal(got[untouched], y0[untouched])
        err = np.abs(got - want)
        tol = FP64_TOL * (abs(alpha) * scale + abs(beta) * np.abs(y0)) + 1e-300
        assert np.all(err <= tol), (alpha, beta, first, last, np.max(err / tol))
    # FP32 values: oracle in float32 on float32-rounded inputs, tolerance scaled by 2^-20
    A32 = tnl.SparseMatrix.from_csr(rp, c, v, format=fmt, dtype="f32")
    want = seg.spmv(sv.astype(np.float32), sc, x.astype(np.float32), dtype=np.float32).astype(np.float64)
    got = run_spmv(tnl, A32, x, dev)
    assert np.all(np.abs(got - want) <= 2.0 ** -20 * scale * 8 + 1e-30)
    # BF16 values, FP32 compute: values rounded to bfloat16 on the device
    Ab = tnl.SparseMatrix.from_csr(rp, c, v, format=fmt, dtype="bf16")
    vb, cb = Ab.storage()
    vb_full = np.zeros(len(c))
    # the bf16 storage read back holds the rounded values; compare with a float64 product of them
    segb = O.Segments(KIND[fmt], np.diff(rp))
    ref_vals, ref_cols = segb.convert(rp, c, v)
    assert np.array_equal(cb, ref_cols[:len(cb)])
    assert np.all(np.abs(vb - ref_vals[:len(vb)]) <= 2.0 ** -8 * np.abs(ref_vals[:len(vb)]))
    want_b = segb.spmv(vb, cb, x.astype(np.float32).astype(np.float64))
    got = run_spmv(tnl, Ab, x, dev)
    assert np.all(np.abs(got - want_b) <= 2.0 ** -20 * scale * 8 + 1e-30)
    del vb_full


@pytest.mark.parametrize("fmt", FORMATS)
def test_full_size_laplacian_160(tnl, dev, fmt):
    """The benchmark matrix: 160^3 7-point Laplacian.  A*1 is exact (row sums are small
    integers), metadata bit-exact against the oracle, and a random x within tolerance."""
    rp, c, v = W.laplacian_3d(160)
    n = len(rp) - 1
    A = tnl.SparseMatrix.from_csr(rp, c, v, format=fmt, dtype="f64")
    assert A.getNonzeroElementsCount() == 28518400
    seg = O.Segments(KIND[fmt], np.diff(rp))
    for name in O.METADATA_NAMES[O.KINDS[KIND[fmt]]]:
        np.testing.assert_array_equal(A.metadata(name), seg.get(name), err_msg=name)
    y = run_spmv(tnl, A, np.ones(n), dev)
    idx = np.arange(n)
    want = np.zeros(n)
    for d in range(3):
        coord = (idx // 160 ** d) % 160
        want += (coord == 0).astype(float) + (coord == 159).astype(float)
    np.testing.assert_array_equal(y, want)
    x = W.uniform_vector(n, 11)
    ov, oc = O.Segments("CSR", np.diff(rp)).convert(rp, c, v)
    y_ref = O.Segments("CSR", np.diff(rp)).spmv(ov, oc, x)
    check_product(fmt, run_spmv(tnl, A, x, dev), y_ref, scale_of(rp, c, v, x))


def test_device_conversion_between_formats(tnl, dev):
    """B = A across formats on the device gives TNL's storage for B, checked against the oracle's
    conversion chain: B's capacities are A's getRowCapacities -- the stored (non-padding) entries
    of each row, not A's segment sizes (SparseMatrix.hpp:1070-1107) -- and the entries follow A's
    local order."""
    rp, c, v = W.irregular(3000, mean=8, sigma=1.0, max_len=300, seed=12)
    n = len(rp) - 1
    x = W.uniform_vector(n, 5)
    A = tnl.SparseMatrix.from_csr(rp, c, v, format="ChunkedEllpack", dtype="f64")
    src = O.Segments("ChunkedEllpack", np.diff(rp))
    seg_sizes = np.array([src.get_segment_size(r) for r in range(n)], np.int32)
    assert np.any(seg_sizes > np.diff(rp)), "the source must hold padding for the case to mean anything"
    sv, sc = src.convert(rp, c, v)
    caps = np.array([np.sum(sc[[src.global_index(r, l) for l in range(seg_sizes[r])]] != -1) for r in range(n)],
                    np.int32)
    np.testing.assert_array_equal(A.getRowCapacities().cpu().numpy(), caps)
    rp2 = np.zeros(n + 1, np.int32)
    rp2[1:] = np.cumsum(caps)
    c2 = np.full(rp2[-1], -1, np.int32)
    v2 = np.zeros(rp2[-1])
    for r in range(n):
        k = rp2[r]
        for l in range(seg_sizes[r]):
            g = src.global_index(r, l)
            if sc[g] != -1:
                c2[k] = sc[g]
                v2[k] = sv[g]
                k += 1
    for fmt in FORMATS:
        B = A.convert(fmt)
        dst = O.Segments(KIND[fmt], caps)
        for nm in O.METADATA_NAMES[dst.kind]:
            np.testing.assert_array_equal(B.metadata(nm), dst.get(nm), err_msg=f"{fmt} {nm}")
        dv, dc = dst.convert(rp2, c2, v2)
        bv, bc = B.storage()
        np.testing.assert_array_equal(bc, dc[:len(bc)], err_msg=fmt)
        np.testing.assert_array_equal(bv, dv[:len(bv)], err_msg=fmt)
        y = run_spmv(tnl, B, x, dev)
        y_ref = dst.spmv(dv, dc, x)
        check_product("x", y, y_ref, scale_of(rp, c, v, x))


def test_empty_and_degenerate_matrices(tnl, dev):
    for fmt in FORMATS:
        for rows, caps in [(1, [0]), (40, [0] * 40), (33, [1] * 33), (5, [0, 3, 0, 0, 2])]:
            rp = np.zeros(rows + 1, np.int32)
            rp[1:] = np.cumsum(caps)
            c = np.array([k % rows for k in range(rp[-1])], np.int32)
            v = np.arange(1, rp[-1] + 1, dtype=np.float64)
            A = tnl.SparseMatrix.from_csr(rp, c, v, format=fmt)
            x = np.linspace(1, 2, rows)
            y = run_spmv(tnl, A, x, dev, np.full(rows, 7.0), 1.0, 0.0)
            want = W.csr_to_dense(rp, c, v, rows) @ x
            np.testing.assert_allclose(y, want, rtol=1e-15, atol=0, err_msg=f"{fmt} {caps}")


def test_invalid_arguments_raise(tnl, dev):
    rp, c, v = W.laplacian_2d(4)
    A = tnl.SparseMatrix.from_csr(rp, c, v)
    x = torch.zeros(16, dtype=torch.float64, device=dev)
    with pytest.raises(ValueError):
        A.vectorProduct(x, torch.zeros(15, dtype=torch.float64, device=dev))
    with pytest.raises(tnl.TnlHipError):
        A.vectorProduct(x, torch.zeros(16, dtype=torch.float64, device=dev), 1.0, 0.0, 5, 3)
    bad = c.copy()
    bad[3] = 99
    with pytest.raises(tnl.TnlHipError):
        tnl.SparseMatrix.from_csr(rp, bad, v)


@pytest.mark.parametrize("fmt,params", [("Ellpack", {"xcd_remap": 3}), ("Ellpack", {"xcd_remap": 3, "pair_rows": 1}),
                                        ("SlicedEllpack", {"xcd_remap": 3}),
                                        ("SlicedEllpack", {"xcd_remap": 3, "pair_rows": 1, "blocks_per_cu": 4}),
                                        ("CSRAdaptive", {"xcd_remap": 1}), ("CSRAdaptive", {"xcd_remap": 3}),
                                        ("CSRScalar", {"xcd_remap": 1}), ("ChunkedEllpack", {"xcd_remap": 1}),
                                        ("SlicedEllpack", {"xcd_remap": 1})])
def test_launch_distributions_give_the_same_bits(tnl, dev, fmt, params):
    """Row distributions over the XCDs change only which workgroup computes a row: every keeper
    mode and row range gives the default launch's bits (a row computed twice or skipped would
    show in the beta != 0 cases)."""
    rng = np.random.default_rng(5)
    for rp, c, v in (W.irregular(30001, mean=10, sigma=1.3, max_len=1500, seed=9), W.laplacian_3d(33)):
        n = len(rp) - 1
        x = rng.uniform(-1, 1, n)
        y0 = rng.uniform(-1, 1, n)
        A = tnl.SparseMatrix.from_csr(rp, c, v, format=fmt, dtype="f64")
        B = tnl.SparseMatrix.from_csr(rp, c, v, format=fmt, dtype="f64")
        for k, val in params.items():
            B.set_param(k, val)
        for alpha, beta, first, last in [(1.0, 0.0, 0, 0), (1.0, -0.5, 17, n - 5), (0.75, 3.0, 1000, 1001)]:
            want = run_spmv(tnl, A, x, dev, y0, alpha, beta, first, last)
            got = run_spmv(tnl, B, x, dev, y0, alpha, beta, first, last)
            np.testing.assert_array_equal(got, want, err_msg=f"{fmt} {params} {alpha} {beta} {first} {last}")


@pytest.mark.parametrize("fmt", FORMATS)
def test_matrix_market_read_into_every_format(tnl, dev, fmt, tmp_path):
    """readMtx into a device matrix of each format (MatrixReader.hpp:19-26: the host read, then
    matrix = hostMatrix): capacities are the host matrix's stored entries per row, explicit zeros
    are dropped; y = A x as the dense product; writeMtx from the device (a host copy by operator=,
    then the host writer) gives TNL's text with the header counting the nonzero entries."""
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mtx.npz"))
    for name in (str(n) for n in d["names"]):
        if f"{name}/error" in d.files:
            continue
        path = tmp_path / f"{name}.mtx"
        path.write_text(str(d[f"{name}/text"]))
        A = tnl.SparseMatrix.from_mtx(path, format=fmt, dtype="f64")
        rows, cols = (int(x) for x in d[f"{name}/dims"])
        rp, ci, v = d[f"{name}/row_ptr"], d[f"{name}/cols"], d[f"{name}/vals"]
        assert (A.info.rows, A.info.columns) == (rows, cols)
        assert A.info.nonzeros == int(np.sum((ci != -1) & (v != 0))), name
        stored = np.array([np.sum(ci[rp[r]:rp[r + 1]] != -1) for r in range(rows)], np.int32)
        seg = O.Segments(KIND[fmt], stored)
        for nm in O.METADATA_NAMES[seg.kind]:
            np.testing.assert_array_equal(A.metadata(nm), seg.get(nm), err_msg=f"{name} {nm}")
        sv, sc = seg.convert(rp, ci, v)
        bv, bc = A.storage()
        np.testing.assert_array_equal(bc, sc[:len(bc)], err_msg=name)
        np.testing.assert_array_equal(bv, sv[:len(bv)], err_msg=name)
        x = np.random.default_rng(1).uniform(-1, 1, cols)
        dense = np.zeros((rows, cols))
        for r in range(rows):
            for k in range(rp[r], rp[r + 1]):
                if ci[k] >= 0:
                    dense[r, ci[k]] = v[k]
        got = run_spmv(tnl, A, x, dev)
        want = dense @ x
        assert np.all(np.abs(got - want) <= 1e-13 * (np.abs(dense) @ np.abs(x)) + 1e-300), name
        out = tmp_path / f"{name}.out.mtx"
        A.write_mtx(out)
        assert out.read_text() == str(d[f"{name}/written_device"]), name


def _symmetric_random(n, seed):
    """Lower triangle of A + A^T for an irregular A."""
    rp, c, v = W.irregular(n, mean=10, sigma=1.0, max_len=60, seed=seed)
    A = W.csr_to_dense(rp, c, v, n)
    L = np.tril(A + A.T)
    rows, cols = np.nonzero(L)
    rp2 = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=n))]).astype(np.int32)
    return rp2, cols.astype(np.int32), L[rows, cols]


@pytest.mark.parametrize("fmt", FORMATS)
def test_symmetric_products(tnl, dev, fmt):
    """SymmetricMatrix products (transposed entries added with atomics): TNL's unit-test values
    exactly, and the oracle (TNL's Host order) to |dy_i| <= 1e-13 * (|alpha| sum_j |a_ij x_j| +
    |beta y_i|) over the full matrix, for the keepers and a row range."""
    for (rp, c, v), x, y in tnl_cases.symmetric_matrices():
        A = tnl.SparseMatrix.symmetric_from_csr(rp, c, v, format=fmt)
        assert A.symmetric
        np.testing.assert_array_equal(run_spmv(tnl, A, x, dev), np.array(y, np.float64))
    n = 700
    rp, c, v = _symmetric_random(n, 13)
    L = W.csr_to_dense(rp, c, v, n)
    full = L + L.T - np.diag(np.diag(L))
    x = W.uniform_vector(n, 7)
    y0 = W.uniform_vector(n, 8)
    A = tnl.SparseMatrix.symmetric_from_csr(rp, c, v, format=fmt)
    for alpha, beta, first, last in ((1.0, 0.0, 0, 0), (-1.0, 1.0, 0, 0), (0.5, 2.0, 0, 0), (1.0, 0.0, 200, 450)):
        print(f"symmetric {fmt} {alpha} {beta} {first} {last}", flush=True)
        got = run_spmv(tnl, A, x, dev, y0, alpha, beta, first, last)
        want = O.symmetric_spmv(rp, c, v, x, y0, alpha, beta, first, last)
        scale = abs(alpha) * np.abs(full) @ np.abs(x) + abs(beta) * np.abs(y0)
        assert np.all(np.abs(got - want) <= 1e-13 * scale + 1e-300), (fmt, alpha, beta, first, last)
    A32 = tnl.SparseMatrix.symmetric_from_csr(rp, c, v, format=fmt, dtype="f32")
    got32 = run_spmv(tnl, A32, x, dev)
    assert np.all(np.abs(got32 - full @ x) <= 1e-5 * (np.abs(full) @ np.abs(x)))
    bad_c = c.copy()
    bad_c[np.argmax(np.diff(rp) > 0)] = n - 1
    with pytest.raises(tnl.TnlHipError, match="lower triangle"):
        tnl.SparseMatrix.symmetric_from_csr(rp, bad_c, v, format=fmt)
    B = A.convert("SlicedEllpack")
    assert B.symmetric


@pytest.mark.parametrize("fmt", FORMATS)
def test_binary_products(tnl, dev, fmt):
    """Binary matrices (no values, int32 vectors): TNL's unit-test values and the oracle (TNL's
    product) exactly, for the keepers and a row range; explicit zeros of the source are dropped,
    the stored entries read back as ones, and the Krylov solvers refuse the type."""
    for (rp, c), xval, y in tnl_cases.binary_matrices():
        n = len(rp) - 1
        A = tnl.SparseMatrix.from_csr(rp, c, None, format=fmt, dtype="binary")
        xd = torch.full((n,), xval, dtype=torch.int32, device=dev)
        yd = torch.zeros(n, dtype=torch.int32, device=dev)
        A.vectorProduct(xd, yd)
        np.testing.assert_array_equal(yd.cpu().numpy(), np.array(y, np.int32))
    rng = np.random.default_rng(5)
    rp, c, v = W.irregular(500, mean=8, sigma=1.0, max_len=40, seed=6)
    v = v.copy()
    v[::7] = 0.0  # dropped by the conversion, as operator= drops explicit zeros
    keep = v != 0
    lens = np.add.reduceat(keep.astype(np.int64), rp[:-1]) * (np.diff(rp) > 0)
    rp_k = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    A = tnl.SparseMatrix.from_csr(rp, c, v, format=fmt, dtype="binary")
    assert A.value_dtype == "binary" and A.getNonzeroElementsCount() == int(keep.sum())
    vals, cols = A.storage()
    assert np.all(vals[cols >= 0] == 1.0)
    x = rng.integers(-1000, 1000, 500).astype(np.int32)
    y0 = rng.integers(-1000, 1000, 500).astype(np.int32)
    for alpha, beta, first, last in ((1, 0, 0, 0), (3, 0, 0, 0), (1, 2, 0, 0), (-2, 5, 50, 200)):
        print(f"binary {fmt} {alpha} {beta} {first} {last}", flush=True)
        xd = torch.tensor(x, device=dev)
        yd = torch.tensor(y0, device=dev)
        A.vectorProduct(xd, yd, alpha, beta, first, last)
        want = O.binary_spmv(rp_k, c[keep], x, y0, alpha, beta, first, last)
        np.testing.assert_array_equal(yd.cpu().numpy(), want)
    B = A.convert("BiEllpack")
    assert B.value_dtype == "binary"
    s = tnl.getLinearSolver("cg")
    s.setMatrix(A)
    with pytest.raises(tnl.TnlHipError, match="binary"):
        s.solve(torch.zeros(500, dtype=torch.float32, device=dev), torch.zeros(500, dtype=torch.float32, device=dev))
