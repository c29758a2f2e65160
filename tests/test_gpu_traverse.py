"""Row traversals on the GPU through the C ABI against the oracle (itself pinned to TNL's Host
results): reduceRows for every fetch / reduce pair, getRowCapacities, getCompressedRowLengths,
getRowCapacity and getRow in every format, bit for bit (one lane per row in TNL's order); the
storage TNL's operator= and its element-by-element construction give a CSR with padding slots and
explicit zeros."""
import numpy as np
import pytest
import torch

import pyoracle as O
from test_traverse import PAIRS, irregular_with_padding_and_zeros, set_row_capacities_case

pytestmark = pytest.mark.gpu

FORMATS = ["CSRScalar", "CSRAdaptive", "Ellpack", "SlicedEllpack", "ChunkedEllpack", "BiEllpack"]
KIND = {"CSRScalar": "CSR", "CSRAdaptive": "CSR", "Ellpack": "Ellpack", "SlicedEllpack": "SlicedEllpack",
        "ChunkedEllpack": "ChunkedEllpack", "BiEllpack": "BiEllpack"}


def stored_counts(rp, cols):
    return np.array([np.sum(cols[rp[r]:rp[r + 1]] != -1) for r in range(len(rp) - 1)], np.int32)


def check_traversals(A, seg, v, c, label):
    n = seg.n
    for fetch, reduce, ident in PAIRS:
        want = O.reduce_rows(seg, v, c, fetch, reduce, ident)
        for dt in (torch.float64, torch.float32):
            out = torch.full((n,), 7.0, dtype=dt, device="cuda")
            A.reduceAllRows(fetch, reduce, ident, out)
            np.testing.assert_array_equal(out.cpu().numpy(), want.astype(out.cpu().numpy().dtype),
                                          err_msg=f"{label} {fetch}/{reduce} {dt}")
    lo, hi = n // 3, n - n // 4
    part = torch.full((n,), -5, dtype=torch.int32, device="cuda")
    A.reduceRows(lo, hi, "one", "sum", 0.0, part)
    p = part.cpu().numpy()
    np.testing.assert_array_equal(p[lo:hi], O.reduce_rows(seg, v, c, "one", "sum", 0.0)[lo:hi], err_msg=label)
    assert np.all(p[:lo] == -5) and np.all(p[hi:] == -5), "rows outside the range are left alone"
    np.testing.assert_array_equal(A.getRowCapacities().cpu().numpy(), O.reduce_rows(seg, v, c, "one", "sum", 0.0),
                                  err_msg=label)
    np.testing.assert_array_equal(A.getCompressedRowLengths().cpu().numpy(),
                                  O.reduce_rows(seg, v, c, "nonzero", "sum", 0.0), err_msg=label)
    for r in sorted({0, n // 2, n - 1, 7 % n}):
        size = seg.get_segment_size(r)
        assert A.getRowCapacity(r) == size, label
        cols, vals = A.getRow(r)
        g = [seg.global_index(r, l) for l in range(size)]
        np.testing.assert_array_equal(cols, c[g], err_msg=f"{label} row {r}")
        np.testing.assert_array_equal(vals, np.where(c[g] == -1, 0.0, v[g]), err_msg=f"{label} row {r}")


@pytest.mark.parametrize("fmt", FORMATS)
def test_traversals_after_operator_assign(tnl, dev, fmt):
    """A = csr: capacities from the source's stored entries, zeros dropped (SparseMatrix.hpp:1070-1107)."""
    for label, (rp, cols, vals, ncols) in (("unit", set_row_capacities_case()),
                                           ("irregular", irregular_with_padding_and_zeros())):
        print(f"case {label} {fmt}", flush=True)
        A = tnl.SparseMatrix.from_csr(rp, cols, vals, columns=ncols, format=fmt, dtype="f64")
        seg = O.Segments(KIND[fmt], stored_counts(rp, cols))
        v, c = seg.convert(rp, cols, vals)
        for nm in O.METADATA_NAMES[seg.kind]:
            np.testing.assert_array_equal(A.metadata(nm), seg.get(nm), err_msg=f"{label} {nm}")
        bv, bc = A.storage()
        np.testing.assert_array_equal(bc, c[:len(bc)])
        np.testing.assert_array_equal(bv, v[:len(bv)])
        check_traversals(A, seg, v, c, f"{label} {fmt}")


@pytest.mark.parametrize("fmt", FORMATS)
def test_traversals_after_element_by_element_build(tnl, dev, fmt):
    """setRowCapacities + setElement: capacities as set, padding slots kept, explicit zeros stored."""
    rp, cols, vals, ncols = irregular_with_padding_and_zeros(seed=9)
    A = tnl.SparseMatrix.staged(rp, cols, vals, columns=ncols, format=fmt, dtype="f64")
    seg = O.Segments(KIND[fmt], np.diff(rp))
    for nm in O.METADATA_NAMES[seg.kind]:
        np.testing.assert_array_equal(A.metadata(nm), seg.get(nm), err_msg=nm)
    # the staged storage: each row's entries (zeros included) in order, then padding
    keep = cols != -1
    rp2 = np.concatenate([[0], np.cumsum(np.diff(rp))]).astype(np.int32)
    c2 = np.full(len(cols), -1, np.int32)
    v2 = np.zeros(len(cols))
    for r in range(len(rp) - 1):
        k = rp[r]
        for g in range(rp[r], rp[r + 1]):
            if keep[g]:
                c2[k], v2[k] = cols[g], vals[g]
                k += 1
    n = len(rp) - 1
    sv = np.zeros(max(seg.storage_size(), 1))
    sc = np.full(max(seg.storage_size(), 1), -1, np.int32)
    for r in range(n):
        for l in range(rp2[r + 1] - rp2[r]):
            g = seg.global_index(r, l)
            sc[g], sv[g] = c2[rp2[r] + l], v2[rp2[r] + l]
    bv, bc = A.storage()
    np.testing.assert_array_equal(bc, sc[:len(bc)])
    np.testing.assert_array_equal(bv, sv[:len(bv)])
    assert A.getNonzeroElementsCount() == int(keep.sum())
    check_traversals(A, seg, sv, sc, fmt)


def test_compressed_row_lengths_of_tnl_unit_test(tnl, dev):
    """SparseMatrixTest.hpp:241-297 on the device: {3, 3, 1, 2, 3, 4, 5, 6, 7, 8}."""
    rp, cols, vals, ncols = set_row_capacities_case()
    for fmt in FORMATS:
        A = tnl.SparseMatrix.staged(rp, cols, vals, columns=ncols, format=fmt, dtype="f64")
        np.testing.assert_array_equal(A.getCompressedRowLengths().cpu().numpy(), [3, 3, 1, 2, 3, 4, 5, 6, 7, 8])


def test_maximum_per_row_of_tnl_unit_test(tnl, dev):
    """SegmentsTest.hpp:101-150 on the device: 20 rows of 5, (i, l) = 5 i + l + 1, max = 5 (i + 1)."""
    n, size = 20, 5
    rp = np.arange(0, n * size + 1, size, dtype=np.int32)
    cols = np.tile(np.arange(size, dtype=np.int32), n)
    vals = np.array([i * size + l + 1 for i in range(n) for l in range(size)], np.float64)
    for fmt in FORMATS:
        A = tnl.SparseMatrix.from_csr(rp, cols, vals, columns=size, format=fmt, dtype="f64")
        out = torch.zeros(n, dtype=torch.int32, device="cuda")
        A.reduceAllRows("value", "max", float(np.iinfo(np.int32).min), out)
        np.testing.assert_array_equal(out.cpu().numpy(), (np.arange(n) + 1) * size, err_msg=fmt)
