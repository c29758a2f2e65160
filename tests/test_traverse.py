"""Row traversals of the oracle (reduceRows, getRowCapacities, getCompressedRowLengths,
getRowCapacity) pinned against TNL's own Host results (the reference build, when present) and
TNL's unit-test values (SparseMatrixTest.hpp:241-297, SegmentsTest.hpp:101-150)."""
import numpy as np
import pytest

import pyoracle as O

KINDS = list(O.KINDS)
PAIRS = [("value", "sum", 0.0), ("abs", "max", 0.0), ("value", "min", np.inf), ("column", "max", -1.0),
         ("one", "sum", 0.0), ("nonzero", "sum", 0.0), ("value", "max", -np.inf)]


def set_row_capacities_case():
    """SparseMatrixTest.hpp:241-297: 10 x 11, capacities {3, 3, 1, 2, ..., 8}, row i filled
    with columns 0..cap-1 and values 1, 2, 3, ..."""
    caps = [3, 3, 1, 2, 3, 4, 5, 6, 7, 8]
    rp = np.zeros(11, np.int32)
    rp[1:] = np.cumsum(caps)
    cols = np.concatenate([np.arange(c) for c in caps]).astype(np.int32)
    vals = np.arange(1, rp[-1] + 1, dtype=np.float64)
    return rp, cols, vals, 11


def irregular_with_padding_and_zeros(seed=3, n=300, ncols=257):
    """CSR rows of random length with padding slots (-1) after the entries and explicit zeros."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 40, n)
    lens[::17] = 0
    pad = rng.integers(0, 3, n)
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum(lens + pad)
    cols, vals = [], []
    for L, p in zip(lens, pad):
        c = np.sort(rng.choice(ncols, L, replace=False))
        v = rng.standard_normal(L)
        v[rng.random(L) < 0.15] = 0.0
        cols += list(c) + [-1] * p
        vals += list(v) + [0.0] * p
    return rp, np.array(cols, np.int32), np.array(vals, np.float64), ncols


def stored_counts(rp, cols):
    """getRowCapacities of a CSR: its stored (non-padding) entries per row."""
    return np.array([np.sum(cols[rp[r]:rp[r + 1]] != -1) for r in range(len(rp) - 1)], np.int32)


def oracle_storage(kind, rp, cols, vals):
    """Storage of the matrix of `kind` holding the CSR: the CSR arrays themselves for CSR (explicit
    zeros and padding kept), TNL's operator= for the others -- capacities from the source's
    getRowCapacities (stored entries per row), zeros dropped."""
    if kind == "CSR":
        return O.Segments(kind, np.diff(rp)), vals, cols
    s = O.Segments(kind, stored_counts(rp, cols))
    v, c = s.convert(rp, cols, vals)
    return s, v, c


@pytest.mark.parametrize("kind", KINDS)
def test_compressed_row_lengths_match_tnl_unit_test(kind):
    rp, cols, vals, ncols = set_row_capacities_case()
    s, v, c = oracle_storage(kind, rp, cols, vals)
    lens = O.reduce_rows(s, v, c, "nonzero", "sum", 0.0)
    np.testing.assert_array_equal(lens, [3, 3, 1, 2, 3, 4, 5, 6, 7, 8])


@pytest.mark.parametrize("kind", KINDS)
def test_maximum_in_segments_matches_tnl_unit_test(kind):
    # SegmentsTest.hpp:101-150: 20 segments of 5, element (i, l) = 5 i + l + 1, max = 5 (i + 1)
    n, size = 20, 5
    rp = np.arange(0, n * size + 1, size, dtype=np.int32)
    cols = np.tile(np.arange(size, dtype=np.int32), n)
    vals = np.array([i * size + l + 1 for i in range(n) for l in range(size)], np.float64)
    s, v, c = oracle_storage(kind, rp, cols, vals)
    mx = O.reduce_rows(s, v, c, "value", "max", float(np.iinfo(np.int32).min))
    np.testing.assert_array_equal(mx, (np.arange(n) + 1) * size)


def test_planted_mistake_is_caught():
    # a fetch that forgot to skip padding counts the padding slots: the stored-entry count differs
    rp, cols, vals, _ = irregular_with_padding_and_zeros()
    s = O.Segments("CSR", np.diff(rp))
    counted = O.reduce_rows(s, vals, cols, "one", "sum", 0.0)
    assert np.any(counted != np.diff(rp)), "the case must hold padding"
    np.testing.assert_array_equal(counted, np.add.reduceat(cols >= 0, rp[:-1]) * (np.diff(rp) > 0))


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("case", ["unit", "irregular"])
def test_row_reductions_match_tnl(kind, case):
    rp, cols, vals, ncols = set_row_capacities_case() if case == "unit" else irregular_with_padding_and_zeros()
    s, v, c = oracle_storage(kind, rp, cols, vals)
    for fetch, reduce, ident in PAIRS:
        ref, caps, lens, sizes = O.ref_row_reductions(kind, rp, cols, vals, ncols, fetch, reduce, ident)
        got = O.reduce_rows(s, v, c, fetch, reduce, ident)
        np.testing.assert_array_equal(got, ref, err_msg=f"{kind} {case} {fetch}/{reduce}")
    np.testing.assert_array_equal(O.reduce_rows(s, v, c, "one", "sum", 0.0), caps)
    np.testing.assert_array_equal(O.reduce_rows(s, v, c, "nonzero", "sum", 0.0), lens)
    np.testing.assert_array_equal([s.get_segment_size(r) for r in range(s.n)], sizes)
