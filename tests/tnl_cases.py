"""Matrices of TNL's SpMV unit tests (src/UnitTests/Matrices/SparseMatrixVectorProductTest.hpp)
as data: row capacities, (row, column, value) entries, input vector value and the expected
product.  Entries are listed in ascending column order per row, the order setElement keeps."""
import numpy as np


def _entries(rows):
    out = []
    for r, row in enumerate(rows):
        for c, v in row:
            out.append((r, c, v))
    return out


def small_matrix_1():
    # SparseMatrixVectorProductTest.hpp:17-68
    rows = [[(0, 1)], [(1, 2), (3, 3)], [(1, 4)], [(2, 5)]]
    return dict(n=4, m=4, caps=[1, 2, 1, 1], entries=_entries(rows), x=2.0, y=[2, 10, 8, 10])


def small_matrix_2():
    # SparseMatrixVectorProductTest.hpp:70-122
    rows = [[(0, 1), (1, 2), (2, 3)], [(3, 4)], [(0, 5), (1, 6), (2, 7)], [(1, 8)]]
    return dict(n=4, m=4, caps=[3, 1, 3, 1], entries=_entries(rows), x=2.0, y=[12, 8, 36, 16])


def small_matrix_3():
    # SparseMatrixVectorProductTest.hpp:124-177
    rows = [[(0, 1), (1, 2), (2, 3)], [(1, 4), (2, 5), (3, 6)], [(0, 7), (1, 8), (2, 9)], [(1, 10), (2, 11), (3, 12)]]
    return dict(n=4, m=4, caps=[3, 3, 3, 3], entries=_entries(rows), x=2.0, y=[12, 30, 48, 66])


def medium_matrix_1():
    # SparseMatrixVectorProductTest.hpp:179-254
    rows = [[(0, 1), (1, 2), (2, 3), (5, 4)], [(1, 5), (2, 6), (3, 7), (4, 8)],
            [(0, 9), (1, 10), (2, 11), (3, 12), (4, 13)], [(1, 14), (2, 15), (3, 16), (4, 17)],
            [(2, 18), (3, 19), (4, 20), (5, 21)], [(3, 22), (4, 23), (5, 24), (6, 25)],
            [(0, 26), (1, 27), (2, 28), (3, 29), (4, 30)], [(0, 31), (1, 32), (2, 33), (3, 34), (4, 35)]]
    return dict(n=8, m=8, caps=[4, 4, 5, 4, 4, 4, 5, 5], entries=_entries(rows), x=2.0,
                y=[20, 52, 110, 124, 156, 188, 280, 330])


def medium_matrix_2():
    # SparseMatrixVectorProductTest.hpp:256-336
    rows = [[(0, 1), (1, 2), (2, 3), (4, 4), (5, 5), (7, 1)], [(1, 6), (3, 7), (7, 1)],
            [(1, 8), (2, 9), (4, 10), (7, 1)], [(1, 11), (2, 12), (3, 13), (4, 14), (7, 1)], [(1, 15), (7, 1)],
            [(1, 16), (2, 17), (3, 18), (4, 19), (5, 20), (6, 21), (7, 1)],
            [(0, 22), (1, 23), (2, 24), (3, 25), (4, 26), (5, 27), (6, 28), (7, 1)],
            [(c, 29 + c) for c in range(8)]]
    return dict(n=8, m=8, caps=[6, 3, 4, 5, 2, 7, 8, 8], entries=_entries(rows), x=2.0,
                y=[32, 28, 56, 102, 32, 224, 352, 520])


def large_diagonal(n=1051):
    # SparseMatrixVectorProductTest.hpp:339-372: a_ii = i + 1, x = 1 -> y_i = i + 1
    return dict(n=n, m=n, caps=[1] * n, entries=[(i, i, float(i + 1)) for i in range(n)], x=1.0,
                y=[i + 1 for i in range(n)])


def large_lower(n=1051):
    # SparseMatrixVectorProductTest.hpp:374-395: a_ij = j + 1 for j <= i -> y_i = (i+1)(i+2)/2
    ent = [(i, j, float(j + 1)) for i in range(n) for j in range(i + 1)]
    return dict(n=n, m=n, caps=[i + 1 for i in range(n)], entries=ent, x=1.0,
                y=[(i + 1) * (i + 2) // 2 for i in range(n)])


def long_rows(columns):
    # SparseMatrixVectorProductTest.hpp:397-432: 33 rows, a_rl = l + r -> y_r = c(c-1)/2 + c r
    rows = 33
    ent = [(r, l, float(l + r)) for r in range(rows) for l in range(columns)]
    return dict(n=rows, m=columns, caps=[columns] * rows, entries=ent, x=1.0,
                y=[columns * (columns - 1) / 2.0 + columns * r for r in range(rows)])


LONG_ROW_COLUMNS = [64, 65, 128, 129, 256, 257, 512, 513, 1024, 1025, 2048, 2049, 3000]


def all_cases():
    cases = {"small1": small_matrix_1(), "small2": small_matrix_2(), "small3": small_matrix_3(),
             "medium1": medium_matrix_1(), "medium2": medium_matrix_2(), "large_diagonal": large_diagonal(),
             "large_lower": large_lower()}
    for c in LONG_ROW_COLUMNS:
        cases[f"long_rows_{c}"] = long_rows(c)
    return cases


def to_csr(case):
    """CSR arrays with the given capacities (entries first, then explicit zeros as filler
    so the capacity is kept: zeros are dropped by the conversion and become padding)."""
    n = case["n"]
    caps = np.asarray(case["caps"], np.int64)
    rp = np.zeros(n + 1, np.int64)
    np.cumsum(caps, out=rp[1:])
    cols = np.zeros(rp[-1], np.int32)
    vals = np.zeros(rp[-1], np.float64)
    fill = rp[:-1].copy()
    for r, c, v in case["entries"]:
        cols[fill[r]] = c
        vals[fill[r]] = v
        fill[r] += 1
    for r in range(n):  # filler slots: column of the last entry, value 0 (dropped on conversion)
        for k in range(fill[r], rp[r + 1]):
            cols[k] = 0
            vals[k] = 0.0
    return rp.astype(np.int32), cols, vals


def symmetric_matrices():
    """SymmetricSparseMatrixTest.hpp:679-851 (test_VectorProduct): the lower triangle (TNL's
    setElement moves an entry above the diagonal to its mirror), the input vector and TNL's
    expected product."""
    def csr(n, entries):
        entries = sorted((r, c, v) if c <= r else (c, r, v) for r, c, v in entries)
        merged = {}
        for r, c, v in entries:
            merged[(r, c)] = v
        rp = np.zeros(n + 1, np.int32)
        for r, _ in merged:
            rp[r + 1] += 1
        rp = np.cumsum(rp).astype(np.int32)
        cols = np.array([c for _, c in merged], np.int32)
        vals = np.array(list(merged.values()), np.float64)
        return rp, cols, vals

    out = []
    out.append((csr(4, [(0, 0, 1), (1, 1, 2), (2, 1, 3), (3, 1, 4), (3, 2, 5)]), np.full(4, 2.0),
                [2, 18, 16, 18]))
    out.append((csr(4, [(0, 0, 1), (1, 0, 2), (2, 0, 3), (2, 1, 6), (2, 2, 7), (3, 1, 8), (3, 3, 9)]),
                np.full(4, 2.0), [12, 32, 32, 34]))
    out.append((csr(4, [(0, 0, 1), (0, 1, 2), (0, 2, 3), (0, 3, 4), (1, 0, 2), (1, 1, 5), (2, 0, 3), (2, 2, 6),
                        (3, 0, 4), (3, 3, 7)]), np.array([0.0, 1, 2, 3]), [20, 5, 12, 21]))
    out.append((csr(8, [(0, 0, 1), (1, 1, 2), (2, 0, 3), (2, 2, 5), (3, 1, 6), (3, 3, 8), (4, 0, 9), (4, 2, 10),
                        (4, 4, 11), (5, 1, 12), (5, 3, 13), (5, 5, 14), (6, 0, 15), (6, 2, 16), (6, 4, 17),
                        (6, 6, 18), (7, 1, 19), (7, 3, 20), (7, 5, 21), (7, 7, 22)]),
                np.array([1.0, 2, 1, 2, 1, 2, 1, 2]), [28, 78, 34, 94, 47, 120, 66, 164]))
    out.append((csr(8, [(0, 0, 1), (1, 1, 2), (2, 2, 3), (3, 2, 4), (3, 3, 5), (4, 2, 6), (4, 3, 7), (4, 4, 8),
                        (5, 2, 9), (5, 3, 10), (5, 4, 11), (5, 5, 12), (6, 6, 13), (7, 7, 14)]),
                np.arange(1.0, 9.0),
                [1, 4, 3 * 3 + 4 * 4 + 5 * 6 + 6 * 9, 3 * 4 + 4 * 5 + 5 * 7 + 6 * 10, 3 * 6 + 4 * 7 + 5 * 8 + 6 * 11,
                 3 * 9 + 4 * 10 + 5 * 11 + 6 * 12, 7 * 13, 8 * 14]))
    return out


def binary_matrices():
    """BinarySparseMatrixTest.hpp:584-790 (test_VectorProduct): the column pattern per row, the
    constant input vector and TNL's expected product (ComputeReal = int)."""
    def csr(rows):
        rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
        return rp, np.array([c for r in rows for c in r], np.int32)

    m4 = [[0, 1, 2, 5], [1, 2, 3, 4], [0, 1, 2, 3, 4], [1, 2, 3, 4], [2, 3, 4, 5], [3, 4, 5, 6], [0, 1, 2, 3, 4],
          [0, 1, 2, 3, 4]]
    return [(csr([[0], [1, 3], [1], [2]]), 2, [2, 4, 2, 2]),
            (csr([[0, 1, 2], [3], [0, 1, 2], [1]]), 2, [6, 2, 6, 2]),
            (csr([[0, 1, 2], [1, 2, 3], [0, 1, 2], [1, 2, 3]]), 2, [6, 6, 6, 6]),
            (csr(m4), 2, [8, 8, 10, 8, 8, 8, 10, 10])]
