"""Synthetic inputs of the benchmark configurations, as host CSR arrays (numpy).

Every matrix is returned as (row_ptr int32[n+1], cols int32[nnz], vals float64[nnz]) with the
columns of each row in ascending order, which is the order TNL's setElement keeps
(src/TNL/Matrices/SparseMatrixView.hpp:270-345).
"""
import numpy as np


def laplacian_2d(N):
    """5-point Laplacian on an N x N grid, Dirichlet truncation: 4 on the diagonal,
    -1 at the S, W, E, N neighbours; row index = i + N*j."""
    return _stencil((N, N), [((0, -1), -1.0), ((-1, 0), -1.0), ((0, 0), 4.0), ((1, 0), -1.0), ((0, 1), -1.0)])


def laplacian_3d(N):
    """7-point Laplacian on an N^3 grid: 6 on the diagonal, -1 at the six neighbours;
    row index = i + N*j + N*N*k."""
    o = [((0, 0, -1), -1.0), ((0, -1, 0), -1.0), ((-1, 0, 0), -1.0), ((0, 0, 0), 6.0),
         ((1, 0, 0), -1.0), ((0, 1, 0), -1.0), ((0, 0, 1), -1.0)]
    return _stencil((N, N, N), o)


def convection_diffusion_3d(N, peclet, velocity=(1.0, 0.5, 0.25)):
    """-Laplace(u) + Pe * (v . grad u), first-order upwind (v > 0: backward differences),
    on an N^3 grid with h = 1/(N+1), rows scaled by h^2:
      diag 6 + Pe*h*(vx+vy+vz); west/south/down -1 - Pe*h*v; east/north/up -1."""
    h = 1.0 / (N + 1)
    vx, vy, vz = velocity
    o = [((0, 0, -1), -1.0 - peclet * h * vz), ((0, -1, 0), -1.0 - peclet * h * vy),
         ((-1, 0, 0), -1.0 - peclet * h * vx), ((0, 0, 0), 6.0 + peclet * h * (vx + vy + vz)),
         ((1, 0, 0), -1.0), ((0, 1, 0), -1.0), ((0, 0, 1), -1.0)]
    return _stencil((N, N, N), o)


def _stencil(shape, offsets):
    """Assemble a constant-coefficient stencil on a box grid (x fastest)."""
    dims = len(shape)
    n = int(np.prod(shape))
    idx = np.arange(n, dtype=np.int64)
    coords = []
    rem = idx.copy()
    for d in range(dims):
        coords.append(rem % shape[d])
        rem //= shape[d]
    strides = [int(np.prod(shape[:d])) for d in range(dims)]
    # offsets are listed in ascending linear-offset order
    lin = [sum(o[d] * strides[d] for d in range(dims)) for o, _ in offsets]
    order = np.argsort(lin, kind="stable")
    offsets = [offsets[i] for i in order]
    valid = []
    for o, _ in offsets:
        ok = np.ones(n, dtype=bool)
        for d in range(dims):
            c = coords[d] + o[d]
            ok &= (c >= 0) & (c < shape[d])
        valid.append(ok)
    valid = np.stack(valid, axis=1)  # n x s
    counts = valid.sum(axis=1)
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=row_ptr[1:])
    col_mat = np.stack([idx + sum(o[d] * strides[d] for d in range(dims)) for o, _ in offsets], axis=1)
    val_mat = np.broadcast_to(np.array([v for _, v in offsets], dtype=np.float64), col_mat.shape)
    cols = col_mat[valid].astype(np.int32)
    vals = np.ascontiguousarray(val_mat[valid])
    return row_ptr.astype(np.int32), cols, vals


def irregular(n, mean=16.0, sigma=1.2, max_len=4096, seed=0):
    """Irregular SPD-like matrix: off-diagonal row lengths round(lognormal) with
    E[L] = mean (mu = ln(mean) - sigma^2/2), clipped to [1, max_len]; columns uniform
    without repeats and excluding the diagonal; values N(0,1); diagonal = 1 + sum|offdiag|."""
    rng = np.random.default_rng(seed)
    mu = np.log(mean) - sigma * sigma / 2.0
    L = np.clip(np.rint(rng.lognormal(mu, sigma, n)), 1, min(max_len, n - 1)).astype(np.int64)
    rows = np.repeat(np.arange(n, dtype=np.int64), L)
    total = int(L.sum())
    cols = rng.integers(0, n - 1, total, dtype=np.int64)
    cols += cols >= rows  # skip the diagonal
    # key = row * n + col sorted: rows stay grouped, columns ascend; redraw repeats
    key = rows * n + cols
    del cols
    key.sort()
    for _ in range(64):
        dup = np.flatnonzero(key[1:] == key[:-1]) + 1
        if len(dup) == 0:
            break
        r = key[dup] // n
        c = rng.integers(0, n - 1, len(dup), dtype=np.int64)
        c += c >= r
        key[dup] = r * n + c
        key.sort()
    del rows
    # the diagonal joins each row in column order
    key = np.concatenate([key, np.arange(n, dtype=np.int64) * (n + 1)])
    key.sort()
    all_rows = key // n
    all_cols = (key - all_rows * n).astype(np.int32)
    del key
    is_diag = all_cols == all_rows
    vals = rng.standard_normal(total + n)
    vals[is_diag] = 0.0
    diag = 1.0 + np.bincount(all_rows, weights=np.abs(vals), minlength=n)
    vals[is_diag] = diag
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(L + 1, out=row_ptr[1:])
    return row_ptr.astype(np.int32), all_cols, vals


def uniform_vector(n, seed, dtype=np.float64):
    """x ~ U[-1, 1] from the seed."""
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n).astype(dtype)


def csr_to_dense(row_ptr, cols, vals, ncols):
    n = len(row_ptr) - 1
    A = np.zeros((n, ncols))
    for r in range(n):
        for k in range(row_ptr[r], row_ptr[r + 1]):
            if cols[k] >= 0:
                A[r, cols[k]] += vals[k]
    return A


def dense_to_csr(A, keep_zeros_mask=None):
    """CSR of the nonzeros of a dense matrix (optionally keeping explicit zeros)."""
    mask = A != 0 if keep_zeros_mask is None else (A != 0) | keep_zeros_mask
    counts = mask.sum(axis=1)
    row_ptr = np.zeros(A.shape[0] + 1, dtype=np.int32)
    np.cumsum(counts, out=row_ptr[1:])
    r, c = np.nonzero(mask)
    return row_ptr, c.astype(np.int32), A[r, c].astype(np.float64)


def algorithmic_bytes(n_rows, n_cols, nnz, value_bytes, index_bytes=4, beta_nonzero=False, rhs=1):
    """CSR-compulsory bytes of one SpMV: values + column indices + row pointers + x read
    once + y written (+ y read when beta != 0), per right-hand side for the vectors."""
    b = nnz * (value_bytes + index_bytes) + (n_rows + 1) * index_bytes
    b += rhs * (n_cols * value_bytes + n_rows * value_bytes)
    if beta_nonzero:
        b += rhs * n_rows * value_bytes
    return b
