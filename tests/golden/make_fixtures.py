"""Writes the golden fixtures under tests/golden/ from TNL's own Host path.

The values come from oracle/_ref/libtnlref.so, which oracle/Makefile compiles from the
reference headers where they lie (no reference source is copied).  Run in the container
that holds the reference tree:

    make -C oracle && python tests/golden/make_fixtures.py

Outputs (numpy .npz, data only):
  segments.npz  per case: CSR input, x, and for each segments kind the metadata arrays,
                the converted storage (values, columns) and y = A x (column-major
                organization, TNL's GPU default; metadata also for row-major)
  adaptive.npz  CSR-Adaptive block descriptors per case and value size
  tnlfile.npz   TNL's binary SparseMatrix::save of small matrices in every matrix type, both
                organizations and both Real types (file bytes)
  solvers.npz   residual histories, iteration counts, convergence flags and solutions of
                CG / BiCGStab / GMRES (all variants, with and without Jacobi) at one thread, and
                the iteration counts of the same runs at 2, 3, 4 and 8 threads
  history.npz   TNL's residual-history-file text for some of those solves
  SHA256SUMS    checksums of every .npz above (tests/conftest.py checks them at session start)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tnl-dev_amd"))

import pyoracle as O  # noqa: E402
import workloads as W  # noqa: E402


def cases():
    """Small matrices covering the edge cases of every format."""
    out = {}
    out["lap2d_16"] = W.laplacian_2d(16)
    out["lap3d_10"] = W.laplacian_3d(10)  # n = 1000: partial last strip / slice
    out["irr_3000"] = W.irregular(3000, mean=16, sigma=1.2, max_len=1024, seed=7)
    out["convdiff_8"] = W.convection_diffusion_3d(8, 10.0)
    # explicit zeros, empty rows and a ragged tail of 33 rows (one-row last strip)
    rng = np.random.default_rng(11)
    rows = []
    for r in range(33):
        L = int(rng.integers(0, 40)) if r % 5 else 0
        cs = np.sort(rng.choice(64, size=L, replace=False))
        vs = rng.standard_normal(L)
        vs[rng.random(L) < 0.2] = 0.0  # explicit zeros are dropped by the conversion
        rows.append((cs, vs))
    rp = np.zeros(34, np.int32)
    rp[1:] = np.cumsum([len(c) for c, _ in rows])
    out["ragged_zeros_33x64"] = (rp, np.concatenate([c for c, _ in rows]).astype(np.int32),
                                 np.concatenate([v for _, v in rows]).astype(np.float64))
    return out


def mtx_cases():
    """Matrix Market texts covering what MatrixReader accepts and rejects."""
    rng = np.random.default_rng(21)
    cases = {
        "general": "%%MatrixMarket matrix coordinate real general\n% comment\n%\n4 5 6\n1 1 1.5\n4 5 -2e-3\n"
                   "2 3 7\n1 4 0.25\n3 2 1e10\n4 1 -0.125\n",
        "unsorted_duplicates_zeros": "%%MatrixMarket matrix coordinate real general\n3 4 7\n2 3 1\n2 1 2\n2 3 5\n"
                                     "1 1 0\n3 3 4\n3 3 0\n2 2 -1\n",
        "err_row_over_capacity": "%%MatrixMarket matrix coordinate real general\n3 3 4\n2 3 1\n2 1 2\n2 3 5\n"
                                 "2 2 -1\n",
        "symmetric": "%%MatrixMarket matrix coordinate real symmetric\n% lower triangle\n4 4 6\n1 1 4\n2 1 -1\n"
                     "2 2 4\n3 2 -1\n4 1 0.5\n4 4 4\n",
        "coordinates_keyword_blank_lines": "%%MatrixMarket matrix coordinates real general\n2 2 2\n\n1 2 3.0\n\n"
                                           "2 1 -3.0\n",
        "extra_spaces": "%%MatrixMarket  matrix coordinate  real general\n  3   2   2\n 1  2   8.5\n3 1    -4\n",
        "integer_tokens_truncate": "%%MatrixMarket matrix coordinate real general\n3 3 2\n2.7 1.2 5\n3 3 1e2\n",
        "empty_rows": "%%MatrixMarket matrix coordinate real general\n5 5 2\n2 2 1\n5 1 2\n",
        "err_header": "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",
        "err_complex": "%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n",
        "err_skew": "%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n2 1 1\n",
        "err_not_mm": "MatrixMarket matrix coordinate real general\n1 1 1\n1 1 1\n",
        "err_dims": "%%MatrixMarket matrix coordinate real general\n2 2\n1 1 1\n",
        "err_tokens": "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1\n",
        "err_tabs": "%%MatrixMarket matrix coordinate real general\n2 2 1\n1\t1\t1.0\n",
        "err_out_of_range": "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n",
        "err_zero_dims": "%%MatrixMarket matrix coordinate real general\n0 2 0\n",
    }
    # a larger random general matrix and a symmetric Laplacian, written with full precision
    n = 300
    lines = []
    for _ in range(2000):
        r, c = rng.integers(1, n + 1, 2)
        lines.append(f"{r} {c} {rng.standard_normal():.17g}")
    cases["random_300"] = (f"%%MatrixMarket matrix coordinate real general\n{n} {n} {len(lines)}\n" +
                           "\n".join(lines) + "\n")
    rp, c, v = W.laplacian_2d(12)
    low = [(i + 1, c[k] + 1, v[k]) for i in range(len(rp) - 1) for k in range(rp[i], rp[i + 1]) if c[k] <= i]
    m = len(rp) - 1
    cases["laplacian_12_symmetric"] = (f"%%MatrixMarket matrix coordinate real symmetric\n{m} {m} {len(low)}\n" +
                                       "".join(f"{a} {b} {val:.17g}\n" for a, b, val in low))
    return cases


def write_mtx_fixtures():
    import tempfile
    out = {}
    names = sorted(mtx_cases())
    out["names"] = np.array(names)
    with tempfile.TemporaryDirectory() as tmp:
        for name in names:
            text = mtx_cases()[name]
            out[f"{name}/text"] = np.array(text)
            path = os.path.join(tmp, name + ".mtx")
            with open(path, "w") as f:
                f.write(text)
            try:
                rows, cols, rp, ci, v = O.ref_read_mtx(path)
            except ValueError as e:
                out[f"{name}/error"] = np.array(str(e))
                continue
            out[f"{name}/dims"] = np.array([rows, cols], np.int32)
            out[f"{name}/row_ptr"] = rp
            out[f"{name}/cols"] = ci
            out[f"{name}/vals"] = v
            wpath = os.path.join(tmp, name + ".out.mtx")
            O.ref_write_mtx(wpath, rows, cols, rp, ci, v)
            with open(wpath) as f:
                out[f"{name}/written"] = np.array(f.read())
            # writeMtx of the device matrix readMtx builds from the file (zeros dropped twice over)
            O.ref_write_mtx_device(wpath, rows, cols, rp, ci, v)
            with open(wpath) as f:
                out[f"{name}/written_device"] = np.array(f.read())
    np.savez_compressed(os.path.join(HERE, "mtx.npz"), **out)
    print("mtx.npz", os.path.getsize(os.path.join(HERE, "mtx.npz")))


def main(only_solvers=False):
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libtnlref.so is missing: run make -C oracle in the reference container")
    if only_solvers:
        return write_solver_fixtures()
    seg = {}
    for name, (rp, c, v) in cases().items():
        n = len(rp) - 1
        ncols = 64 if name.startswith("ragged") else n
        x = W.uniform_vector(ncols, 1234)
        seg[f"{name}/row_ptr"] = rp
        seg[f"{name}/cols"] = c
        seg[f"{name}/vals"] = v
        seg[f"{name}/x"] = x
        seg[f"{name}/ncols"] = np.array([ncols], np.int32)
        caps = np.diff(rp).astype(np.int32)
        for kind, k in O.KINDS.items():
            for cm in (True, False):
                s = O.Segments(k, caps, cm, backend="ref")
                org = "col" if cm else "row"
                for nm, arr in s.metadata().items():
                    seg[f"{name}/{kind}/{org}/{nm}"] = arr
                seg[f"{name}/{kind}/{org}/storage"] = np.array([s.storage_size()], np.int64)
            vals, cols, y = O.ref_convert_spmv(k, rp, c, v, ncols, x, column_major=True)
            if kind == "BiEllpack":
                # TNL allocates 32 x groupPointers[last]; everything past groupPointers[last]
                # is padding, so only the used part is kept
                used = int(seg[f"{name}/{kind}/col/groupPointers"][-1])
                assert np.all(cols[used:] == -1) and np.all(vals[used:] == 0.0)
                vals, cols = vals[:used], cols[:used]
            if len(vals) <= 300000:  # padding-heavy Ellpack storage of the irregular case is left out
                seg[f"{name}/{kind}/col/values"] = vals
                seg[f"{name}/{kind}/col/columns"] = cols
            seg[f"{name}/{kind}/col/y"] = y
    np.savez_compressed(os.path.join(HERE, "segments.npz"), **seg)

    ad = {}
    for name, (rp, c, v) in cases().items():
        for vs in (4, 8):
            ad[f"{name}/{vs}"] = O.ref_csr_adaptive_blocks(rp, vs)
    np.savez_compressed(os.path.join(HERE, "adaptive.npz"), **ad)
    write_solver_fixtures()


def write_solver_fixtures():
    sol = {}
    systems = {
        "lap2d_24": W.laplacian_2d(24),
        "irr_2000": W.irregular(2000, mean=16, sigma=1.2, max_len=512, seed=3),
        "convdiff_8_pe10": W.convection_diffusion_3d(8, 10.0),
    }
    runs = [("lap2d_24", "cg", {}), ("lap2d_24", "cg", {"preconditioner": 1}),
            ("lap2d_24", "bicgstab", {}), ("lap2d_24", "bicgstab", {"preconditioner": 1}),
            ("lap2d_24", "bicgstab", {"bicgstab_exact_residue": 1})]
    runs += [("lap2d_24", "gmres", {"gmres_variant": var}) for var in range(5)]
    runs += [("lap2d_24", "gmres", {"gmres_variant": 4, "preconditioner": 1, "restarting_min": 5,
                                    "restarting_max": 15})]
    runs += [("irr_2000", "bicgstab", {"preconditioner": 1, "convergence_residue": 1e-10})]
    runs += [("convdiff_8_pe10", "gmres", {"gmres_variant": var, "preconditioner": 1, "restarting_min": 30,
                                           "restarting_max": 30, "convergence_residue": 1e-10})
             for var in (3, 4)]
    runs += [("lap2d_24", "tfqmr", {}), ("lap2d_24", "tfqmr", {"preconditioner": 1}),
             ("convdiff_8_pe10", "tfqmr", {"preconditioner": 1, "convergence_residue": 1e-10}),
             ("irr_2000", "tfqmr", {"preconditioner": 1, "convergence_residue": 1e-10, "min_iterations": 3})]
    # ILU(0)
    runs += [("lap2d_24", "cg", {"preconditioner": 2}), ("lap2d_24", "bicgstab", {"preconditioner": 2}),
             ("lap2d_24", "gmres", {"gmres_variant": 4, "preconditioner": 2}),
             ("convdiff_8_pe10", "gmres", {"gmres_variant": 3, "preconditioner": 2, "restarting_min": 30,
                                           "restarting_max": 30, "convergence_residue": 1e-10}),
             ("convdiff_8_pe10", "tfqmr", {"preconditioner": 2, "convergence_residue": 1e-10}),
             ("irr_2000", "bicgstab", {"preconditioner": 2, "convergence_residue": 1e-10})]
    # BiCGStab(ell)
    runs += [("lap2d_24", "bicgstabl", {"bicgstab_ell": 1}), ("lap2d_24", "bicgstabl", {"bicgstab_ell": 2}),
             ("lap2d_24", "bicgstabl", {"bicgstab_ell": 4, "preconditioner": 1}),
             ("convdiff_8_pe10", "bicgstabl", {"bicgstab_ell": 2, "preconditioner": 1, "convergence_residue": 1e-10}),
             ("convdiff_8_pe10", "bicgstabl", {"bicgstab_ell": 3, "preconditioner": 2, "bicgstab_exact_residue": 1,
                                               "convergence_residue": 1e-10}),
             ("irr_2000", "bicgstabl", {"bicgstab_ell": 2, "preconditioner": 1, "convergence_residue": 1e-10})]
    for name, (rp, c, v) in systems.items():
        sol[f"{name}/row_ptr"] = rp
        sol[f"{name}/cols"] = c
        sol[f"{name}/vals"] = v
        sol[f"{name}/b"] = W.uniform_vector(len(rp) - 1, 42)
    for i, (name, method, kw) in enumerate(runs):
        rp, c, v = systems[name]
        kw = dict(kw)
        kw.setdefault("convergence_residue", 1e-8)
        kw.setdefault("max_iterations", 5000)
        r = O.solve(method, rp, c, v, sol[f"{name}/b"], backend="ref", threads=1, **kw)
        key = f"run{i:02d}"
        sol[f"{key}/system"] = np.array(name)
        sol[f"{key}/method"] = np.array(method)
        sol[f"{key}/params"] = np.array(repr(sorted(kw.items())))
        sol[f"{key}/history"] = r["history"]
        sol[f"{key}/iterations"] = np.array([r["iterations"]])
        # TNL's own count with 2, 3, 4 and 8 OpenMP threads (another reduction order): how far
        # the count of this run moves with the rounding of its dot products alone
        sol[f"{key}/iterations_threads"] = np.array(
            [r["iterations"]] + [O.solve(method, rp, c, v, sol[f"{name}/b"], backend="ref", threads=t, **kw)["iterations"]
                                 for t in (2, 3, 4, 8)])
        sol[f"{key}/converged"] = np.array([r["converged"]])
        sol[f"{key}/x"] = r["x"]
    np.savez_compressed(os.path.join(HERE, "solvers.npz"), **sol)
    print("solvers.npz", os.path.getsize(os.path.join(HERE, "solvers.npz")))


HISTORY_RUNS = [("lap2d_24", "cg", {}), ("lap2d_24", "bicgstab", {"preconditioner": 1}),
                ("lap2d_24", "gmres", {"gmres_variant": 4}), ("lap2d_24", "gmres", {"gmres_variant": 1}),
                ("convdiff_8_pe10", "tfqmr", {"preconditioner": 1, "convergence_residue": 1e-10}),
                ("lap2d_24", "bicgstabl", {"bicgstab_ell": 2})]


def write_history_fixtures():
    """TNL's residual-history-file text (IterativeSolver.hpp:57-61, 189-200) for solves of the
    systems held in solvers.npz."""
    import tempfile
    sol = np.load(os.path.join(HERE, "solvers.npz"))
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for i, (name, method, kw) in enumerate(HISTORY_RUNS):
            kw = dict(kw)
            kw.setdefault("convergence_residue", 1e-8)
            kw.setdefault("max_iterations", 5000)
            path = os.path.join(tmp, f"h{i}.txt")
            O.solve(method, sol[f"{name}/row_ptr"], sol[f"{name}/cols"], sol[f"{name}/vals"], sol[f"{name}/b"],
                    backend="ref", threads=1, history_file=path, **kw)
            key = f"hist{i:02d}"
            out[f"{key}/system"] = np.array(name)
            out[f"{key}/method"] = np.array(method)
            out[f"{key}/params"] = np.array(repr(sorted(kw.items())))
            with open(path) as f:
                out[f"{key}/text"] = np.array(f.read())
    np.savez_compressed(os.path.join(HERE, "history.npz"), **out)
    print("history.npz", os.path.getsize(os.path.join(HERE, "history.npz")))


def write_checksums():
    """SHA256SUMS in sha256sum's format, one line per fixture file."""
    import hashlib
    lines = []
    for name in sorted(f for f in os.listdir(HERE) if f.endswith(".npz")):
        with open(os.path.join(HERE, name), "rb") as f:
            lines.append(f"{hashlib.sha256(f.read()).hexdigest()}  {name}\n")
    with open(os.path.join(HERE, "SHA256SUMS"), "w") as f:
        f.writelines(lines)


def tnlfile_cases():
    """Matrices for TNL binary file fixtures: no explicit zeros, so that every conversion keeps
    the same entries."""
    rng = np.random.default_rng(31)

    def random_rows(n, ncols, lengths):
        rp = np.zeros(n + 1, np.int32)
        rp[1:] = np.cumsum(lengths)
        ci = np.concatenate([np.sort(rng.choice(ncols, size=L, replace=False)) for L in lengths] or [[]])
        # eighths keep the fixtures compressible; exact in float as in double
        v = rng.integers(1, 17, rp[-1]) / 8.0 * rng.choice([-1.0, 1.0], rp[-1])
        return rp, ci.astype(np.int32), v, ncols

    out = {}
    out["tiny_5x5"] = (np.array([0, 2, 3, 6, 6, 7], np.int32), np.array([0, 3, 1, 0, 2, 4, 3], np.int32),
                       np.arange(1, 8, dtype=np.float64), 5)
    lengths = rng.integers(0, 26, 70)
    lengths[::9] = 0
    out["ragged_70x60"] = random_rows(70, 60, lengths)
    out["many_slices_600x400"] = random_rows(600, 400, rng.integers(1, 21, 600))
    out["empty_rows_40x40"] = random_rows(40, 40, np.zeros(40, np.int64))
    return out


def write_tnlfile_fixtures():
    """TNL's own SparseMatrix::save of every case, matrix type, organization and Real type."""
    import tempfile
    out = {}
    cases = tnlfile_cases()
    out["names"] = np.array(sorted(cases))
    with tempfile.TemporaryDirectory() as tmp:
        for name in sorted(cases):
            rp, ci, v, ncols = cases[name]
            out[f"{name}/row_ptr"] = rp
            out[f"{name}/cols"] = ci
            out[f"{name}/vals"] = v
            out[f"{name}/ncols"] = np.array([ncols], np.int32)
            for variant in O.TNL_FILE_VARIANTS:
                # a CSR file does not depend on the organization
                orgs = (("cm", True),) if variant.startswith("CSR") else (("cm", True), ("rm", False))
                for org, cm in orgs:
                    for dt, single in (("f64", False), ("f32", True)):
                        path = os.path.join(tmp, f"{name}_{variant}_{org}_{dt}.tnl")
                        O.ref_save_tnl(variant, rp, ci, v, ncols, path, column_major=cm, single=single)
                        with open(path, "rb") as f:
                            out[f"{name}/{variant}/{org}/{dt}"] = np.frombuffer(f.read(), np.uint8)
    np.savez_compressed(os.path.join(HERE, "tnlfile.npz"), **out)
    print("tnlfile.npz", os.path.getsize(os.path.join(HERE, "tnlfile.npz")))


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "solvers":
        main(only_solvers=True)
    elif len(sys.argv) > 1 and sys.argv[1] == "tnlfile":
        if not O.ref_available():
            raise SystemExit("oracle/_ref/libtnlref.so is missing: run make -C oracle in the reference container")
        write_tnlfile_fixtures()
    elif len(sys.argv) > 1 and sys.argv[1] == "mtx":
        if not O.ref_available():
            raise SystemExit("oracle/_ref/libtnlref.so is missing: run make -C oracle in the reference container")
        write_mtx_fixtures()
    else:
        main()
        write_mtx_fixtures()
        write_tnlfile_fixtures()
    write_checksums()
