"""ctypes bindings for the CPU oracle (liboracle.so) and, when built, for TNL's own Host
path compiled from the reference headers (_ref/libtnlref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, always as the checker, never as the thing measured or shipped.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libtnlref.so")

KINDS = {"CSR": 0, "Ellpack": 1, "SlicedEllpack": 2, "ChunkedEllpack": 3, "BiEllpack": 4}
METHODS = {"cg": 0, "bicgstab": 1, "gmres": 2, "tfqmr": 3, "bicgstabl": 4}
GMRES_VARIANTS = {"CGS": 0, "CGSR": 1, "MGS": 2, "MGSR": 3, "CWY": 4}

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")

METADATA_NAMES = {
    0: ["offsets"],
    1: ["ellpack"],
    2: ["sliceOffsets", "sliceSegmentSizes", "sliced"],
    3: ["rowToChunkMapping", "rowToSliceMapping", "chunksToSegmentsMapping", "rowPointers", "slices", "chunked"],
    4: ["rowPermArray", "groupPointers", "biellpack"],
}


class SolverParams(C.Structure):
    _fields_ = [
        ("max_iterations", C.c_int32),
        ("min_iterations", C.c_int32),
        ("convergence_residue", C.c_double),
        ("divergence_residue", C.c_double),
        ("gmres_variant", C.c_int32),
        ("restarting_min", C.c_int32),
        ("restarting_max", C.c_int32),
        ("restarting_step_min", C.c_int32),
        ("restarting_step_max", C.c_int32),
        ("bicgstab_exact_residue", C.c_int32),
        ("preconditioner", C.c_int32),
        ("bicgstab_ell", C.c_int32),
    ]


def make_params(lib_default, **kw):
    p = SolverParams()
    lib_default(C.byref(p))
    for k, v in kw.items():
        if k == "gmres_variant" and isinstance(v, str):
            v = GMRES_VARIANTS[v]
        if k == "preconditioner" and isinstance(v, str):
            v = {"none": 0, "diagonal": 1, "ilu0": 2}[v]
        setattr(p, k, v)
    return p


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle not built: {ORACLE_SO} (run __graft_entry__.build())")
        L = C.CDLL(ORACLE_SO)
        L.orc_segments_create.restype = C.c_void_p
        L.orc_segments_create.argtypes = [C.c_int32, C.c_int32, _i32p, C.c_int32]
        L.orc_segments_create_sliced.restype = C.c_void_p
        L.orc_segments_create_sliced.argtypes = [C.c_int32, _i32p, C.c_int32, C.c_int32]
        L.orc_segments_destroy.argtypes = [C.c_void_p]
        L.orc_segments_storage_size.restype = C.c_int64
        L.orc_segments_storage_size.argtypes = [C.c_void_p]
        L.orc_segments_segment_size.restype = C.c_int32
        L.orc_segments_segment_size.argtypes = [C.c_void_p, C.c_int32]
        L.orc_segments_global_index.restype = C.c_int64
        L.orc_segments_global_index.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
        L.orc_segments_get.restype = C.c_int64
        L.orc_segments_get.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]
        L.orc_convert_from_csr.argtypes = [C.c_void_p, _i32p, _i32p, _f64p, _f64p, _i32p]
        L.orc_spmv_f64.argtypes = [C.c_void_p, _f64p, _i32p, _f64p, _f64p, C.c_double, C.c_double, C.c_int32, C.c_int32]
        L.orc_spmv_f32.argtypes = [C.c_void_p, _f32p, _i32p, _f32p, _f32p, C.c_float, C.c_float, C.c_int32, C.c_int32]
        L.orc_csr_adaptive_blocks.restype = C.c_int64
        L.orc_csr_adaptive_blocks.argtypes = [C.c_int32, _i32p] + [C.c_int32] * 5 + [C.c_void_p]
        L.orc_reduce_rows.restype = None
        L.orc_reduce_rows.argtypes = [C.c_void_p, _f64p, _i32p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                      C.c_double, _f64p]
        L.orc_dot_f64.restype = C.c_double
        L.orc_dot_f64.argtypes = [C.c_int64, _f64p, _f64p]
        L.orc_dot_f32.restype = C.c_float
        L.orc_dot_f32.argtypes = [C.c_int64, _f32p, _f32p]
        L.orc_solver_params_default.argtypes = [C.c_void_p]
        L.orc_ilu0.restype = C.c_int32
        L.orc_ilu0.argtypes = [C.c_int32, _i32p, _i32p, _f64p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)] + \
            [C.c_void_p] * 8
        L.orc_symmetric_spmv.restype = None
        L.orc_symmetric_spmv.argtypes = [C.c_int32, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_double, C.c_double,
                                         C.c_int32, C.c_int32]
        L.orc_binary_spmv.restype = None
        L.orc_binary_spmv.argtypes = [C.c_int32, _i32p, _i32p, _i32p, _i32p, C.c_int32, C.c_int32, C.c_int32,
                                      C.c_int32]
        L.orc_solve.restype = C.c_int32
        L.orc_solve.argtypes = [C.c_int32, C.c_int32, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_void_p,
                                C.POINTER(C.c_int32), C.POINTER(C.c_double), C.c_void_p, C.c_int32,
                                C.POINTER(C.c_int32)]
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"TNL reference build not present: {REF_SO}")
        L = C.CDLL(REF_SO)
        L.ref_segments_create.restype = C.c_void_p
        L.ref_segments_create.argtypes = [C.c_int, C.c_int, _i32p, C.c_int]
        L.ref_sliced_create.restype = C.c_void_p
        L.ref_sliced_create.argtypes = [C.c_int, _i32p, C.c_int, C.c_int]
        L.ref_segments_destroy.argtypes = [C.c_void_p]
        L.ref_segments_storage_size.restype = C.c_longlong
        L.ref_segments_storage_size.argtypes = [C.c_void_p]
        L.ref_segments_get.restype = C.c_longlong
        L.ref_segments_get.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]
        L.ref_convert_spmv.restype = C.c_int
        L.ref_convert_spmv.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p,
                                       _f64p, _i32p, C.c_longlong]
        L.ref_row_reductions.restype = C.c_int
        L.ref_row_reductions.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _i32p, _i32p, _f64p, C.c_int, C.c_int,
                                         C.c_double, _f64p, _i32p, _i32p, _i32p]
        L.ref_csr_adaptive_blocks.restype = C.c_longlong
        L.ref_csr_adaptive_blocks.argtypes = [C.c_int, _i32p, C.c_int, C.c_void_p]
        L.ref_time_spmv_csr.restype = C.c_double
        L.ref_time_spmv_csr.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_int, C.c_int]
        L.ref_solve.restype = C.c_int
        L.ref_solve.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_void_p,
                                C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_void_p, C.c_int,
                                C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_double)]
        L.ref_read_mtx.restype = C.c_int
        L.ref_read_mtx.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_longlong),
                                   C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_int]
        L.ref_keep_history_file.restype = None
        L.ref_keep_history_file.argtypes = [C.c_char_p]
        L.ref_write_mtx_device.restype = C.c_int
        L.ref_write_mtx_device.argtypes = [C.c_char_p, C.c_int, C.c_int, _i32p, _i32p, _f64p]
        L.ref_write_mtx.restype = C.c_int
        L.ref_write_mtx.argtypes = [C.c_char_p, C.c_int, C.c_int, _i32p, _i32p, _f64p]
        L.ref_save_tnl.restype = C.c_int
        L.ref_save_tnl.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _i32p, _i32p, _f64p, C.c_char_p,
                                   C.c_char_p, C.c_int]
        L.ref_load_tnl.restype = C.c_int
        L.ref_load_tnl.argtypes = [C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_void_p, C.c_void_p,
                                   C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_char_p, C.c_int]
        L.ref_symmetric_spmv.restype = C.c_int
        L.ref_symmetric_spmv.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_double, C.c_double,
                                         C.c_int, C.c_int]
        L.ref_binary_spmv.restype = C.c_int
        L.ref_binary_spmv.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _i32p, _i32p, C.c_int, C.c_int, C.c_int,
                                      C.c_int]
        L.ref_ilu0.restype = C.c_int
        L.ref_ilu0.argtypes = [C.c_int, _i32p, _i32p, _f64p, C.POINTER(C.c_int), C.POINTER(C.c_int)] + \
            [C.c_void_p] * 8
        _ref = L
    return _ref


def ref_read_mtx(path):
    """TNL's MatrixReader::readMtx into a Host CSR matrix: (rows, cols, row_ptr, cols, vals), or
    raises ValueError with TNL's message."""
    L = ref()
    rows, cols, st = C.c_int(), C.c_int(), C.c_longlong()
    err = C.create_string_buffer(1024)
    if L.ref_read_mtx(str(path).encode(), C.byref(rows), C.byref(cols), C.byref(st), None, None, None, err, 1024):
        raise ValueError(err.value.decode(errors="replace"))
    rp = np.zeros(rows.value + 1, np.int32)
    ci = np.zeros(max(st.value, 1), np.int32)
    v = np.zeros(max(st.value, 1), np.float64)
    L.ref_read_mtx(str(path).encode(), C.byref(rows), C.byref(cols), C.byref(st), rp.ctypes.data, ci.ctypes.data,
                   v.ctypes.data, err, 1024)
    return rows.value, cols.value, rp, ci[:st.value], v[:st.value]


def ref_write_mtx_device(path, rows, cols, row_ptr, col_idx, vals):
    """writeMtx of a device matrix holding TNL's host read result (two operator= conversions)."""
    ref().ref_write_mtx_device(str(path).encode(), int(rows), int(cols), np.ascontiguousarray(row_ptr, np.int32),
                               np.ascontiguousarray(col_idx, np.int32), np.ascontiguousarray(vals, np.float64))


def ref_write_mtx(path, rows, cols, row_ptr, col_idx, vals):
    """TNL's MatrixWriter::writeMtx of a Host CSR matrix."""
    ref().ref_write_mtx(str(path).encode(), int(rows), int(cols), np.ascontiguousarray(row_ptr, np.int32),
                        np.ascontiguousarray(col_idx, np.int32), np.ascontiguousarray(vals, np.float64))


# TNL matrix types of ref_save_tnl / ref_load_tnl: CSR with each kernel, then the Ellpack family
TNL_FILE_VARIANTS = ["CSRScalar", "CSRVector", "CSRHybrid", "CSRLight", "CSRAdaptive", "Ellpack", "SlicedEllpack",
                     "ChunkedEllpack", "BiEllpack"]


def ref_save_tnl(variant, row_ptr, col_idx, vals, ncols, path, column_major=True, single=False):
    """TNL's SparseMatrix< double or float, Host, int, General, variant >::operator=( CSR ) then
    save( path ); column_major selects the organization of the Ellpack family."""
    v = TNL_FILE_VARIANTS.index(variant) if isinstance(variant, str) else variant
    rp = np.ascontiguousarray(row_ptr, np.int32)
    err = C.create_string_buffer(1024)
    if ref().ref_save_tnl(v, 1 if column_major else 0, 1 if single else 0, len(rp) - 1, int(ncols), rp,
                          np.ascontiguousarray(col_idx, np.int32), np.ascontiguousarray(vals, np.float64),
                          str(path).encode(), err, 1024):
        raise ValueError(err.value.decode(errors="replace"))


def ref_load_tnl(variant, path, x=None, column_major=True, single=False):
    """TNL's SparseMatrix::load( path ) into the given type, then y = A x when x is given:
    (rows, columns, y); raises ValueError with TNL's message."""
    v = TNL_FILE_VARIANTS.index(variant) if isinstance(variant, str) else variant
    rows, cols = C.c_int(), C.c_int()
    err = C.create_string_buffer(1024)
    xs = None if x is None else np.ascontiguousarray(x, np.float64)
    y = None
    if xs is not None:
        # the row count is known only after loading: load once without the product
        if ref().ref_load_tnl(v, 1 if column_major else 0, 1 if single else 0, str(path).encode(), None, None,
                              C.byref(rows), C.byref(cols), err, 1024):
            raise ValueError(err.value.decode(errors="replace"))
        y = np.zeros(max(rows.value, 1), np.float64)
    if ref().ref_load_tnl(v, 1 if column_major else 0, 1 if single else 0, str(path).encode(),
                          None if xs is None else xs.ctypes.data, None if y is None else y.ctypes.data,
                          C.byref(rows), C.byref(cols), err, 1024):
        raise ValueError(err.value.decode(errors="replace"))
    return rows.value, cols.value, None if y is None else y[:rows.value]


# ------------------------------------------------------------------ segments


class Segments:
    """Oracle segments of one kind, built from row capacities."""

    def __init__(self, kind, capacities, column_major=True, backend="oracle", slice_size=32):
        self.kind = KINDS[kind] if isinstance(kind, str) else kind
        self.caps = np.ascontiguousarray(capacities, dtype=np.int32)
        self.n = len(self.caps)
        self.column_major = column_major
        self.backend = backend
        sliced = slice_size != 32 and self.kind == KINDS["SlicedEllpack"]
        if backend == "oracle":
            self._L = lib()
            if sliced:
                self._h = self._L.orc_segments_create_sliced(self.n, self.caps, 1 if column_major else 0, slice_size)
            else:
                self._h = self._L.orc_segments_create(self.kind, self.n, self.caps, 1 if column_major else 0)
        else:
            self._L = ref()
            if sliced:
                self._h = self._L.ref_sliced_create(self.n, self.caps, 1 if column_major else 0, slice_size)
            else:
                self._h = self._L.ref_segments_create(self.kind, self.n, self.caps, 1 if column_major else 0)
        if not self._h:
            raise ValueError("bad segments kind")

    def __del__(self):
        try:
            if self._h:
                if self.backend == "oracle":
                    self._L.orc_segments_destroy(self._h)
                else:
                    self._L.ref_segments_destroy(self._h)
        except Exception:
            pass

    def storage_size(self):
        if self.backend == "oracle":
            return int(self._L.orc_segments_storage_size(self._h))
        return int(self._L.ref_segments_storage_size(self._h))

    def get(self, name):
        f = self._L.orc_segments_get if self.backend == "oracle" else self._L.ref_segments_get
        cnt = f(self._h, name.encode(), None)
        if cnt < 0:
            raise KeyError(name)
        out = np.zeros(max(cnt, 1), dtype=np.int32)
        f(self._h, name.encode(), out.ctypes.data)
        return out[:cnt]

    def metadata(self):
        return {nm: self.get(nm) for nm in METADATA_NAMES[self.kind]}

    def get_segment_size(self, seg):
        assert self.backend == "oracle"
        return int(self._L.orc_segments_segment_size(self._h, seg))

    def global_index(self, seg, local):
        assert self.backend == "oracle"
        return int(self._L.orc_segments_global_index(self._h, seg, local))

    def convert(self, row_ptr, cols, vals):
        """SparseMatrix::operator=( CSR ) storage arrays (oracle only)."""
        st = self.storage_size()
        ov = np.zeros(max(st, 1), np.float64)
        oc = np.zeros(max(st, 1), np.int32)
        self._L.orc_convert_from_csr(self._h, np.ascontiguousarray(row_ptr, np.int32),
                                     np.ascontiguousarray(cols, np.int32), np.ascontiguousarray(vals, np.float64), ov, oc)
        return ov[:st], oc[:st]

    def spmv(self, vals, cols, x, y=None, alpha=1.0, beta=0.0, first=0, last=0, dtype=np.float64):
        n = self.n
        if y is None:
            y = np.zeros(n, dtype)
        y = np.ascontiguousarray(y, dtype).copy()
        v = np.ascontiguousarray(vals, dtype)
        xx = np.ascontiguousarray(x, dtype)
        c = np.ascontiguousarray(cols, np.int32)
        if len(v) == 0:
            v = np.zeros(1, dtype)
            c = np.full(1, -1, np.int32)
        f = self._L.orc_spmv_f64 if dtype == np.float64 else self._L.orc_spmv_f32
        f(self._h, v, c, xx, y, alpha, beta, first, last)
        return y


FETCHES = {"value": 0, "abs": 1, "one": 2, "nonzero": 3, "column": 4}
REDUCES = {"sum": 0, "min": 1, "max": 2}


def reduce_rows(seg, vals, cols, fetch, reduce, identity, first=0, last=0):
    """SparseMatrixView::reduceRows over the oracle segments' storage (vals, cols in storage order)."""
    out = np.zeros(max(seg.n, 1), np.float64)
    v = np.ascontiguousarray(vals, np.float64)
    c = np.ascontiguousarray(cols, np.int32)
    if len(v) == 0:
        v = np.zeros(1, np.float64)
        c = np.full(1, -1, np.int32)
    seg._L.orc_reduce_rows(seg._h, v, c, int(first), int(last), FETCHES[fetch], REDUCES[reduce], float(identity), out)
    return out[:seg.n]


def ref_row_reductions(kind, row_ptr, cols, vals, ncols, fetch, reduce, identity, column_major=True):
    """TNL's reduceAllRows / getRowCapacities / getCompressedRowLengths / getRowCapacity on a Host
    matrix of `kind` assigned from the CSR (the CSR kind keeps the CSR storage as given)."""
    k = KINDS[kind] if isinstance(kind, str) else kind
    rp = np.ascontiguousarray(row_ptr, np.int32)
    n = len(rp) - 1
    c = np.ascontiguousarray(cols, np.int32)
    v = np.ascontiguousarray(vals, np.float64)
    if len(c) == 0:
        c = np.full(1, -1, np.int32)
        v = np.zeros(1, np.float64)
    out = np.zeros(max(n, 1), np.float64)
    caps = np.zeros(max(n, 1), np.int32)
    lens = np.zeros(max(n, 1), np.int32)
    sizes = np.zeros(max(n, 1), np.int32)
    rc = ref().ref_row_reductions(k, 1 if column_major else 0, n, int(ncols), rp, c, v, FETCHES[fetch],
                                  REDUCES[reduce], float(identity), out, caps, lens, sizes)
    if rc:
        raise RuntimeError(f"ref_row_reductions failed: {rc}")
    return out[:n], caps[:n], lens[:n], sizes[:n]


def ref_convert_spmv(kind, row_ptr, cols, vals, ncols, x, column_major=True, y=None):
    """TNL SparseMatrix<Host, kind> = CSR; returns (values, columns, y = A x)."""
    L = ref()
    k = KINDS[kind] if isinstance(kind, str) else kind
    rp = np.ascontiguousarray(row_ptr, np.int32)
    n = len(rp) - 1
    seg = Segments(k, np.diff(rp), column_major, backend="ref")
    st = seg.storage_size()
    ov = np.zeros(max(st, 1), np.float64)
    oc = np.zeros(max(st, 1), np.int32)
    yy = np.zeros(n, np.float64) if y is None else np.ascontiguousarray(y, np.float64).copy()
    c = np.ascontiguousarray(cols, np.int32)
    v = np.ascontiguousarray(vals, np.float64)
    if len(c) == 0:
        c = np.zeros(1, np.int32)
        v = np.zeros(1, np.float64)
    rc = L.ref_convert_spmv(k, 1 if column_major else 0, n, ncols, rp, c, v,
                            np.ascontiguousarray(x, np.float64), yy, ov, oc, max(st, 1))
    if rc != 0:
        raise RuntimeError(f"ref_convert_spmv failed: {rc}")
    return ov[:st], oc[:st], yy


def csr_adaptive_blocks(offsets, value_size, warp_size=32, block_size=256, streamed_shared_bytes=24576,
                        max_adaptive_elements=512):
    L = lib()
    off = np.ascontiguousarray(offsets, np.int32)
    n = len(off) - 1
    cnt = L.orc_csr_adaptive_blocks(n, off, value_size, warp_size, block_size, streamed_shared_bytes,
                                    max_adaptive_elements, None)
    out = np.zeros(max(6 * cnt, 1), np.int32)
    L.orc_csr_adaptive_blocks(n, off, value_size, warp_size, block_size, streamed_shared_bytes,
                              max_adaptive_elements, out.ctypes.data)
    return out[:6 * cnt].reshape(-1, 6)


def ref_csr_adaptive_blocks(offsets, value_size):
    L = ref()
    off = np.ascontiguousarray(offsets, np.int32)
    n = len(off) - 1
    log = int(np.log2(value_size))
    cnt = L.ref_csr_adaptive_blocks(n, off, log, None)
    out = np.zeros(max(6 * cnt, 1), np.int32)
    L.ref_csr_adaptive_blocks(n, off, log, out.ctypes.data)
    return out[:6 * cnt].reshape(-1, 6)


def dot(a, b):
    if a.dtype == np.float32:
        return float(lib().orc_dot_f32(len(a), np.ascontiguousarray(a), np.ascontiguousarray(b)))
    return float(lib().orc_dot_f64(len(a), np.ascontiguousarray(a, np.float64), np.ascontiguousarray(b, np.float64)))


# ------------------------------------------------------------------ solvers


def symmetric_spmv(row_ptr, cols, vals, x, y0=None, alpha=1.0, beta=0.0, first=0, last=0, backend="oracle"):
    """The SymmetricMatrix product of a lower-triangle CSR matrix: y = beta y0 (scaled only when
    beta != 1) plus alpha A x with A = L + L^T - D."""
    rp = np.ascontiguousarray(row_ptr, np.int32)
    n = len(rp) - 1
    y = np.zeros(n) if y0 is None else np.array(y0, np.float64)
    f = lib().orc_symmetric_spmv if backend == "oracle" else ref().ref_symmetric_spmv
    f(n, rp, np.ascontiguousarray(cols, np.int32), np.ascontiguousarray(vals, np.float64),
      np.ascontiguousarray(x, np.float64), y, float(alpha), float(beta), int(first), int(last))
    return y


def binary_spmv(row_ptr, cols, x, y0=None, alpha=1, beta=0, first=0, last=0, ncols=None, backend="oracle"):
    """Binary matrix product in int32: y = alpha * (sum of x over the stored columns) + beta * y0."""
    rp = np.ascontiguousarray(row_ptr, np.int32)
    n = len(rp) - 1
    xs = np.ascontiguousarray(x, np.int32)
    y = np.zeros(n, np.int32) if y0 is None else np.array(y0, np.int32)
    c = np.ascontiguousarray(cols, np.int32)
    if backend == "oracle":
        lib().orc_binary_spmv(n, rp, c, xs, y, int(alpha), int(beta), int(first), int(last))
    else:
        ref().ref_binary_spmv(n, len(xs) if ncols is None else ncols, rp, c, xs, y, int(alpha), int(beta),
                              int(first), int(last))
    return y


def ilu0(row_ptr, cols, vals, b=None, backend="oracle"):
    """ILU(0) factors ((Lrp, Lc, Lv), (Urp, Uc, Uv)) and (LU)^-1 b.  The oracle keeps U's rows in
    natural order; TNL stores them reversed (row N-1-i of its U is row i), so backend="ref"
    returns them reordered to natural order."""
    rp = np.ascontiguousarray(row_ptr, np.int32)
    c = np.ascontiguousarray(cols, np.int32)
    v = np.ascontiguousarray(vals, np.float64)
    n = len(rp) - 1
    f = lib().orc_ilu0 if backend == "oracle" else ref().ref_ilu0
    nl, nu = C.c_int32(), C.c_int32()
    f(n, rp, c, v, C.byref(nl), C.byref(nu), *([None] * 8))
    Lrp, Urp = np.zeros(n + 1, np.int32), np.zeros(n + 1, np.int32)
    Lc, Uc = np.zeros(max(nl.value, 1), np.int32), np.zeros(max(nu.value, 1), np.int32)
    Lv, Uv = np.zeros(max(nl.value, 1)), np.zeros(max(nu.value, 1))
    x = np.zeros(n)
    bb = None if b is None else np.ascontiguousarray(b, np.float64)
    f(n, rp, c, v, C.byref(nl), C.byref(nu), Lrp.ctypes.data, Lc.ctypes.data, Lv.ctypes.data, Urp.ctypes.data,
      Uc.ctypes.data, Uv.ctypes.data, None if bb is None else bb.ctypes.data, None if bb is None else x.ctypes.data)
    Lc, Lv, Uc, Uv = Lc[:nl.value], Lv[:nl.value], Uc[:nu.value], Uv[:nu.value]
    if backend != "oracle":
        # TNL's U row k holds matrix row n-1-k
        lens = np.diff(Urp)[::-1]
        starts = Urp[:-1][::-1]
        order = np.concatenate([np.arange(s, s + L) for s, L in zip(starts, lens)] or [np.zeros(0, np.int64)])
        Uc, Uv = Uc[order], Uv[order]
        Urp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    return (Lrp, Lc, Lv), (Urp, Uc, Uv), (None if b is None else x)


def solve(method, row_ptr, cols, vals, b, x0=None, backend="oracle", threads=1, history_cap=100000, history_file=None,
          **params):
    """Returns dict(x, converged, iterations, residue, history[, seconds]).  With backend "ref" and a
    history_file, TNL's residual-history-file of the solve is left at that path."""
    rp = np.ascontiguousarray(row_ptr, np.int32)
    n = len(rp) - 1
    c = np.ascontiguousarray(cols, np.int32)
    v = np.ascontiguousarray(vals, np.float64)
    bb = np.ascontiguousarray(b, np.float64)
    x = np.zeros(n) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
    m = METHODS[method] if isinstance(method, str) else method
    hist = np.zeros(history_cap, np.float64)
    hl = C.c_int32(0)
    res = C.c_double(0)
    its = C.c_int32(0)
    out = {}
    if backend == "oracle":
        L = lib()
        p = make_params(L.orc_solver_params_default, **params)
        rc = L.orc_solve(m, n, rp, c, v, bb, x, C.byref(p), C.byref(its), C.byref(res), hist.ctypes.data,
                         history_cap, C.byref(hl))
    else:
        L = ref()
        if history_file:
            L.ref_keep_history_file(str(history_file).encode())
        p = make_params(lib().orc_solver_params_default, **params)
        secs = C.c_double(0)
        its_i = C.c_int(0)
        hl_i = C.c_int(0)
        rc = L.ref_solve(m, n, rp, c, v, bb, x, C.byref(p), C.byref(its_i), C.byref(res), hist.ctypes.data,
                         history_cap, C.byref(hl_i), threads, C.byref(secs))
        its, hl = its_i, hl_i
        out["seconds"] = secs.value
    out.update(x=x, converged=rc, iterations=its.value, residue=res.value,
               history=hist[:min(hl.value, history_cap)].copy())
    return out


def ref_time_spmv_csr(row_ptr, cols, vals, ncols, x, loops=10, threads=0):
    """TNL Host CSR vectorProduct (OpenMP), median seconds per call over `loops` calls."""
    L = ref()
    rp = np.ascontiguousarray(row_ptr, np.int32)
    n = len(rp) - 1
    y = np.zeros(n, np.float64)
    t = L.ref_time_spmv_csr(n, ncols, rp, np.ascontiguousarray(cols, np.int32), np.ascontiguousarray(vals, np.float64),
                            np.ascontiguousarray(x, np.float64), y, loops, threads)
    return t, y
