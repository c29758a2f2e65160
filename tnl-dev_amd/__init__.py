"""Python mirror of TNL's SpMV + Krylov interface over the tnlhip C ABI (libtnlhip.so).

Names follow TNL (src/TNL/Matrices/SparseMatrix.h, src/TNL/Solvers/Linear/*.h):
SparseMatrix.vectorProduct, LinearSolver.setMatrix / setPreconditioner / solve, and the
IterativeSolver setters.  Vectors are torch tensors on the GPU (float64 for FP64 matrices,
float32 for FP32 / BF16 matrices); torch provides device memory and streams only.

The HIP library is required: importing this package raises if libtnlhip.so is missing,
and there is no CPU fallback.
"""
import ctypes as C
import os
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtnlhip.so")

FORMATS = {
    "CSRScalar": 0, "CSRVector": 1, "CSRLight": 2, "CSRHybrid": 3, "CSRAdaptive": 4,
    "Ellpack": 5, "SlicedEllpack": 6, "ChunkedEllpack": 7, "BiEllpack": 8,
}
FORMAT_NAMES = {v: k for k, v in FORMATS.items()}
DTYPES = {"f64": 0, "f32": 1, "bf16": 2, "binary": 3}
METHODS = {"cg": 0, "bicgstab": 1, "gmres": 2, "tfqmr": 3, "bicgstabl": 4}
GMRES_VARIANTS = {"CGS": 0, "CGSR": 1, "MGS": 2, "MGSR": 3, "CWY": 4}

# every function declared in include/tnlhip.h
EXPORTS = [
    "tnlhip_status_string", "tnlhip_last_error", "tnlhip_device_info",
    "tnlhip_matrix_create_csr", "tnlhip_matrix_convert", "tnlhip_matrix_destroy", "tnlhip_matrix_get_info",
    "tnlhip_matrix_get_metadata", "tnlhip_segments_metadata", "tnlhip_matrix_get_storage", "tnlhip_matrix_get_diagonal",
    "tnlhip_matrix_set_param", "tnlhip_spmv", "tnlhip_dot", "tnlhip_norm2", "tnlhip_axpby", "tnlhip_multidot",
    "tnlhip_gemv", "tnlhip_exclusive_scan_i32", "tnlhip_solver_params_default", "tnlhip_solver_create",
    "tnlhip_solver_destroy", "tnlhip_solver_set_matrix", "tnlhip_solver_set_params", "tnlhip_solver_solve",
    "tnlhip_solver_get_status", "tnlhip_solver_get_history", "tnlhip_refinement_params_default",
    "tnlhip_refine_solve", "tnlhip_refine_context_create", "tnlhip_refine_context_destroy",
    "tnlhip_refine_solve_in", "tnlhip_server_config_default", "tnlhip_request_params_default",
    "tnlhip_server_footprint", "tnlhip_server_create", "tnlhip_server_destroy", "tnlhip_server_add_matrix",
    "tnlhip_server_remove_matrix", "tnlhip_server_submit", "tnlhip_server_progress", "tnlhip_server_result",
    "tnlhip_server_stats", "tnlhip_autotune_csr", "tnlhip_mtx_read_csr", "tnlhip_mtx_read", "tnlhip_mtx_write",
    "tnlhip_mtx_write_csr", "tnlhip_matrix_save", "tnlhip_matrix_load", "tnlhip_file_inspect",
    "tnlhip_matrix_create_symmetric_csr", "tnlhip_matrix_create_staged", "tnlhip_matrix_reduce_rows",
    "tnlhip_matrix_get_row_capacities", "tnlhip_matrix_get_compressed_row_lengths", "tnlhip_matrix_get_row_capacity",
    "tnlhip_matrix_get_row", "tnlhip_matrix_get_view", "tnlhip_matrix_values_changed",
    "tnlhip_solver_get_history_iterations", "tnlhip_solver_set_monitor", "tnlhip_solver_set_residual_history_file",
    "tnlhip_solver_set_preconditioner_matrix", "tnlhip_multidot_noalloc", "tnlhip_spmv_multi",
]
# element order of TNL binary files: "cuda" = column-major (TNL's GPU matrices), "host" = row-major
FILE_LAYOUTS = {"cuda": 0, "host": 1}


class TnlHipError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"tnlhip status {status}: {msg}")
        self.status = status


class CapacityError(TnlHipError):
    """std::logic_error in TNL: a row capacity was exceeded."""


class BreakdownError(TnlHipError):
    """GMRES hit a zero Hessenberg diagonal (TNL throws 1)."""


class FileInfo(C.Structure):
    _fields_ = [("format", C.c_int32), ("value_dtype", C.c_int32), ("rows", C.c_int32), ("columns", C.c_int32),
                ("storage_size", C.c_int64), ("symmetric", C.c_int32), ("reserved", C.c_int32)]


class MatrixInfo(C.Structure):
    _fields_ = [("rows", C.c_int32), ("columns", C.c_int32), ("format", C.c_int32), ("value_dtype", C.c_int32),
                ("storage_size", C.c_int64), ("allocated", C.c_int64), ("nonzeros", C.c_int64),
                ("lanes_per_row", C.c_int32), ("symmetric", C.c_int32)]


MONITOR_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int32, C.c_double)


class MatrixView(C.Structure):
    _fields_ = [("format", C.c_int32), ("value_dtype", C.c_int32), ("rows", C.c_int32), ("columns", C.c_int32),
                ("values", C.c_void_p), ("column_indexes", C.c_void_p), ("offsets", C.c_void_p),
                ("segment_size", C.c_int32), ("aligned_size", C.c_int32), ("slice_offsets", C.c_void_p),
                ("slice_segment_sizes", C.c_void_p), ("row_to_chunk", C.c_void_p), ("row_to_slice", C.c_void_p),
                ("slices", C.c_void_p), ("row_perm", C.c_void_p), ("group_pointers", C.c_void_p),
                ("symmetric", C.c_int32), ("reserved", C.c_int32)]


class SolverParams(C.Structure):
    _fields_ = [("max_iterations", C.c_int32), ("min_iterations", C.c_int32), ("convergence_residue", C.c_double),
                ("divergence_residue", C.c_double), ("gmres_variant", C.c_int32), ("restarting_min", C.c_int32),
                ("restarting_max", C.c_int32), ("restarting_step_min", C.c_int32),
                ("restarting_step_max", C.c_int32), ("bicgstab_exact_residue", C.c_int32),
                ("preconditioner", C.c_int32), ("use_graph", C.c_int32), ("bicgstab_ell", C.c_int32),
                ("poly_degree", C.c_int32)]


REFINEMENT_HISTORY = 64


class RefinementParams(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("max_outer_iterations", C.c_int32), ("inner_method", C.c_int32),
                ("inner_value_dtype", C.c_int32), ("stagnation_limit", C.c_int32), ("inner", SolverParams),
                ("poly_value_dtype", C.c_int32), ("poly_matrix", C.c_void_p)]


class RefinementResult(C.Structure):
    _fields_ = [("converged", C.c_int32), ("outer_iterations", C.c_int32), ("inner_iterations", C.c_int64),
                ("residue", C.c_double), ("history", C.c_double * REFINEMENT_HISTORY)]


class Tuning(C.Structure):
    _fields_ = [("format", C.c_int32), ("lanes", C.c_int32), ("blocks_per_cu", C.c_int32), ("pair_rows", C.c_int32),
                ("slice_size", C.c_int32), ("column_stripes", C.c_int32), ("seconds", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [("matrix_capacity", C.c_int64), ("slice_capacity", C.c_int64), ("vector_capacity", C.c_int64),
                ("staging_capacity", C.c_int64), ("max_requests", C.c_int32), ("iterations_per_launch", C.c_int32),
                ("use_graph", C.c_int32), ("gmres_workers", C.c_int32)]


class RequestParams(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("inner_tolerance", C.c_double), ("max_outer_iterations", C.c_int32),
                ("max_inner_iterations", C.c_int32), ("method", C.c_int32), ("stagnation_limit", C.c_int32)]


class RequestResult(C.Structure):
    _fields_ = [("converged", C.c_int32), ("outer_iterations", C.c_int32), ("inner_iterations", C.c_int64),
                ("residue", C.c_double), ("launches", C.c_int64)]


class ServerStats(C.Structure):
    _fields_ = [("resident_requests", C.c_int32), ("queued_requests", C.c_int32), ("resident_entries", C.c_int64),
                ("vector_free", C.c_int64), ("matrix_free", C.c_int64), ("launches", C.c_int64),
                ("tiles", C.c_int64)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libtnlhip.so is not built ({LIB_PATH}); run __graft_entry__.build()")
    # torch's wheel ships its own HIP runtime under the same soname (libamdhip64.so.7).  Loaded
    # first, it is the one runtime of the process and libtnlhip.so binds to it; loaded after the
    # library, the two runtimes see no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, f64 = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    sig = {
        "tnlhip_status_string": (C.c_char_p, [C.c_int]),
        "tnlhip_last_error": (C.c_char_p, []),
        "tnlhip_device_info": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(i64)]),
        "tnlhip_matrix_create_csr": (C.c_int, [C.c_int, C.c_int, i32, i32, vp, vp, vp, C.c_int, C.POINTER(vp)]),
        "tnlhip_matrix_convert": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(vp)]),
        "tnlhip_matrix_create_staged": (C.c_int, [C.c_int, C.c_int, C.c_int, i32, i32, vp, vp, vp, C.POINTER(vp)]),
        "tnlhip_matrix_create_symmetric_csr": (C.c_int, [C.c_int, C.c_int, i32, vp, vp, vp, C.c_int, C.POINTER(vp)]),
        "tnlhip_matrix_destroy": (C.c_int, [vp]),
        "tnlhip_matrix_get_info": (C.c_int, [vp, C.POINTER(MatrixInfo)]),
        "tnlhip_matrix_get_metadata": (i64, [vp, C.c_char_p, vp]),
        "tnlhip_segments_metadata": (i64, [C.c_int, i32, vp, C.c_char_p, vp]),
        "tnlhip_matrix_get_storage": (C.c_int, [vp, vp, vp]),
        "tnlhip_matrix_get_diagonal": (C.c_int, [vp, vp, vp]),
        "tnlhip_matrix_set_param": (C.c_int, [vp, C.c_char_p, i32]),
        "tnlhip_spmv": (C.c_int, [vp, vp, vp, f64, f64, i32, i32, vp]),
        "tnlhip_spmv_multi": (C.c_int, [vp, i32, vp, vp, f64, f64, i32, i32, vp]),
        "tnlhip_dot": (C.c_int, [C.c_int, i64, vp, vp, C.POINTER(f64), vp]),
        "tnlhip_norm2": (C.c_int, [C.c_int, i64, vp, C.POINTER(f64), vp]),
        "tnlhip_axpby": (C.c_int, [C.c_int, i64, f64, vp, f64, vp, vp]),
        "tnlhip_multidot": (C.c_int, [C.c_int, i64, i32, vp, i64, vp, vp, vp]),
        "tnlhip_multidot_noalloc": (C.c_int, [C.c_int, i64, i32, vp, i64, vp, vp, vp]),
        "tnlhip_gemv": (C.c_int, [C.c_int, i64, i32, f64, vp, i64, vp, f64, vp, vp]),
        "tnlhip_exclusive_scan_i32": (C.c_int, [i64, vp, vp, C.c_int, vp]),
        "tnlhip_solver_params_default": (None, [C.POINTER(SolverParams)]),
        "tnlhip_solver_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "tnlhip_solver_destroy": (C.c_int, [vp]),
        "tnlhip_solver_set_matrix": (C.c_int, [vp, vp]),
        "tnlhip_solver_set_preconditioner_matrix": (C.c_int, [vp, vp]),
        "tnlhip_solver_set_params": (C.c_int, [vp, C.POINTER(SolverParams)]),
        "tnlhip_solver_solve": (C.c_int, [vp, vp, vp, C.POINTER(i32), vp]),
        "tnlhip_solver_get_status": (C.c_int, [vp, C.POINTER(i32), C.POINTER(f64)]),
        "tnlhip_solver_get_history": (i64, [vp, vp, i64]),
        "tnlhip_solver_get_history_iterations": (i64, [vp, vp, i64]),
        "tnlhip_solver_set_monitor": (C.c_int, [vp, MONITOR_FN, vp]),
        "tnlhip_solver_set_residual_history_file": (C.c_int, [vp, C.c_char_p]),
        "tnlhip_refinement_params_default": (None, [C.POINTER(RefinementParams)]),
        "tnlhip_refine_solve": (C.c_int, [vp, vp, C.POINTER(RefinementParams), vp, vp,
                                          C.POINTER(RefinementResult), vp]),
        "tnlhip_refine_context_create": (C.c_int, [C.POINTER(C.c_void_p)]),
        "tnlhip_refine_context_destroy": (C.c_int, [vp]),
        "tnlhip_refine_solve_in": (C.c_int, [vp, vp, vp, C.POINTER(RefinementParams), vp, vp,
                                             C.POINTER(RefinementResult), vp]),
        "tnlhip_server_config_default": (None, [C.POINTER(ServerConfig)]),
        "tnlhip_request_params_default": (None, [C.POINTER(RequestParams)]),
        "tnlhip_server_footprint": (C.c_int, [i32, i64, i32, i32, C.POINTER(i64), C.POINTER(i64)]),
        "tnlhip_server_create": (C.c_int, [C.POINTER(ServerConfig), C.POINTER(vp)]),
        "tnlhip_server_destroy": (C.c_int, [vp]),
        "tnlhip_server_add_matrix": (C.c_int, [vp, i32, vp, vp, vp, C.POINTER(i32)]),
        "tnlhip_server_remove_matrix": (C.c_int, [vp, i32]),
        "tnlhip_server_submit": (C.c_int, [vp, i32, f64, i32, vp, vp, C.POINTER(RequestParams), C.POINTER(i64)]),
        "tnlhip_server_progress": (C.c_int, [vp, vp, i32, C.POINTER(i32)]),
        "tnlhip_server_result": (C.c_int, [vp, i64, C.POINTER(RequestResult)]),
        "tnlhip_server_stats": (C.c_int, [vp, C.POINTER(ServerStats)]),
        "tnlhip_mtx_read_csr": (C.c_int, [C.c_char_p, C.POINTER(i32), C.POINTER(i32), C.POINTER(i64), vp, vp, vp]),
        "tnlhip_mtx_read": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.POINTER(vp)]),
        "tnlhip_mtx_write": (C.c_int, [vp, C.c_char_p]),
        "tnlhip_mtx_write_csr": (C.c_int, [C.c_char_p, i32, i32, vp, vp, vp]),
        "tnlhip_matrix_save": (C.c_int, [vp, C.c_char_p, C.c_int]),
        "tnlhip_matrix_load": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]),
        "tnlhip_file_inspect": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(FileInfo), C.c_char_p, C.c_int]),
        "tnlhip_matrix_reduce_rows": (C.c_int, [vp, i32, i32, C.c_int, C.c_int, f64, vp, C.c_int, vp]),
        "tnlhip_matrix_get_row_capacities": (C.c_int, [vp, vp, vp]),
        "tnlhip_matrix_get_compressed_row_lengths": (C.c_int, [vp, vp, vp]),
        "tnlhip_matrix_get_row_capacity": (C.c_int, [vp, i32, C.POINTER(i32)]),
        "tnlhip_matrix_get_row": (C.c_int, [vp, i32, vp, vp, i32, C.POINTER(i32)]),
        "tnlhip_matrix_values_changed": (C.c_int, [vp, vp]),
        "tnlhip_matrix_get_view": (C.c_int, [vp, C.POINTER(MatrixView)]),
        "tnlhip_autotune_csr": (C.c_int, [C.c_int, i32, i32, vp, vp, vp, i32, C.c_uint32, i32, C.POINTER(Tuning),
                                          C.POINTER(Tuning), i32, C.POINTER(i32), C.POINTER(vp)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


lib = _load()


def _check(status):
    if status != 0:
        msg = lib.tnlhip_last_error().decode(errors="replace")
        if status == 4:
            raise CapacityError(status, msg)
        if status == 6:
            raise BreakdownError(status, msg)
        raise TnlHipError(status, msg)


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def _stream(stream):
    if stream is None:
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    return C.c_void_p(stream if isinstance(stream, int) else stream.cuda_stream)


def device_info():
    name = C.create_string_buffer(128)
    cus = C.c_int(0)
    mem = C.c_int64(0)
    _check(lib.tnlhip_device_info(name, 128, C.byref(cus), C.byref(mem)))
    return {"name": name.value.decode(), "compute_units": cus.value, "memory": mem.value}


def _dtype_id(dtype):
    if isinstance(dtype, int):
        return dtype
    return DTYPES[str(dtype).replace("torch.", "").replace("float64", "f64").replace("float32", "f32")
                  .replace("bfloat16", "bf16")]


FETCH = {"value": 0, "abs": 1, "one": 2, "nonzero": 3, "column": 4}
REDUCE = {"sum": 0, "min": 1, "max": 2}


class SparseMatrix:
    """SparseMatrix< Real, Devices::Hip, int, GeneralMatrix, Segments > (Matrices/SparseMatrix.h:40-49)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle) if isinstance(handle, int) else handle
        info = MatrixInfo()
        _check(lib.tnlhip_matrix_get_info(self._h, C.byref(info)))
        self.info = info

    @classmethod
    def from_csr(cls, row_pointers, column_indexes, values, columns=None, format="CSRScalar", dtype="f64"):
        """A = csr (SparseMatrix::operator=, Matrices/SparseMatrix.hpp:1070-1180) from host numpy
        arrays or device torch tensors (int32 row pointers / columns, float64 values)."""
        import numpy as np
        fmt = FORMATS[format] if isinstance(format, str) else format
        dt = _dtype_id(dtype)
        h = C.c_void_p(0)
        if isinstance(row_pointers, np.ndarray):
            rp = np.ascontiguousarray(row_pointers, np.int32)
            ci = np.ascontiguousarray(column_indexes, np.int32)
            va = None if values is None else np.ascontiguousarray(values, np.float64)  # None: a binary matrix of ones
            rows = len(rp) - 1
            cols = rows if columns is None else columns
            _check(lib.tnlhip_matrix_create_csr(fmt, dt, rows, cols, rp.ctypes.data_as(C.c_void_p),
                                                ci.ctypes.data_as(C.c_void_p),
                                                None if va is None else va.ctypes.data_as(C.c_void_p), 1, C.byref(h)))
        else:
            rows = row_pointers.numel() - 1
            cols = rows if columns is None else columns
            _check(lib.tnlhip_matrix_create_csr(fmt, dt, rows, cols, _ptr(row_pointers), _ptr(column_indexes),
                                                _ptr(values), 0, C.byref(h)))
        return cls(h)

    @classmethod
    def symmetric_from_csr(cls, row_pointers, column_indexes, values, format="CSRScalar", dtype="f64"):
        """SparseMatrix< ..., SymmetricMatrix, ... > = csr of the lower triangle (diagonal included);
        host numpy arrays."""
        import numpy as np
        rp = np.ascontiguousarray(row_pointers, np.int32)
        ci = np.ascontiguousarray(column_indexes, np.int32)
        va = np.ascontiguousarray(values, np.float64)
        h = C.c_void_p(0)
        _check(lib.tnlhip_matrix_create_symmetric_csr(FORMATS[format] if isinstance(format, str) else format,
                                                      _dtype_id(dtype), len(rp) - 1, rp.ctypes.data_as(C.c_void_p),
                                                      ci.ctypes.data_as(C.c_void_p), va.ctypes.data_as(C.c_void_p),
                                                      1, C.byref(h)))
        return cls(h)

    @classmethod
    def staged(cls, row_pointers, column_indexes, values, columns=None, format="CSRScalar", dtype="f64",
               symmetric=False):
        """setRowCapacities( row lengths ) + setElement of every stored entry (host numpy arrays):
        capacities kept, slots with column -1 left as padding, explicit zeros stored."""
        import numpy as np
        rp = np.ascontiguousarray(row_pointers, np.int32)
        ci = np.ascontiguousarray(column_indexes, np.int32)
        va = np.ascontiguousarray(values, np.float64)
        rows = len(rp) - 1
        h = C.c_void_p(0)
        _check(lib.tnlhip_matrix_create_staged(FORMATS[format] if isinstance(format, str) else format,
                                               _dtype_id(dtype), 1 if symmetric else 0, rows,
                                               rows if columns is None else columns, rp.ctypes.data_as(C.c_void_p),
                                               ci.ctypes.data_as(C.c_void_p), va.ctypes.data_as(C.c_void_p),
                                               C.byref(h)))
        return cls(h)

    @property
    def symmetric(self):
        return bool(self.info.symmetric)

    @classmethod
    def from_mtx(cls, path, format="CSRScalar", dtype="f64"):
        """MatrixReader< SparseMatrix< Real, Devices::Hip, ... > >::readMtx( path, A )."""
        fmt = FORMATS[format] if isinstance(format, str) else format
        h = C.c_void_p(0)
        _check(lib.tnlhip_mtx_read(str(path).encode(), fmt, _dtype_id(dtype), C.byref(h)))
        return cls(h)

    def write_mtx(self, path):
        """MatrixWriter::writeMtx( path, A )."""
        _check(lib.tnlhip_mtx_write(self._h, str(path).encode()))

    def save(self, path, layout="cuda"):
        """SparseMatrix::save( fileName ): TNL's binary File / Object format (SparseMatrix.hpp:1266-1270)."""
        _check(lib.tnlhip_matrix_save(self._h, str(path).encode(), FILE_LAYOUTS[layout]))

    @classmethod
    def load(cls, path, format=None, dtype=None, layout="cuda"):
        """SparseMatrix::load( fileName ) (SparseMatrix.hpp:1282-1286); format / dtype None take the file's."""
        fmt = -1 if format is None else (FORMATS[format] if isinstance(format, str) else format)
        dt = -1 if dtype is None else _dtype_id(dtype)
        h = C.c_void_p(0)
        _check(lib.tnlhip_matrix_load(str(path).encode(), fmt, dt, FILE_LAYOUTS[layout], C.byref(h)))
        return cls(h)

    def __del__(self):
        try:
            if self._h:
                lib.tnlhip_matrix_destroy(self._h)
                self._h = None
        except Exception:
            pass

    # --- TNL-style accessors
    def getRows(self):
        return self.info.rows

    def getColumns(self):
        return self.info.columns

    def getNonzeroElementsCount(self):
        return self.info.nonzeros

    def getStorageSize(self):
        return self.info.storage_size

    @property
    def format(self):
        return FORMAT_NAMES[self.info.format]

    @property
    def value_dtype(self):
        return {0: "f64", 1: "f32", 2: "bf16", 3: "binary"}[self.info.value_dtype]

    @property
    def vector_dtype(self):
        import torch
        return {0: torch.float64, 3: torch.int32}.get(self.info.value_dtype, torch.float32)

    def convert(self, format, dtype=None):
        """B = A into another format / value type, on the device."""
        fmt = FORMATS[format] if isinstance(format, str) else format
        dt = self.info.value_dtype if dtype is None else _dtype_id(dtype)
        h = C.c_void_p(0)
        _check(lib.tnlhip_matrix_convert(self._h, fmt, dt, C.byref(h)))
        return SparseMatrix(h)

    def metadata(self, name):
        import numpy as np
        n = lib.tnlhip_matrix_get_metadata(self._h, name.encode(), None)
        if n < 0:
            raise KeyError(name)
        out = np.zeros(max(n, 1), np.int32)
        lib.tnlhip_matrix_get_metadata(self._h, name.encode(), out.ctypes.data_as(C.c_void_p))
        return out[:n]

    def storage(self):
        """(values as float64, column indexes) of the allocated storage, host copies."""
        import numpy as np
        n = self.info.allocated
        v = np.zeros(max(n, 1), np.float64)
        c = np.zeros(max(n, 1), np.int32)
        _check(lib.tnlhip_matrix_get_storage(self._h, v.ctypes.data_as(C.c_void_p), c.ctypes.data_as(C.c_void_p)))
        return v[:n], c[:n]

    def diagonal(self, out, stream=None):
        _check(lib.tnlhip_matrix_get_diagonal(self._h, _ptr(out), _stream(stream)))
        return out

    def set_param(self, name, value):
        _check(lib.tnlhip_matrix_set_param(self._h, name.encode(), int(value)))
        info = MatrixInfo()
        _check(lib.tnlhip_matrix_get_info(self._h, C.byref(info)))
        self.info = info

    # ---- row traversal (SparseMatrixView.hpp:116-250, 495-595)

    def reduceRows(self, begin, end, fetch, reduce, identity, out, stream=None):
        """out[row] = identity (+) fetch(row's slots in local order) for rows [begin, end) (end 0 =
        all), with fetch in FETCH ("value", "abs", "one", "nonzero", "column") and reduce in REDUCE
        ("sum", "min", "max"); out is a float64 / float32 / int32 device tensor of length rows."""
        import torch
        odt = {torch.float64: 0, torch.float32: 1, torch.int32: 4}[out.dtype]
        if out.numel() != self.info.rows:
            raise ValueError("out must have one entry per row")
        _check(lib.tnlhip_matrix_reduce_rows(self._h, int(begin), int(end), FETCH[fetch], REDUCE[reduce],
                                             float(identity), _ptr(out), odt, _stream(stream)))
        return out

    def reduceAllRows(self, fetch, reduce, identity, out, stream=None):
        return self.reduceRows(0, 0, fetch, reduce, identity, out, stream)

    def getRowCapacities(self, out=None, stream=None):
        """getRowCapacities: stored (non-padding) entries per row, as TNL's reduceAllRows counts them."""
        import torch
        out = torch.empty(self.info.rows, dtype=torch.int32, device="cuda") if out is None else out
        _check(lib.tnlhip_matrix_get_row_capacities(self._h, _ptr(out), _stream(stream)))
        return out

    def getCompressedRowLengths(self, out=None, stream=None):
        """getCompressedRowLengths: entries per row whose value is not zero."""
        import torch
        out = torch.empty(self.info.rows, dtype=torch.int32, device="cuda") if out is None else out
        _check(lib.tnlhip_matrix_get_compressed_row_lengths(self._h, _ptr(out), _stream(stream)))
        return out

    def getRowCapacity(self, row):
        """getRowCapacity( row ): the segment size of the row."""
        c = C.c_int32(0)
        _check(lib.tnlhip_matrix_get_row_capacity(self._h, int(row), C.byref(c)))
        return c.value

    def getRow(self, row):
        """getRow( row ) read on the host: (columns, values) of the row's slots in local order
        (padding slots: column -1, value 0)."""
        import numpy as np
        size = C.c_int32(0)
        _check(lib.tnlhip_matrix_get_row(self._h, int(row), None, None, 0, C.byref(size)))
        n = size.value
        cols = np.zeros(max(n, 1), np.int32)
        vals = np.zeros(max(n, 1), np.float64)
        _check(lib.tnlhip_matrix_get_row(self._h, int(row), cols.ctypes.data_as(C.c_void_p),
                                         vals.ctypes.data_as(C.c_void_p), max(n, 1), C.byref(size)))
        return cols[:n], vals[:n]

    def vectorProduct(self, inVector, outVector, matrixMultiplicator=1.0, outVectorMultiplicator=0.0,
                      firstRow=0, lastRow=0, stream=None):
        """outVector = matrixMultiplicator * A inVector + outVectorMultiplicator * outVector over rows
        [firstRow, lastRow) (lastRow 0 = all), SparseMatrix.h:1078-1085.  Asynchronous on the stream."""
        dt = self.vector_dtype
        if inVector.dtype != dt or outVector.dtype != dt:
            raise TypeError(f"vectors must be {dt}")
        if inVector.numel() != self.info.columns or outVector.numel() != self.info.rows:
            raise ValueError("vector sizes do not match the matrix")
        _check(lib.tnlhip_spmv(self._h, _ptr(inVector), _ptr(outVector), float(matrixMultiplicator),
                               float(outVectorMultiplicator), int(firstRow), int(lastRow), _stream(stream)))
        return outVector

    def vectorProductMulti(self, inVectors, outVectors, matrixMultiplicator=1.0, outVectorMultiplicator=0.0,
                           firstRow=0, lastRow=0, stream=None):
        """vectorProduct for k right-hand sides in one pass over the matrix: inVectors (columns, k)
        and outVectors (rows, k) are contiguous row-major device tensors, 1 <= k <= 64."""
        dt = self.vector_dtype
        if inVectors.dtype != dt or outVectors.dtype != dt:
            raise TypeError(f"vectors must be {dt}")
        if inVectors.dim() != 2 or outVectors.dim() != 2 or inVectors.shape[1] != outVectors.shape[1]:
            raise ValueError("blocks of vectors are (size, k) tensors with the same k")
        if not (inVectors.is_contiguous() and outVectors.is_contiguous()):
            raise ValueError("blocks of vectors must be contiguous (row-major)")
        if inVectors.shape[0] != self.info.columns or outVectors.shape[0] != self.info.rows:
            raise ValueError("vector sizes do not match the matrix")
        _check(lib.tnlhip_spmv_multi(self._h, int(inVectors.shape[1]), _ptr(inVectors), _ptr(outVectors),
                                     float(matrixMultiplicator), float(outVectorMultiplicator), int(firstRow),
                                     int(lastRow), _stream(stream)))
        return outVectors


def segments_metadata(format, capacities, name):
    """Host-only segments metadata (no GPU needed): setSegmentsSizes on the host."""
    import numpy as np
    fmt = FORMATS[format] if isinstance(format, str) else format
    caps = np.ascontiguousarray(capacities, np.int32)
    n = lib.tnlhip_segments_metadata(fmt, len(caps), caps.ctypes.data_as(C.c_void_p), name.encode(), None)
    if n < 0:
        raise KeyError(name)
    out = np.zeros(max(n, 1), np.int32)
    lib.tnlhip_segments_metadata(fmt, len(caps), caps.ctypes.data_as(C.c_void_p), name.encode(),
                                 out.ctypes.data_as(C.c_void_p))
    out = out[:n]
    if name == "storage":
        return int(out[0].astype(np.uint32)) + (int(out[1]) << 32)
    return out


# ------------------------------------------------------------------ BLAS-1


def _vdt(t):
    import torch
    if t.dtype == torch.float64:
        return 0
    if t.dtype == torch.float32:
        return 1
    raise TypeError("vectors are float64 or float32")


def dot(a, b, stream=None):
    """(a, b) — deterministic device reduction."""
    r = C.c_double(0)
    _check(lib.tnlhip_dot(_vdt(a), a.numel(), _ptr(a), _ptr(b), C.byref(r), _stream(stream)))
    return r.value


def lpNorm2(a, stream=None):
    r = C.c_double(0)
    _check(lib.tnlhip_norm2(_vdt(a), a.numel(), _ptr(a), C.byref(r), _stream(stream)))
    return r.value


def axpby(alpha, x, beta, y, stream=None):
    _check(lib.tnlhip_axpby(_vdt(x), x.numel(), float(alpha), _ptr(x), float(beta), _ptr(y), _stream(stream)))
    return y


def multidot(V, ld, count, w, stream=None):
    import numpy as np
    out = np.zeros(count, np.float64)
    _check(lib.tnlhip_multidot(_vdt(w), w.numel(), count, _ptr(V), ld, _ptr(w), out.ctypes.data_as(C.c_void_p),
                               _stream(stream)))
    return out


def gemv(alpha, V, ld, count, x, beta, y, stream=None):
    import numpy as np
    xx = np.ascontiguousarray(x, np.float64)
    _check(lib.tnlhip_gemv(_vdt(y), y.numel(), count, float(alpha), _ptr(V), ld, xx.ctypes.data_as(C.c_void_p),
                           float(beta), _ptr(y), _stream(stream)))
    return y


def exclusive_scan(inp, out, total=True, stream=None):
    n = inp.numel()
    _check(lib.tnlhip_exclusive_scan_i32(n, _ptr(inp), _ptr(out), 1 if total else 0, _stream(stream)))
    return out


# ------------------------------------------------------------------ solvers


class Diagonal:
    """Preconditioners::Diagonal (Jacobi), Solvers/Linear/Preconditioners/Diagonal.h."""
    kind = 1


class ILU0:
    """Preconditioners::ILU0, Solvers/Linear/Preconditioners/ILU0.h: factors with A's sparsity,
    factorized and applied level by level on the device."""
    kind = 2


class JacobiPolynomial:
    """Jacobi-polynomial preconditioner (no TNL counterpart): z = sum_{j<=k} (I - D^-1 P)^j D^-1 s,
    k = setup(poly_degree=...), P the solver's preconditioner matrix (setPreconditionerMatrix; the
    solved matrix by default)."""
    kind = 3


class LinearSolver:
    """LinearSolver< Matrix > (Solvers/Linear/LinearSolver.h:115-175) with IterativeSolver's setters."""
    method = None

    def __init__(self):
        h = C.c_void_p(0)
        _check(lib.tnlhip_solver_create(METHODS[self.method], C.byref(h)))
        self._h = h
        self.params = SolverParams()
        lib.tnlhip_solver_params_default(C.byref(self.params))
        self._matrix = None

    def __del__(self):
        try:
            if self._h:
                lib.tnlhip_solver_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def setMatrix(self, matrix):
        self._matrix = matrix
        _check(lib.tnlhip_solver_set_matrix(self._h, matrix._h))

    def setPreconditioner(self, preconditioner):
        self.params.preconditioner = 0 if preconditioner is None else getattr(preconditioner, "kind", 1)

    def setPreconditionerMatrix(self, matrix):
        """P of the Jacobi-polynomial preconditioner (None: the solved matrix); kept alive here."""
        self._poly_matrix = matrix
        _check(lib.tnlhip_solver_set_preconditioner_matrix(self._h, None if matrix is None else matrix._h))

    def setMaxIterations(self, n):
        self.params.max_iterations = int(n)

    def setMinIterations(self, n):
        self.params.min_iterations = int(n)

    def setConvergenceResidue(self, r):
        self.params.convergence_residue = float(r)

    def setDivergenceResidue(self, r):
        self.params.divergence_residue = float(r)

    def setup(self, **parameters):
        """Config-style parameters: max_iterations, convergence_residue, gmres_variant, ..."""
        for k, v in parameters.items():
            k = k.replace("-", "_")
            if k == "gmres_variant" and isinstance(v, str):
                v = GMRES_VARIANTS[v]
            if k == "preconditioner" and isinstance(v, str):
                v = {"none": 0, "diagonal": 1, "ilu0": 2, "jacobi-polynomial": 3}[v]
            setattr(self.params, k, v)
        return True

    def solve(self, b, x, stream=None):
        """bool solve( b, x ); x holds the initial guess.  Synchronous."""
        _check(lib.tnlhip_solver_set_params(self._h, C.byref(self.params)))
        conv = C.c_int32(0)
        _check(lib.tnlhip_solver_solve(self._h, _ptr(b), _ptr(x), C.byref(conv), _stream(stream)))
        return bool(conv.value)

    def getIterations(self):
        it = C.c_int32(0)
        r = C.c_double(0)
        _check(lib.tnlhip_solver_get_status(self._h, C.byref(it), C.byref(r)))
        return it.value

    def getResidue(self):
        it = C.c_int32(0)
        r = C.c_double(0)
        _check(lib.tnlhip_solver_get_status(self._h, C.byref(it), C.byref(r)))
        return r.value

    def getResidualHistory(self):
        import numpy as np
        n = lib.tnlhip_solver_get_history(self._h, None, 0)
        out = np.zeros(max(n, 1), np.float64)
        lib.tnlhip_solver_get_history(self._h, out.ctypes.data_as(C.c_void_p), n)
        return out[:n]

    def getResidualHistoryIterations(self):
        """getIterations() at each setResidue call of the last solve."""
        import numpy as np
        n = lib.tnlhip_solver_get_history_iterations(self._h, None, 0)
        out = np.zeros(max(n, 1), np.int32)
        lib.tnlhip_solver_get_history_iterations(self._h, out.ctypes.data_as(C.c_void_p), n)
        return out[:n]

    def setSolverMonitor(self, monitor):
        """IterativeSolver::setSolverMonitor: monitor(iterations, residue) is called each time the
        solve reads its device state; None removes it."""
        if monitor is None:
            self._monitor = None
            _check(lib.tnlhip_solver_set_monitor(self._h, MONITOR_FN(), None))
            return
        self._monitor = MONITOR_FN(lambda _user, it, res: monitor(int(it), float(res)))
        _check(lib.tnlhip_solver_set_monitor(self._h, self._monitor, None))

    def setResidualHistoryFile(self, path):
        """setup( residual-history-file ): each solve appends its residues in TNL's format."""
        _check(lib.tnlhip_solver_set_residual_history_file(self._h, None if not path else str(path).encode()))


class CG(LinearSolver):
    method = "cg"


class BICGStab(LinearSolver):
    method = "bicgstab"


class GMRES(LinearSolver):
    method = "gmres"


class TFQMR(LinearSolver):
    method = "tfqmr"


class BICGStabL(LinearSolver):
    """BiCGStab(ell), Solvers/Linear/BICGStabL.h; setup(bicgstab_ell=..., bicgstab_exact_residue=...)."""
    method = "bicgstabl"


class RefineContext:
    """Buffers and inner solver of refine_solve kept between solves (no TNL counterpart): pass it
    as refine_solve(..., context=ctx) and only the first solve of the largest size allocates."""

    def __init__(self):
        h = C.c_void_p()
        _check(lib.tnlhip_refine_context_create(C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib.tnlhip_refine_context_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def refine_solve(A, b, x, A_low=None, tolerance=1e-10, inner_method="gmres", inner_dtype="f32",
                 max_outer_iterations=50, stagnation_limit=2, stream=None, poly_dtype=None, P=None, context=None,
                 **inner):
    """FP64 iterative refinement around a low-precision inner solve (no TNL counterpart).
    Returns a dict with converged, outer_iterations, inner_iterations, residue, history.  With a
    RefineContext the solve reuses its buffers and inner solver."""
    p = RefinementParams()
    lib.tnlhip_refinement_params_default(C.byref(p))
    p.tolerance = float(tolerance)
    p.max_outer_iterations = int(max_outer_iterations)
    p.stagnation_limit = int(stagnation_limit)
    p.inner_method = METHODS[inner_method] if isinstance(inner_method, str) else inner_method
    p.inner_value_dtype = _dtype_id(inner_dtype)
    if poly_dtype is not None:
        p.poly_value_dtype = _dtype_id(poly_dtype)
    if P is not None:
        p.poly_matrix = P._h
    for k, v in inner.items():
        if k == "gmres_variant" and isinstance(v, str):
            v = GMRES_VARIANTS[v]
        setattr(p.inner, k, v)
    r = RefinementResult()
    low = A_low._h if A_low is not None else None
    if context is not None:
        _check(lib.tnlhip_refine_solve_in(context._h, A._h, low, C.byref(p), _ptr(b), _ptr(x), C.byref(r),
                                          _stream(stream)))
    else:
        _check(lib.tnlhip_refine_solve(A._h, low, C.byref(p), _ptr(b), _ptr(x), C.byref(r), _stream(stream)))
    return {"converged": bool(r.converged), "outer_iterations": r.outer_iterations,
            "inner_iterations": r.inner_iterations, "residue": r.residue,
            "history": list(r.history[:min(r.outer_iterations + 1, REFINEMENT_HISTORY)])}


class Server:
    """Solve server: resident matrices, requests (A + shift I) X = B with 1-16 right-hand sides,
    FP64 refinement over FP32 CG, advanced together by replayed persistent kernels (no TNL
    counterpart).  B and X are host float64 arrays of shape (rows, k) in row-major order."""

    def __init__(self, vector_capacity=None, matrix_capacity=None, slice_capacity=None, max_requests=None,
                 iterations_per_launch=None, use_graph=True, staging_capacity=None, gmres_workers=None):
        cfg = ServerConfig()
        lib.tnlhip_server_config_default(C.byref(cfg))
        for k, v in dict(vector_capacity=vector_capacity, matrix_capacity=matrix_capacity,
                         slice_capacity=slice_capacity, max_requests=max_requests,
                         iterations_per_launch=iterations_per_launch, staging_capacity=staging_capacity,
                         gmres_workers=gmres_workers).items():
            if v is not None:
                setattr(cfg, k, int(v))
        cfg.use_graph = 1 if use_graph else 0
        self.config = cfg
        h = C.c_void_p()
        _check(lib.tnlhip_server_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self._live = {}  # ticket -> (b, x) kept alive until finished

    def close(self):
        if self._h:
            lib.tnlhip_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_matrix(self, row_ptr, cols, vals):
        rp = np.ascontiguousarray(row_ptr, np.int32)
        c = np.ascontiguousarray(cols, np.int32)
        v = np.ascontiguousarray(vals, np.float64)
        mid = C.c_int32()
        _check(lib.tnlhip_server_add_matrix(self._h, len(rp) - 1, rp.ctypes.data, c.ctypes.data, v.ctypes.data,
                                            C.byref(mid)))
        return mid.value

    def remove_matrix(self, mid):
        _check(lib.tnlhip_server_remove_matrix(self._h, int(mid)))

    def submit(self, mid, B, shift=0.0, tolerance=1e-10, inner_tolerance=3e-6, max_outer_iterations=20,
               max_inner_iterations=5000, stagnation_limit=2, method="cg"):
        """Queue (A + shift I) X = B; method "cg" (batched track) or "gmres" (worker track)."""
        B = np.ascontiguousarray(B, np.float64)
        if B.ndim == 1:
            B = B[:, None]
        X = np.zeros_like(B)
        p = RequestParams()
        lib.tnlhip_request_params_default(C.byref(p))
        p.tolerance, p.inner_tolerance = float(tolerance), float(inner_tolerance)
        p.max_outer_iterations, p.max_inner_iterations = int(max_outer_iterations), int(max_inner_iterations)
        p.stagnation_limit = int(stagnation_limit)
        p.method = METHODS[method] if isinstance(method, str) else int(method)
        t = C.c_int64()
        _check(lib.tnlhip_server_submit(self._h, int(mid), float(shift), B.shape[1], B.ctypes.data, X.ctypes.data,
                                        C.byref(p), C.byref(t)))
        self._live[t.value] = (B, X)
        return t.value

    def progress(self):
        """Admit, run one replay, retire; returns the finished tickets."""
        cap = max(1, self.config.max_requests) + 64 * max(1, self.config.gmres_workers)
        out = (C.c_int64 * cap)()
        n = C.c_int32()
        _check(lib.tnlhip_server_progress(self._h, out, cap, C.byref(n)))
        return list(out[:min(n.value, cap)])

    def result(self, ticket):
        """(X, info) of a finished ticket."""
        r = RequestResult()
        _check(lib.tnlhip_server_result(self._h, int(ticket), C.byref(r)))
        B, X = self._live.pop(ticket)
        return X, {"converged": bool(r.converged), "outer_iterations": r.outer_iterations,
                   "inner_iterations": r.inner_iterations, "residue": r.residue, "launches": r.launches}

    def stats(self):
        s = ServerStats()
        _check(lib.tnlhip_server_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def run_all(self):
        """Progress until nothing is resident or queued; returns {ticket: (X, info)}."""
        done = {}
        while True:
            for t in self.progress():
                done[t] = self.result(t)
            s = self.stats()
            if s["resident_requests"] == 0 and s["queued_requests"] == 0:
                return done


def autotune(row_pointers, column_indexes, values, dtype="f64", formats=None, repetitions=10, columns=None):
    """Time y = A x for every format (or the given list) and launch configuration; returns
    (the fastest SparseMatrix, its configuration, every timed configuration).  Host numpy CSR."""
    rp = np.ascontiguousarray(row_pointers, np.int32)
    ci = np.ascontiguousarray(column_indexes, np.int32)
    va = np.ascontiguousarray(values, np.float64)
    rows = len(rp) - 1
    cols = rows if columns is None else columns
    names = list(FORMATS) if formats is None else list(formats)
    mask = 0
    for f in names:
        mask |= 1 << FORMATS[f]
    best = Tuning()
    cap = 64
    allt = (Tuning * cap)()
    cnt = C.c_int32()
    h = C.c_void_p()
    _check(lib.tnlhip_autotune_csr(_dtype_id(dtype), rows, cols, rp.ctypes.data, ci.ctypes.data, va.ctypes.data, 1,
                                   mask, int(repetitions), C.byref(best), allt, cap, C.byref(cnt), C.byref(h)))

    def as_dict(t):
        return {"format": FORMAT_NAMES[t.format], "lanes": t.lanes, "blocks_per_cu": t.blocks_per_cu,
                "pair_rows": t.pair_rows, "slice_size": t.slice_size, "column_stripes": t.column_stripes,
                "seconds": t.seconds}

    return SparseMatrix(h), as_dict(best), [as_dict(allt[i]) for i in range(min(cnt.value, cap))]


def mtx_read_csr(path):
    """MatrixReader::readMtx into a host CSR matrix (Matrices/MatrixReader.hpp): returns
    (rows, columns, row_pointers, column_indexes, values) with TNL's capacities and padding."""
    rows, cols, st = C.c_int32(), C.c_int32(), C.c_int64()
    p = str(path).encode()
    _check(lib.tnlhip_mtx_read_csr(p, C.byref(rows), C.byref(cols), C.byref(st), None, None, None))
    rp = np.zeros(rows.value + 1, np.int32)
    ci = np.zeros(max(st.value, 1), np.int32)
    v = np.zeros(max(st.value, 1), np.float64)
    _check(lib.tnlhip_mtx_read_csr(p, C.byref(rows), C.byref(cols), C.byref(st), rp.ctypes.data, ci.ctypes.data,
                                   v.ctypes.data))
    return rows.value, cols.value, rp, ci[:st.value], v[:st.value]


def mtx_write_csr(path, rows, columns, row_pointers, column_indexes, values):
    """MatrixWriter::writeMtx of a host CSR matrix (Matrices/MatrixWriter.hpp:105-130)."""
    rp = np.ascontiguousarray(row_pointers, np.int32)
    ci = np.ascontiguousarray(column_indexes, np.int32)
    v = np.ascontiguousarray(values, np.float64)
    _check(lib.tnlhip_mtx_write_csr(str(path).encode(), int(rows), int(columns), rp.ctypes.data, ci.ctypes.data,
                                    v.ctypes.data))


def file_inspect(path, layout="cuda"):
    """getObjectType( fileName ) of a TNL binary matrix file plus the checks of SparseMatrix.load,
    on the host: {"type", "format", "dtype", "rows", "columns", "storage_size"}."""
    info = FileInfo()
    t = C.create_string_buffer(4096)
    _check(lib.tnlhip_file_inspect(str(path).encode(), FILE_LAYOUTS[layout], C.byref(info), t, 4096))
    return {"type": t.value.decode(), "format": FORMAT_NAMES[info.format],
            "dtype": {v: k for k, v in DTYPES.items()}[info.value_dtype], "rows": info.rows, "columns": info.columns,
            "storage_size": info.storage_size, "symmetric": bool(info.symmetric)}


def server_footprint(rows, nnz, max_row, nrhs):
    mb, rb = C.c_int64(), C.c_int64()
    _check(lib.tnlhip_server_footprint(int(rows), int(nnz), int(max_row), int(nrhs), C.byref(mb), C.byref(rb)))
    return mb.value, rb.value


def getLinearSolver(name):
    """LinearSolverTypeResolver::getLinearSolver by name (Solvers/LinearSolverTypeResolver.h:59-90)."""
    return {"cg": CG, "bicgstab": BICGStab, "gmres": GMRES, "tfqmr": TFQMR, "bicgstabl": BICGStabL}[name]()


def getPreconditioner(name):
    """LinearSolverTypeResolver::getPreconditioner by name (Solvers/LinearSolverTypeResolver.h:92-115):
    None for "none"."""
    return {"none": lambda: None, "diagonal": Diagonal, "ilu0": ILU0}[name]()
