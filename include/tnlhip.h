/*
 * tnlhip — C ABI of the MI355X (gfx950) implementation of TNL's sparse matrix-vector
 * product, its segment formats, the reduction / BLAS-1 kernels and the Krylov solvers
 * CG, BiCGStab and GMRES.
 *
 * TNL selects the device by a template tag (Devices::Host / Devices::Cuda); its GPU
 * boundary is SparseMatrix<Real, Device, Index, GeneralMatrix, Segments> with the Device
 * tag routed to the CUDA allocator and kernel launches.  This header is the boundary a
 * new Devices::Hip tag binds to: plain pointers, sizes and status codes, no C++ or torch
 * types.  Each entry point names the reference interface it replaces (paths relative to
 * TNL's src/).  The C++ front in tnl-dev_amd/TNL/ wraps it with TNL's own class names.
 *
 * Conventions
 *  - Index type is int32 (TNL's default `int`).
 *  - Device pointers are HIP device pointers; `stream` is a hipStream_t (NULL = default).
 *  - Vectors are float64 for F64 matrices, float32 for F32 and BF16 matrices and int32 for
 *    BINARY ones.
 *  - Every call returns a tnlhip_status; tnlhip_last_error() describes the last failure
 *    on the calling thread.
 *  - Calls that take a stream are asynchronous on it; calls that return host values
 *    (dot, norm, solve) synchronise that stream before returning, as TNL's do.
 */
#ifndef TNLHIP_H
#define TNLHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */

typedef enum
{
   TNLHIP_SUCCESS = 0,
   TNLHIP_ERROR_INVALID_ARGUMENT = 1, /* TNL_ASSERT_* / std::invalid_argument */
   TNLHIP_ERROR_RUNTIME = 2,          /* Exceptions::CudaRuntimeError (TNL_CHECK_CUDA_DEVICE) */
   TNLHIP_ERROR_OUT_OF_MEMORY = 3,    /* std::bad_alloc from Allocators::Cuda */
   TNLHIP_ERROR_CAPACITY = 4,         /* std::logic_error on row capacity overflow, SparseMatrixView.hpp:304-312 */
   TNLHIP_ERROR_UNSUPPORTED = 5,
   TNLHIP_ERROR_BREAKDOWN = 6         /* GMRES zero Hessenberg diagonal: `throw 1`, Solvers/Linear/GMRES.hpp:648-660 */
} tnlhip_status;

const char* tnlhip_status_string( int status );
const char* tnlhip_last_error( void );

/* Library version, device name and compute-unit count of the current HIP device. */
int tnlhip_device_info( char* name, int name_len, int* compute_units, int64_t* global_memory );

/* ------------------------------------------------------------------ sparse matrix */

/* Segments of SparseMatrix<..., Segments> and, for CSR, its kernel
 * (Algorithms/Segments/CSR.h:577-605, Ellpack.h, SlicedEllpack.h, ChunkedEllpack.h, BiEllpack.h).
 * All padded formats use TNL's GPU default ElementsOrganization, column-major
 * (Algorithms/Segments/ElementsOrganization.h:21-32). */
typedef enum
{
   TNLHIP_CSR_SCALAR = 0,   /* CSRScalar: one lane per row, sequential left-to-right sum */
   TNLHIP_CSR_VECTOR = 1,   /* CSRVector: one wavefront (64 lanes) per row */
   TNLHIP_CSR_LIGHT = 2,    /* CSRLight: 2..64 lanes per row chosen from nnz / rows */
   TNLHIP_CSR_HYBRID = 3,   /* CSRHybrid: 2^ceil(log2(nnz/rows)) lanes per row, at most 64 */
   TNLHIP_CSR_ADAPTIVE = 4, /* CSRAdaptive: STREAM / VECTOR / LONG blocks */
   TNLHIP_ELLPACK = 5,      /* Ellpack, alignment 32 */
   TNLHIP_SLICED_ELLPACK = 6, /* SlicedEllpack, slice size 32 */
   TNLHIP_CHUNKED_ELLPACK = 7, /* ChunkedEllpack, 256 chunks per slice, desired chunk size 16 */
   TNLHIP_BIELLPACK = 8     /* BiEllpack, strips of 32 rows */
} tnlhip_format;

typedef enum
{
   TNLHIP_F64 = 0,
   TNLHIP_F32 = 1,
   TNLHIP_BF16 = 2, /* values stored as bfloat16, ComputeReal = float (SparseMatrix.h:40-49) */
   TNLHIP_BINARY = 3, /* Real = bool: no values stored, every stored entry is 1; ComputeReal = Index
                        (SparseMatrixView.h:20-29), so vectors are int32 and a product sums x over
                        the stored columns in int32 */
   TNLHIP_INT32 = 4   /* int32 output vectors of the row reductions (not a matrix value type) */
} tnlhip_dtype;

typedef struct tnlhip_matrix_s* tnlhip_matrix;

typedef struct
{
   int32_t rows;
   int32_t columns;
   int32_t format;
   int32_t value_dtype;
   int64_t storage_size;  /* Segments::getStorageSize() as TNL reports it */
   int64_t allocated;     /* elements held in HBM (BiEllpack: groupPointers[last], not 32x) */
   int64_t nonzeros;      /* getNonzeroElementsCount() */
   int32_t lanes_per_row; /* CSR light / hybrid / vector kernel parameter (0 otherwise) */
   int32_t symmetric;     /* 1 for a SymmetricMatrix (the lower triangle is stored) */
} tnlhip_matrix_info;

/* SparseMatrix<Real, Devices::Hip, int, GeneralMatrix, Segments> A;  A = csr;
 *   Matrices/SparseMatrix.hpp:1070-1180 (operator= from a CSR matrix): row capacities are
 *   the source's getRowCapacities, i.e. its stored (column != -1) entries per row, entries are
 *   copied in source order, explicit zeros are dropped and the remaining slots hold the
 *   padding index -1
 *   (SparseMatrix.hpp:1315-1318).  setRowCapacities, SparseMatrix.hpp:284-305.
 * row_pointers[rows+1], column_indexes[nnz], values[nnz] (float64) are host pointers when
 * host_pointers != 0, HIP device pointers otherwise. */
int tnlhip_matrix_create_csr( int format, int value_dtype, int32_t rows, int32_t columns, const int32_t* row_pointers,
                              const int32_t* column_indexes, const double* values, int host_pointers,
                              tnlhip_matrix* out );

/* A matrix built as TNL user code builds one element by element: setDimensions, setRowCapacities
 * with the CSR row lengths, then setElement of every stored entry (SparseMatrix.hpp:284-305,
 * SparseMatrixView.hpp:270-345) -- unlike tnlhip_matrix_create_csr (operator=), the capacities are
 * kept as given, slots with column -1 stay padding and explicit zeros are stored.  symmetric != 0
 * builds a SymmetricMatrix from its lower triangle.  Host arrays.  This is what the C++ front's
 * SparseMatrix< ..., Devices::Hip > hands over when it first moves to the device. */
int tnlhip_matrix_create_staged( int format, int value_dtype, int symmetric, int32_t rows, int32_t columns,
                                 const int32_t* row_pointers, const int32_t* column_indexes, const double* values,
                                 tnlhip_matrix* out );

/* SparseMatrix<Real, Devices::Hip, int, SymmetricMatrix, Segments> A;  A = csr: the lower
 *   triangle, diagonal included, of a square matrix (an entry above the diagonal is an error).
 *   Products scale y by beta when beta != 1, then add alpha a_rc x_r to y_c for every stored
 *   a_rc with c < r and the row sums alpha sum_c a_rc x_c to y_r, with floating-point atomics
 *   (SparseMatrixView.hpp:390-491): y matches TNL's to rounding, not bit for bit.  Conversions,
 *   solvers (with Jacobi), save / load keep the matrix type; ILU(0) refuses it. */
int tnlhip_matrix_create_symmetric_csr( int format, int value_dtype, int32_t rows, const int32_t* row_pointers,
                                        const int32_t* column_indexes, const double* values, int host_pointers,
                                        tnlhip_matrix* out );

/* B = A across formats / value types on the device (SparseMatrix.hpp:1070-1107): B's capacities are
 * A's getRowCapacities (stored entries per row, not A's segment sizes), explicit zeros dropped. */
int tnlhip_matrix_convert( tnlhip_matrix source, int format, int value_dtype, tnlhip_matrix* out );

/* Matrix Market files (Matrices/MatrixReader.hpp, MatrixWriter.hpp:105-130).
 * tnlhip_mtx_read_csr: MatrixReader::readMtx into a host CSR matrix -- header "%%MatrixMarket
 *   matrix coordinate real general|symmetric", row capacities counted from the file (padding
 *   column -1 where entries repeat), setElement semantics (sorted, last repeated value kept,
 *   explicit zeros stored).  Pass NULL arrays to query rows / columns / storage first.
 * tnlhip_mtx_read: the same into a matrix of any format on the device (readMtx for a
 *   Devices::Hip matrix, MatrixReader.hpp:19-26: read on the host, then matrix = hostMatrix, which
 *   takes the stored entries per row as capacities and drops explicit zeros).
 * tnlhip_mtx_write: MatrixWriter::writeMtx of a device matrix (MatrixWriter.hpp:35-40, 105-130: a
 *   host copy by operator=, then the entries whose value is not zero, 6 digits). */
int tnlhip_mtx_read_csr( const char* path, int32_t* rows, int32_t* columns, int64_t* storage, int32_t* row_pointers,
                         int32_t* column_indexes, double* values );
int tnlhip_mtx_read( const char* path, int format, int value_dtype, tnlhip_matrix* out );
int tnlhip_mtx_write( tnlhip_matrix A, const char* path );
/* writeMtx of a host CSR matrix (padding column -1 allowed). */
int tnlhip_mtx_write_csr( const char* path, int32_t rows, int32_t columns, const int32_t* row_pointers,
                          const int32_t* column_indexes, const double* values );

/* TNL binary files (File / Object format, Object.hpp:17-75): "TNLMN", the serialization type
 * string, then SparseMatrix::save -- rows, columns, the values and column-index arrays and the
 * segments' own fields (Matrix.hpp:129-150, SparseMatrixView.hpp:911-920; CSRView.hpp:221-224,
 * EllpackView.hpp:404-410, SlicedEllpackView.hpp:310-318, ChunkedEllpackView.hpp:369-378,
 * BiEllpackView.hpp:345-351).  The element order of the Ellpack family depends on the
 * organization TNL was built with, which the file does not record, hence the layout:
 *   TNLHIP_FILE_CUDA_LAYOUT  column-major, as TNL's Devices::Cuda matrices write it
 *                            (CSR kernel names carry TNL::Devices::Cuda)
 *   TNLHIP_FILE_HOST_LAYOUT  row-major, as TNL's Devices::Host matrices write it (Ellpack's
 *                            alignedSize = rows; CSR kernel names carry TNL::Devices::Host) */
typedef enum
{
   TNLHIP_FILE_CUDA_LAYOUT = 0,
   TNLHIP_FILE_HOST_LAYOUT = 1
} tnlhip_file_layout;

typedef struct
{
   int32_t format;      /* TNLHIP_CSR_* from the CSR kernel name, or the Ellpack-family format */
   int32_t value_dtype; /* TNLHIP_F64 ("double") or TNLHIP_F32 ("float") */
   int32_t rows;
   int32_t columns;
   int64_t storage_size; /* length of the stored values array */
   int32_t symmetric;    /* 1 when the type names a Symmetric matrix */
   int32_t reserved;
} tnlhip_file_info;

/* SparseMatrix::save( fileName ) (SparseMatrix.hpp:1266-1270).  f64 and f32 matrices only: TNL
 * has no bfloat16 Real. */
int tnlhip_matrix_save( tnlhip_matrix A, const char* path, int layout );
/* SparseMatrix::load( fileName ) (SparseMatrix.hpp:1248-1255, 1282-1286).  format and value_dtype
 * -1 take them from the file; otherwise the file's type must match them, where any CSR kernel
 * name matches every CSR format (the stored arrays do not depend on the kernel).  The metadata
 * is checked to be the one TNL builds (every slot in range, every column index in [-1,
 * columns)) before anything reaches the device. */
int tnlhip_matrix_load( const char* path, int format, int value_dtype, int layout, tnlhip_matrix* out );
/* getObjectType( fileName ) (Object.hpp:64-84) and the checks of tnlhip_matrix_load, on the host
 * only; type (may be NULL) receives the serialization type string. */
int tnlhip_file_inspect( const char* path, int layout, tnlhip_file_info* info, char* type, int type_len );

int tnlhip_matrix_destroy( tnlhip_matrix A );

int tnlhip_matrix_get_info( tnlhip_matrix A, tnlhip_matrix_info* info );

/* Copies one segments metadata array to the host; returns its length (or -1 for an
 * unknown name), out may be NULL to query the length.  Names follow TNL's members:
 *   CSR:            offsets                                   (detail/CSR.h:22-34)
 *   Ellpack:        ellpack = {segmentSize, size, alignedSize} (Ellpack.hpp:72-83)
 *   SlicedEllpack:  sliceOffsets, sliceSegmentSizes, sliced = {size, alignedSize}
 *   ChunkedEllpack: rowToChunkMapping, rowToSliceMapping, chunksToSegmentsMapping,
 *                   rowPointers, slices (4 ints per slice: size, chunkSize,
 *                   firstSegment, pointer), chunked = {numberOfSlices, storageSize}
 *   BiEllpack:      rowPermArray, groupPointers, biellpack = {virtualRows, storage lo, hi}
 *   CSRAdaptive:    adaptiveBlocks (6 ints per block: type, firstSegment, blockSize,
 *                   segmentsInBlock, warpIdx, warpsCount) */
int64_t tnlhip_matrix_get_metadata( tnlhip_matrix A, const char* name, int32_t* out );

/* Host-only: the segments metadata `format` builds for n segments of the given capacities,
 * without touching the GPU (Segments::setSegmentsSizes on Devices::Host; e.g.
 * ChunkedEllpack.hpp:194-235, BiEllpack.hpp:246-279).  Same names and return convention as
 * tnlhip_matrix_get_metadata, plus "storage" = {lo, hi} of getStorageSize(). */
int64_t tnlhip_segments_metadata( int format, int32_t n, const int32_t* capacities, const char* name, int32_t* out );

/* Copies the allocated storage (values as float64, column indexes) to the host. */
int tnlhip_matrix_get_storage( tnlhip_matrix A, double* values, int32_t* columns );

/* Diagonal a_ii of every row (SparseMatrixView::getElement(i, i)), device pointer
 * of the vector type. */
int tnlhip_matrix_get_diagonal( tnlhip_matrix A, void* diagonal, void* stream );

/* Overrides a kernel parameter chosen at construction (ChunkedEllpack and SlicedEllpack with
 * FP32 / BF16 values start with "xcd_remap" 1, FP64 / FP32 Ellpack with "xcd_remap" 3 and
 * "blocks_per_cu" 32, FP64 SlicedEllpack with "pair_rows" 1 and "blocks_per_cu" 32, every other
 * matrix with "xcd_remap" 0, "pair_rows" 0, "blocks_per_cu" 16): "lanes_per_row" (CSR light / hybrid /
 * vector: a power of two in [1, 64]; once set, that many lanes share each row even when rows
 * average at most 8 stored elements, where these kernels otherwise take CSRScalar's staged
 * one-lane-per-row path); "xcd_remap" (row distribution: 0 the plain order, 1 contiguous row
 * blocks -- slices for ChunkedEllpack -- with each XCD owning a contiguous eighth; 3 per-XCD
 * grid-stride, Ellpack / SlicedEllpack only); for the Ellpack / SlicedEllpack kernels
 * "pair_rows" (0/1: two adjacent rows per lane with 8 / 16-byte loads; SlicedEllpack with slice
 * size 1 always takes one row per lane) and "blocks_per_cu"
 * (workgroups per CU the grid is sized for); "stream_y" (-1 the kernel's default, 0 stores of y
 * through L2, 1 non-temporal stores); "column_stripes" (CSR: 0 drops the column-striped plan,
 * 2 / 4 / 8 / 16 / 32 builds one, a second, stripe-major copy of the entries whose stripes of x
 * stay in the L2 of the XCDs that own them; sums reassociated by stripe, to 1e-13 of sum |a x|);
 * "slice_size" (SlicedEllpack: TNL's SliceSize template parameter, a power of two in [1, 256],
 * default 32; the storage is re-laid out on the device with capacities = stored entries per row).
 * Every other name or value fails with TNLHIP_ERROR_INVALID_ARGUMENT. */
int tnlhip_matrix_set_param( tnlhip_matrix A, const char* name, int32_t value );

/* SparseMatrix::vectorProduct( x, y, alpha, beta, firstRow, lastRow ):
 *   y[first, last) = alpha * A x + beta * y; last == 0 means all rows; beta == 0 does not
 *   read y.  x may be NULL when A has no columns (nothing is read from it), y when A has no
 *   rows.  Matrices/SparseMatrix.h:1078-1085, SparseMatrixView.hpp:390-491. */
int tnlhip_spmv( tnlhip_matrix A, const void* x, void* y, double alpha, double beta, int32_t first, int32_t last,
                 void* stream );

/* SparseMatrix::vectorProduct generalized to k right-hand sides read with one pass over A (no TNL
 * entry point takes several vectors; SparseMatrixView.hpp:390-491 per column):
 *   Y[first, last) = alpha * A X + beta * Y, X columns x k and Y rows x k row-major device arrays
 *   (entry row * k + j), 1 <= k <= 64; last == 0 means all rows; beta == 0 does not read Y.
 *   Column j of Y equals the one-vector product of column j of X in TNL's Host order (each product
 *   rounded, the row's slots summed left to right), bit for bit.  Any format; general real-valued
 *   matrices (TNLHIP_ERROR_UNSUPPORTED for symmetric and binary ones).  X may be NULL when A has
 *   no columns, Y when it has no rows.  X and Y 16-byte aligned with k a multiple of 2 (FP64) or
 *   4 (FP32 / BF16 values) take the 16-byte-lane kernel; any other layout works, more slowly. */
int tnlhip_spmv_multi( tnlhip_matrix A, int32_t k, const void* X, void* Y, double alpha, double beta, int32_t first,
                       int32_t last, void* stream );

/* ------------------------------------------------------------------ row traversal */

/* Fetch and reduction of tnlhip_matrix_reduce_rows: the fetch( rowIdx, columnIdx, value ) and
 * reduce( a, b ) lambdas of SparseMatrixView::reduceRows that C callers can name. */
typedef enum
{
   TNLHIP_FETCH_VALUE = 0,     /* value */
   TNLHIP_FETCH_ABS_VALUE = 1, /* |value| */
   TNLHIP_FETCH_ONE = 2,       /* 1 for every stored entry (getRowCapacities' fetch) */
   TNLHIP_FETCH_NONZERO = 3,   /* value != 0 (getCompressedRowLengths' fetch) */
   TNLHIP_FETCH_COLUMN = 4     /* the column index */
} tnlhip_fetch;

typedef enum
{
   TNLHIP_REDUCE_SUM = 0, /* std::plus */
   TNLHIP_REDUCE_MIN = 1, /* TNL::min */
   TNLHIP_REDUCE_MAX = 2  /* TNL::max */
} tnlhip_reduce;

/* SparseMatrixView::reduceRows( first, last, fetch, reduce, keep, identity )
 * (Matrices/SparseMatrixView.hpp:495-595; Segments' reduceSegments, e.g. CSRView.hpp:185-199):
 * out[row] = identity (+) fetch(a_row,0) (+) fetch(a_row,1) ... over the row's slots in local order,
 * a padding slot (column -1) fetching the identity; keep stores into out (a device vector of
 * out_dtype: TNLHIP_F64, TNLHIP_F32 or TNLHIP_INT32).  last == 0 means all rows.  Each row is
 * folded left to right as TNL's Host traversal does, so every result is TNL's bit for bit. */
int tnlhip_matrix_reduce_rows( tnlhip_matrix A, int32_t first, int32_t last, int fetch, int reduce, double identity,
                               void* out, int out_dtype, void* stream );

/* SparseMatrix::getRowCapacities( vector ) (SparseMatrixView.hpp:143-160): reduceAllRows with fetch 1,
 * i.e. the number of stored (non-padding) entries of every row; int32 device vector of rows. */
int tnlhip_matrix_get_row_capacities( tnlhip_matrix A, int32_t* capacities, void* stream );

/* SparseMatrix::getCompressedRowLengths( vector ) (SparseMatrixView.hpp:116-135): the entries of
 * every row whose value is not zero; int32 device vector of rows. */
int tnlhip_matrix_get_compressed_row_lengths( tnlhip_matrix A, int32_t* lengths, void* stream );

/* SparseMatrix::getRowCapacity( row ) (SparseMatrixView.hpp:163-172): the segment size of the row. */
int tnlhip_matrix_get_row_capacity( tnlhip_matrix A, int32_t row, int32_t* capacity );

/* SparseMatrix::getRow( row ) read on the host (SparseMatrixView.hpp:222-250, SparseMatrixRowView):
 * *size = the segment size; columns[l] / values[l] receive getColumnIndex( l ) / getValue( l ) for
 * l < size in local order (padding slots: column -1, value 0; binary matrices: 1 per stored entry).
 * columns / values are host arrays of at least `capacity` entries; pass NULL for both to query the size. */
int tnlhip_matrix_get_row( tnlhip_matrix A, int32_t row, int32_t* columns, double* values, int32_t capacity,
                           int32_t* size );

/* Device arrays of a matrix, for traversal kernels compiled by the caller (the C++ front's
 * forElements / forRows / reduceRows with device lambdas, SparseMatrixView.hpp:495-760).  The
 * index maps are TNL's column-major ones (the segment views of Algorithms/Segments): CSR
 * offsets[r] + l; Ellpack r + aligned_size * l; SlicedEllpack slice_offsets[r / 32] + r % 32 + 32 l;
 * ChunkedEllpack and BiEllpack as detail/ChunkedEllpack.h:110-145 and detail/BiEllpack.h:150-185.
 * The pointers stay valid until the matrix is destroyed; a caller that writes values or column
 * indexes through them calls tnlhip_matrix_values_changed afterwards. */
typedef struct
{
   int32_t format, value_dtype, rows, columns;
   void* values;             /* NULL for binary matrices */
   int32_t* column_indexes;
   const int32_t* offsets;   /* CSR */
   int32_t segment_size, aligned_size; /* Ellpack */
   const int32_t* slice_offsets;       /* SlicedEllpack */
   const int32_t* slice_segment_sizes;
   const int32_t* row_to_chunk;        /* ChunkedEllpack: rowToChunkMapping, rowToSliceMapping, */
   const int32_t* row_to_slice;        /* slices (size, chunkSize, firstSegment, pointer) */
   const int32_t* slices;
   const int32_t* row_perm;            /* BiEllpack: rowPermArray, groupPointers */
   const int32_t* group_pointers;
   int32_t symmetric;
   int32_t reserved;
} tnlhip_matrix_view;

int tnlhip_matrix_get_view( tnlhip_matrix A, tnlhip_matrix_view* view );

/* After values or column indexes were written through the view (forElements): waits for the
 * stream, then refreshes what the library derived from them (cached diagonal, column-striped
 * plan, nonzero count). */
int tnlhip_matrix_values_changed( tnlhip_matrix A, void* stream );

/* ------------------------------------------------------------------ reductions / BLAS-1 */

/* (a, b) = Algorithms::reduce( a * b, Plus ), Containers/Expressions/ExpressionTemplates.h:393-403.
 * Deterministic: fixed partition and fixed combining order. */
int tnlhip_dot( int dtype, int64_t n, const void* a, const void* b, double* result, void* stream );

/* lpNorm( a, 2.0 ) = sqrt( (a, a) ), ExpressionTemplates.h:458-478. */
int tnlhip_norm2( int dtype, int64_t n, const void* a, double* result, void* stream );

/* y = alpha * x + beta * y (VectorAssignment over ParallelFor, Containers/detail/VectorAssignment.h). */
int tnlhip_axpby( int dtype, int64_t n, double alpha, const void* x, double beta, void* y, void* stream );

/* result[k] = sum_i V[i + k*ld] * w[i], k < count: Algorithms::Multireduction::reduce with the
 * GMRES fetch (Algorithms/Multireduction.hpp:30-233, Solvers/Linear/GMRES.hpp:218-230). */
int tnlhip_multidot( int dtype, int64_t n, int32_t count, const void* V, int64_t ld, const void* w, double* result,
                     void* stream );

/* tnlhip_multidot without allocating: no counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same kernels and bits as tnlhip_multidot, but count is limited to 1..32 columns (one
 * pass; larger counts fail with TNLHIP_ERROR_INVALID_ARGUMENT), `result` (count doubles) is DEVICE
 * memory, and the call is asynchronous on `stream` (no synchronisation, nothing copied to the
 * host).  The partial sums live in a fixed 256 KiB device array that each GPU holds until the
 * process ends and that every call on that GPU shares: calls that overlap on one GPU (different
 * streams or threads) race and must be serialised by the caller. */
int tnlhip_multidot_noalloc( int dtype, int64_t n, int32_t count, const void* V, int64_t ld, const void* w,
                             double* result, void* stream );

/* y = alpha * V x + beta * y with V column-major n x count, x on the host:
 * Matrices::MatrixOperations::gemv, Matrices/MatrixOperations.h:46-150, 232-345. */
int tnlhip_gemv( int dtype, int64_t n, int32_t count, double alpha, const void* V, int64_t ld, const double* x,
                 double beta, void* y, void* stream );

/* In-place exclusive prefix sum of int32 (Algorithms::inplaceExclusiveScan, Algorithms/scan.h);
 * out[n] receives the total when out has n+1 elements (total_out != 0). */
int tnlhip_exclusive_scan_i32( int64_t n, const int32_t* in, int32_t* out, int total_out, void* stream );

/* ------------------------------------------------------------------ Krylov solvers */

typedef enum
{
   TNLHIP_CG = 0,       /* Solvers/Linear/CG.hpp */
   TNLHIP_BICGSTAB = 1, /* Solvers/Linear/BICGStab.hpp */
   TNLHIP_GMRES = 2,    /* Solvers/Linear/GMRES.hpp */
   TNLHIP_TFQMR = 3,    /* Solvers/Linear/TFQMR.hpp */
   TNLHIP_BICGSTABL = 4 /* Solvers/Linear/BICGStabL.hpp: BiCGStab(ell) */
} tnlhip_method;

typedef enum
{
   TNLHIP_GMRES_CGS = 0,
   TNLHIP_GMRES_CGSR = 1,
   TNLHIP_GMRES_MGS = 2,
   TNLHIP_GMRES_MGSR = 3,
   TNLHIP_GMRES_CWY = 4
} tnlhip_gmres_variant;

/* IterativeSolver / solver setup parameters; defaults as IterativeSolver.h:226-245 and
 * GMRES.h:220-237 (CWY, restarting 10/10, steps 3/3). */
typedef struct
{
   int32_t max_iterations;
   int32_t min_iterations;
   double convergence_residue;
   double divergence_residue;
   int32_t gmres_variant;
   int32_t restarting_min;
   int32_t restarting_max;
   int32_t restarting_step_min;
   int32_t restarting_step_max;
   int32_t bicgstab_exact_residue;
   int32_t preconditioner; /* 0 none, 1 Preconditioners::Diagonal (Jacobi), 2 Preconditioners::ILU0,
                              3 Jacobi polynomial (no TNL counterpart; see poly_degree) */
   int32_t use_graph;      /* 1: CG / BiCGStab iterations are captured once into a hipGraph and
                              replayed (16 per replay); 0: launched one by one */
   int32_t bicgstab_ell;   /* BiCGStab(ell): Bi-CG steps per minimal-residual update, 1..16
                              ("bicgstab-ell", default 1) */
   int32_t poly_degree;    /* preconditioner 3: z = p(D^-1 P) D^-1 s, the Neumann series of P's
                              Jacobi splitting cut after poly_degree products with P (z_0 = D^-1 s,
                              z_{j+1} = z_j + D^-1 (s - P z_j)), D the diagonal of the solved matrix,
                              P its preconditioner matrix (tnlhip_solver_set_preconditioner_matrix;
                              the solved matrix itself by default).  0..32, default 3: with k
                              products the preconditioned operator is I - (I - D^-1 P)^(k+1), and
                              an odd k (even power) keeps its eigenvalues inside (0, 1] for a
                              splitting whose Jacobi matrix has spectrum in (-1, 1); an even k lets
                              them reach 2 (twice the condition number). */
} tnlhip_solver_params;

void tnlhip_solver_params_default( tnlhip_solver_params* params );

typedef struct tnlhip_solver_s* tnlhip_solver;

int tnlhip_solver_create( int method, tnlhip_solver* out );
int tnlhip_solver_destroy( tnlhip_solver solver );

/* LinearSolver::setMatrix( shared_ptr<const Matrix> ), Solvers/Linear/LinearSolver.h:115-131.
 * The matrix must outlive the solver's use of it. */
int tnlhip_solver_set_matrix( tnlhip_solver solver, tnlhip_matrix A );

/* setup( parameters ) / setMaxIterations, setConvergenceResidue, ...; when
 * params->preconditioner == 1 this is setPreconditioner( Diagonal ) followed by its
 * update( matrix ) (Preconditioners/Diagonal.hpp:22-41); with 2, setPreconditioner( ILU0 ) and
 * its update (Preconditioners/ILU0.hpp:17-117: factors with A's sparsity, TNL's Host arithmetic
 * bit for bit), done at the next solve.  ILU(0) needs every row to store its columns below the
 * diagonal before the others and an entry on or above the diagonal (TNLHIP_ERROR_UNSUPPORTED /
 * INVALID_ARGUMENT otherwise). */
int tnlhip_solver_set_params( tnlhip_solver solver, const tnlhip_solver_params* params );

/* The matrix P of the Jacobi-polynomial preconditioner (params->preconditioner == 3); NULL: the
 * solved matrix.  No counterpart in the reference.  P has the solved matrix's dimensions and
 * vector type (FP64 values with FP64 vectors; FP32 or BF16 values with FP32 vectors), is general
 * (not symmetric or binary) and must outlive the solver's use of it.  A BF16 copy of an FP32
 * system makes the preconditioner's products, which carry most of the bytes, read 2-byte values
 * while the Krylov recurrence runs on the FP32 operator. */
int tnlhip_solver_set_preconditioner_matrix( tnlhip_solver solver, tnlhip_matrix P );

/* bool LinearSolver::solve( b, x ), x holds the initial guess on entry; *converged is
 * checkConvergence() (IterativeSolver.hpp:127-150).  Synchronous. */
int tnlhip_solver_solve( tnlhip_solver solver, const void* b, void* x, int32_t* converged, void* stream );

/* getIterations(), getResidue() (IterativeSolver.h:70-200). */
int tnlhip_solver_get_status( tnlhip_solver solver, int32_t* iterations, double* residue );

/* Every value passed to setResidue during the last solve, in order (the residual history
 * file of IterativeSolver.hpp:189-200); returns the count, out may be NULL, -1 for a NULL solver or
 * a negative capacity. */
int64_t tnlhip_solver_get_history( tnlhip_solver solver, double* out, int64_t capacity );

/* getIterations() at each of those setResidue calls (the first field of a history-file line). */
int64_t tnlhip_solver_get_history_iterations( tnlhip_solver solver, int32_t* out, int64_t capacity );

/* IterativeSolver::setSolverMonitor (Solvers/IterativeSolver.h:177-188, IterativeSolverMonitor):
 * monitor( user, iterations, residue ) is called on the calling thread each time the solve reads
 * its device state (after each replayed batch of iterations, each GMRES restart cycle, and at the
 * end); NULL removes it.  The refresh rate of TNL's monitor thread belongs to the monitor, which
 * the C++ front runs. */
typedef void ( *tnlhip_monitor_fn )( void* user, int32_t iterations, double residue );
int tnlhip_solver_set_monitor( tnlhip_solver solver, tnlhip_monitor_fn monitor, void* user );

/* setup( "residual-history-file" ) (IterativeSolver.hpp:57-61, 189-200): opens the file (NULL or ""
 * closes it); each solve appends its setResidue calls as TNL writes them, a blank line at
 * iteration 0, then "<iteration>\t<residue>" in std::scientific. */
int tnlhip_solver_set_residual_history_file( tnlhip_solver solver, const char* path );

/* ------------------------------------------------------------------ mixed-precision refinement */

/* FP64 iterative refinement around a low-precision inner Krylov solve.  No counterpart in the
 * reference (TNL's solvers run in one precision); the inner solver is the CG / BiCGStab /
 * GMRES of this library.  Outer loop: r = b - A x in FP64, inner solve of A_low d = r/|r| in
 * FP32 vectors (A_low values FP32 or BF16) to inner.convergence_residue, x += |r| d in FP64,
 * until |r| / |b| < tolerance. */
#define TNLHIP_REFINEMENT_HISTORY 64

typedef struct
{
   double tolerance;             /* outer FP64 relative residual target */
   int32_t max_outer_iterations;
   int32_t inner_method;         /* tnlhip_method */
   int32_t inner_value_dtype;    /* TNLHIP_F32 or TNLHIP_BF16 */
   int32_t stagnation_limit;     /* stop (not converged) after this many consecutive outer steps that
                                    do not halve the residual; 0 = never.  Refinement converges only
                                    when cond(A) times the unit roundoff of A_low is below 1. */
   tnlhip_solver_params inner;   /* inner stopping rule, GMRES variant / restart, preconditioner */
   int32_t poly_value_dtype;     /* inner.preconditioner == 3: value type of the preconditioner
                                    matrix P, a copy of A in A_low's format (TNLHIP_BF16, the default,
                                    or TNLHIP_F32, which reuses A_low when that is FP32).  With FP32
                                    A_low and BF16 P the inner Krylov space is built on the FP32
                                    operator, so the BF16 rounding of A limits only how well P
                                    preconditions, not the accuracy refinement reaches. */
   tnlhip_matrix poly_matrix;    /* that P, already built (NULL: converted from A on each call) */
} tnlhip_refinement_params;

typedef struct
{
   int32_t converged;
   int32_t outer_iterations;
   int64_t inner_iterations;     /* summed over the outer iterations */
   double residue;               /* final FP64 |b - A x| / |b| */
   double history[ TNLHIP_REFINEMENT_HISTORY ]; /* FP64 residue before each outer step */
} tnlhip_refinement_result;

/* defaults: tolerance 1e-10, 50 outer steps, FP32 Jacobi GMRES(30) CWY inner solve to 1e-2,
 * stagnation limit 2 */
void tnlhip_refinement_params_default( tnlhip_refinement_params* params );

/* A: FP64 matrix; A_low: the same matrix in FP32 / BF16 (NULL: converted internally from A,
 * same format); b, x: FP64 device vectors, x holds the initial guess.  Synchronous. */
int tnlhip_refine_solve( tnlhip_matrix A, tnlhip_matrix A_low, const tnlhip_refinement_params* params,
                         const double* b, double* x, tnlhip_refinement_result* result, void* stream );

/* A refinement context keeps the outer loop's vectors and the inner solver (its Krylov basis,
 * reductions scratch) between solves, so repeated solves allocate and free nothing after the
 * first one of the largest size (tnlhip_refine_solve allocates and frees them on every call, and
 * a device free waits for the whole device).  No TNL counterpart.  A context is used by one host
 * thread at a time, on the device it was first used on; it holds its buffers until destroyed. */
typedef struct tnlhip_refine_context_s* tnlhip_refine_context;
int tnlhip_refine_context_create( tnlhip_refine_context* out );
int tnlhip_refine_context_destroy( tnlhip_refine_context ctx );
/* tnlhip_refine_solve on the context's buffers (grown as needed) */
int tnlhip_refine_solve_in( tnlhip_refine_context ctx, tnlhip_matrix A, tnlhip_matrix A_low,
                            const tnlhip_refinement_params* params, const double* b, double* x,
                            tnlhip_refinement_result* result, void* stream );

/* ------------------------------------------------------------------ autotuner */

/* One timed configuration: format, lanes per row (CSR light / hybrid; 0 = the kernel's
 * choice), workgroups per CU and rows per lane (padded formats; 0 / -1 = defaults), the slice
 * size of SlicedEllpack (TNL's SliceSize template parameter: 32 is TNL's default and keeps its
 * metadata; 16 and 64 are tried too, 0 for other formats), the column stripes of CSRAdaptive's
 * striped plan (0, and 4 and 8 when x exceeds 4 MiB, one XCD's L2), seconds per y = A x. */
typedef struct
{
   int32_t format;
   int32_t lanes;
   int32_t blocks_per_cu;
   int32_t pair_rows;
   int32_t slice_size;
   int32_t column_stripes;
   double seconds;
} tnlhip_tuning;

/* Build the CSR matrix (arguments as tnlhip_matrix_create_csr) in every format whose bit is set
 * in format_mask (bit = tnlhip_format value), time `repetitions` products y = A x for each
 * candidate launch configuration, and report the fastest in *best (all timings in `all`, up to
 * all_capacity; *all_count gets the number timed).  With best_matrix non-NULL the fastest
 * matrix is returned, configured, and owned by the caller.  No counterpart in the reference
 * (the format is TNL's Segments template parameter, SparseMatrix.h:40-49). */
int tnlhip_autotune_csr( int value_dtype, int32_t rows, int32_t columns, const int32_t* row_pointers,
                         const int32_t* column_indexes, const double* values, int32_t host_pointers,
                         uint32_t format_mask, int32_t repetitions, tnlhip_tuning* best, tnlhip_tuning* all,
                         int32_t all_capacity, int32_t* all_count, tnlhip_matrix* best_matrix );

/* ------------------------------------------------------------------ solve server (batched) */

/* Many uneven systems with 1-16 right-hand sides each, resident together and advanced by one
 * fixed sequence of persistent kernels (replayed as a hipGraph): continuous batching with
 * admission by HBM footprint.  Each request solves (A + shift I) X = B for a resident
 * matrix A to an FP64 relative residual per right-hand side, by FP64 iterative refinement
 * over an FP32 CG inner solve (Solvers/Linear/CG.hpp:17-107 recurrences).  GMRES requests
 * (method TNLHIP_GMRES, Solvers/Linear/GMRES.hpp:79-180) run beside them: `gmres_workers` of them at
 * once, each on its own HIP stream, by FP64 refinement over an FP32 GMRES(30) of (A + shift I)
 * preconditioned by the degree-4 Jacobi polynomial over a BF16 copy, admitted FIFO as workers
 * free up.  No counterpart in the reference: TNL solves one
 * system per LinearSolver::solve call (LinearSolver.h:115-162); a TNL caller that loops over
 * solve() for many small systems is what this replaces. */
typedef struct tnlhip_server_s* tnlhip_server;

typedef struct
{
   int64_t matrix_capacity;        /* SlicedEllpack slots of all resident matrices (16 B each) */
   int64_t slice_capacity;         /* slice offsets of all resident matrices (4 B each) */
   int64_t vector_capacity;        /* (row, right-hand side) entries of all resident requests (40 B each) */
   int64_t staging_capacity;       /* largest nnz + rows + 1 of a matrix upload */
   int32_t max_requests;           /* resident requests at once */
   int32_t iterations_per_launch;  /* iterations per replay (one host synchronization each) */
   int32_t use_graph;              /* 1: capture the replay in a hipGraph */
   int32_t gmres_workers;          /* GMRES requests solved at once, each by its own host thread and
                                      HIP stream (0: GMRES requests are refused) */
} tnlhip_server_config;

typedef struct
{
   double tolerance;               /* FP64 |b - (A + shift I) x| / |b| per right-hand side */
   double inner_tolerance;         /* FP32 inner CG relative residual */
   int32_t max_outer_iterations;
   int32_t max_inner_iterations;   /* per outer iteration */
   int32_t method;                 /* TNLHIP_CG (batched track) or TNLHIP_GMRES (worker track) */
   int32_t stagnation_limit;       /* as tnlhip_refinement_params */
} tnlhip_request_params;

typedef struct
{
   int32_t converged;              /* every right-hand side reached the tolerance */
   int32_t outer_iterations;       /* max over the right-hand sides */
   int64_t inner_iterations;       /* summed over the right-hand sides */
   double residue;                 /* max over the right-hand sides */
   int64_t launches;               /* server replays completed when the request finished */
} tnlhip_request_result;

typedef struct
{
   int32_t resident_requests, queued_requests;
   int64_t resident_entries, vector_free, matrix_free, launches, tiles;
} tnlhip_server_stats_t;

void tnlhip_server_config_default( tnlhip_server_config* config );
void tnlhip_request_params_default( tnlhip_request_params* params );
/* HBM bytes a matrix (rows, nnz, longest row) and a request (rows, nrhs) occupy. */
int tnlhip_server_footprint( int32_t rows, int64_t nnz, int32_t max_row, int32_t nrhs, int64_t* matrix_bytes,
                             int64_t* request_bytes );
int tnlhip_server_create( const tnlhip_server_config* config, tnlhip_server* out );
int tnlhip_server_destroy( tnlhip_server server );
/* Upload a square CSR matrix (host arrays) and keep it resident; TNLHIP_ERROR_CAPACITY when the
 * matrix arena has no room.  Synchronous. */
int tnlhip_server_add_matrix( tnlhip_server server, int32_t rows, const int32_t* row_pointers, const int32_t* columns,
                              const double* values, int32_t* matrix_id );
/* The matrix is freed once no resident request uses it. */
int tnlhip_server_remove_matrix( tnlhip_server server, int32_t matrix_id );
/* Queue a request: b and x are host arrays of rows x nrhs (entry row * nrhs + j) that stay valid
 * until the ticket finishes; x receives the solution (the initial guess is 0). */
int tnlhip_server_submit( tnlhip_server server, int32_t matrix_id, double shift, int32_t nrhs, const double* b,
                          double* x, const tnlhip_request_params* params, int64_t* ticket );
/* Admit queued requests that fit (FIFO), run one replay, retire finished requests (their x is
 * written).  Up to `capacity` finished tickets are stored in `finished`; *count gets how many
 * finished. */
int tnlhip_server_progress( tnlhip_server server, int64_t* finished, int32_t capacity, int32_t* count );
/* Result of a finished ticket (once). */
int tnlhip_server_result( tnlhip_server server, int64_t ticket, tnlhip_request_result* result );
int tnlhip_server_stats( tnlhip_server server, tnlhip_server_stats_t* stats );

#ifdef __cplusplus
}
#endif

#endif /* TNLHIP_H */
