/*
 * CPU oracle for the SpMV + Krylov hot path — TEST INFRASTRUCTURE ONLY.
 *
 * This library restates, in plain single-threaded C++, what TNL's
 * Devices::Host path computes: segment metadata for every format, the
 * CSR -> format conversion of SparseMatrix::operator=, each format's Host
 * reduceSegments traversal order for vectorProduct, the sequential
 * 128-block / 4-accumulator reduction, and the CG / BiCGStab / GMRES
 * solvers with the Diagonal (Jacobi) preconditioner.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it, and only as a checker.  The product library never links it.
 *
 * Each function's reference file:line is given in oracle.cpp.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* segments kinds (independent of the CSR kernel flavour) */
enum { ORC_CSR = 0, ORC_ELLPACK = 1, ORC_SLICED_ELLPACK = 2, ORC_CHUNKED_ELLPACK = 3, ORC_BIELLPACK = 4 };

typedef struct orc_segments orc_segments;

/* Build the segments of `kind` for n segments of the given capacities.
 * column_major selects ElementsOrganization (TNL's GPU default is column-major). */
orc_segments* orc_segments_create( int32_t kind, int32_t n, const int32_t* capacities, int32_t column_major );
/* SlicedEllpack with SliceSize = slice_size (SlicedEllpack.h:22 template parameter; 32 by default) */
orc_segments* orc_segments_create_sliced( int32_t n, const int32_t* capacities, int32_t column_major,
                                          int32_t slice_size );
void orc_segments_destroy( orc_segments* s );
int64_t orc_segments_storage_size( const orc_segments* s );
int32_t orc_segments_segment_size( const orc_segments* s, int32_t segment );
int64_t orc_segments_global_index( const orc_segments* s, int32_t segment, int32_t local );
/* Copy a metadata array out by name; returns its element count (out may be NULL to query).
 * Names: offsets | ellpack (segmentSize,size,alignedSize) | sliceOffsets | sliceSegmentSizes |
 * sliced (size,alignedSize) | rowToChunkMapping | rowToSliceMapping | chunksToSegmentsMapping |
 * rowPointers | slices (size,chunkSize,firstSegment,pointer per slice) | chunked (numberOfSlices,storageSize) |
 * rowPermArray | groupPointers | biellpack (virtualRows, storageSize lo, storageSize hi) */
int64_t orc_segments_get( const orc_segments* s, const char* name, int32_t* out );

/* SparseMatrix::operator=( CSR source ) into these segments: entries in source order,
 * explicit zeros dropped, padding column = -1, padding value = 0. */
void orc_convert_from_csr( const orc_segments* s, const int32_t* row_ptr, const int32_t* cols, const double* vals,
                           double* out_vals, int32_t* out_cols );

/* vectorProduct following the format's Host reduceSegments traversal. */
void orc_spmv_f64( const orc_segments* s, const double* vals, const int32_t* cols, const double* x, double* y,
                   double alpha, double beta, int32_t first, int32_t last );
void orc_spmv_f32( const orc_segments* s, const float* vals, const int32_t* cols, const float* x, float* y,
                   float alpha, float beta, int32_t first, int32_t last );

/* SparseMatrixView::reduceRows over the segments' slots in local order: fetch 0 value, 1 |value|,
 * 2 one (getRowCapacities), 3 value != 0 (getCompressedRowLengths), 4 column; reduce 0 sum, 1 min,
 * 2 max; padding slots fetch the identity.  out[r] for r in [first, last) (last 0 = all). */
void orc_reduce_rows( const orc_segments* s, const double* vals, const int32_t* cols, int32_t first, int32_t last,
                      int32_t fetch, int32_t reduce, double identity, double* out );

/* CSR-Adaptive block descriptors (CSRAdaptiveKernel::initValueSize) for one value size.
 * Returns the number of descriptors; out (6 int32 per descriptor: type, firstSegment,
 * blockSize, segmentsInBlock, warpIdx, warpsCount) may be NULL to query.
 * LONG and the final sentinel leave blockSize/segmentsInBlock undefined in TNL; they are 0 here. */
int64_t orc_csr_adaptive_blocks( int32_t n, const int32_t* offsets, int32_t value_size, int32_t warp_size,
                                 int32_t block_size, int32_t streamed_shared_bytes, int32_t max_adaptive_elements,
                                 int32_t* out );

/* Algorithms::reduce( a*b, Plus ) sequential order (128-element blocks, 4 accumulators). */
double orc_dot_f64( int64_t n, const double* a, const double* b );
float orc_dot_f32( int64_t n, const float* a, const float* b );

/* Krylov solvers on a host CSR matrix. method: 0 CG, 1 BiCGStab, 2 GMRES, 3 TFQMR, 4 BiCGStab(ell).
 * variant (GMRES): 0 CGS, 1 CGSR, 2 MGS, 3 MGSR, 4 CWY.  precond: 0 none, 1 diagonal, 2 ILU(0). */
typedef struct {
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
   int32_t preconditioner;
   int32_t bicgstab_ell;
} orc_solver_params;

void orc_solver_params_default( orc_solver_params* p );

/* Returns 1 when converged (checkConvergence), 0 otherwise, -1 on GMRES breakdown (TNL throws).
 * history receives every setResidue value (up to history_cap). */
/* SymmetricMatrix product of a lower-triangle CSR matrix (TNL's Host order, one thread):
 * y *= beta when beta != 1, then per row r in [first, last): y_c += (alpha a_rc) x_r for c < r,
 * y_r += alpha * sum_c a_rc x_c. */
void orc_symmetric_spmv( int32_t n, const int32_t* row_ptr, const int32_t* cols, const double* vals, const double* x,
                         double* y, double alpha, double beta, int32_t first, int32_t last );

/* Binary matrix product (Real = bool, ComputeReal = int): y_r = alpha * sum_c x_c + beta * y_r over
 * the stored columns, in int32 with TNL's keeper choice by alpha / beta. */
void orc_binary_spmv( int32_t n, const int32_t* row_ptr, const int32_t* cols, const int32_t* x, int32_t* y,
                      int32_t alpha, int32_t beta, int32_t first, int32_t last );

/* ILU(0) factors of a square CSR matrix (rows of U in natural order, U's first entry of a row is
 * the one divided by) and x = (LU)^-1 b when b and x are given; NULL arrays query the sizes. */
int32_t orc_ilu0( int32_t n, const int32_t* row_ptr, const int32_t* cols, const double* vals, int32_t* nnzL,
                  int32_t* nnzU, int32_t* Lrp, int32_t* Lc, double* Lv, int32_t* Urp, int32_t* Uc, double* Uv,
                  const double* b, double* x );

int32_t orc_solve( int32_t method, int32_t n, const int32_t* row_ptr, const int32_t* cols, const double* vals,
                   const double* b, double* x, const orc_solver_params* params, int32_t* iterations, double* residue,
                   double* history, int32_t history_cap, int32_t* history_len );

#ifdef __cplusplus
}
#endif
