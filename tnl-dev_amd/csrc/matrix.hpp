// Sparse matrix handle: segments metadata (host copy for export, device copy for the
// kernels), value / column storage in HBM and the per-matrix kernel parameters.
#pragma once

#include "common.hpp"

#include <map>
#include <memory>

namespace tnlhip {

enum SegKind
{
   KIND_CSR = 0,
   KIND_ELLPACK = 1,
   KIND_SLICED = 2,
   KIND_CHUNKED = 3,
   KIND_BIELLPACK = 4
};

inline int kindOf( int format )
{
   switch( format ) {
      case TNLHIP_ELLPACK: return KIND_ELLPACK;
      case TNLHIP_SLICED_ELLPACK: return KIND_SLICED;
      case TNLHIP_CHUNKED_ELLPACK: return KIND_CHUNKED;
      case TNLHIP_BIELLPACK: return KIND_BIELLPACK;
      default: return KIND_CSR;
   }
}

// TNL constants of the segment formats
constexpr int32_t kEllpackAlignment = 32;   // Ellpack.h:21
constexpr int32_t kSliceSize = 32;          // SlicedEllpack.h:22
constexpr int32_t kChunksInSlice = 256;     // ChunkedEllpack.h:182
constexpr int32_t kDesiredChunkSize = 16;   // ChunkedEllpack.h:182
constexpr int32_t kStrip = 32;              // BiEllpack.h:22 (WarpSize template parameter)
constexpr int32_t kLogStrip = 5;
constexpr int32_t kGroupsPerStrip = kLogStrip + 1;

// Segment metadata built on the host from the row capacities (the way TNL builds it for
// every device, e.g. ChunkedEllpack.hpp:194-235 builds a Host copy and assigns it).
struct HostSegments
{
   int kind = KIND_CSR;
   int32_t n = 0;
   int64_t storage = 0;    // getStorageSize()
   int64_t allocated = 0;  // elements that can hold data
   // CSR
   std::vector< int32_t > offsets;
   // Ellpack
   int32_t segmentSize = 0, alignedSize = 0;
   // SlicedEllpack (sliceSize: TNL's SliceSize template parameter, 32 by default; a power of two)
   std::vector< int32_t > sliceOffsets, sliceSegmentSizes;
   int32_t slicedSize = 0;
   int32_t sliceSize = kSliceSize;
   // ChunkedEllpack
   int32_t numberOfSlices = 0;
   std::vector< int32_t > rowToChunk, rowToSlice, chunksToSegments, rowPointers, slices;  // slices: 4 per slice
   // BiEllpack
   int32_t virtualRows = 0;
   std::vector< int32_t > rowPerm, groupPointers;
};

void buildSegments( HostSegments& s, int kind, int32_t n, const int32_t* capacities );

// CSR-Adaptive blocks (same greedy split as CSRAdaptiveKernel::initValueSize) for the
// given wave width / block size / streamed LDS budget.  6 int32 per block.
std::vector< int32_t > buildAdaptiveBlocks( int32_t n, const int32_t* offsets, int value_size, int wave_size,
                                            int block_size, int streamed_bytes, int max_adaptive_elements );

// LDS bytes per wavefront for CSR-Adaptive STREAM blocks (512 FP64 / 1024 FP32 products):
// 4 KiB per wave keeps 32 waves per CU within the 160 KiB of LDS.
constexpr int kAdaptiveStreamBytesPerWave = 4096;

// Adaptive block types (detail/CSRAdaptiveKernelBlockDescriptor.h)
enum { BLOCK_LONG = 0, BLOCK_STREAM = 1, BLOCK_VECTOR = 2 };

// Column-striped copy of a CSR matrix (stripes.hip): entries stripe-major, one partial sum per
// (stripe, row).  Built on request (tnlhip_matrix_set_param "column_stripes").
struct StripePlan
{
   int32_t stripes = 0;               // S
   int32_t width = 0;                 // columns per stripe
   int64_t entries = 0;               // stored entries copied
   DeviceArray< int32_t > offsets;    // S * rows + 1: start of (stripe s, row r) at s * rows + r
   DeviceArray< int32_t > columns;    // entries
   DeviceArray< uint8_t > values;     // entries of the matrix value type
   DeviceArray< uint8_t > partial;    // S * rows partial sums in the compute type
};

struct Matrix
{
   int format = TNLHIP_CSR_SCALAR;
   int dtype = TNLHIP_F64;
   int32_t rows = 0, cols = 0;
   int64_t nnz = 0;
   HostSegments meta;
   DeviceArray< uint8_t > values;   // dtypeBytes(dtype) * meta.allocated
   DeviceArray< int32_t > columns;  // meta.allocated
   // device metadata
   DeviceArray< int32_t > offsets, sliceOffsets, sliceSizes, rowToChunk, rowToSlice, slices, rowPerm, groupPointers;
   DeviceArray< int32_t > adaptive;  // 4 int32 per block: type|warpIdx<<8|warpsCount<<16, first, size, segs
   DeviceArray< uint8_t > biRank;    // BiEllpack: rowPermArray[row] - 32 * strip, one byte per row
   std::vector< int32_t > adaptiveHost;
   int32_t adaptiveBlocks = 0;
   int32_t lanes = 0;      // lanes per row for CSR light / hybrid / vector
   int32_t defaultLanes = 0;  // the format's own choice, restored by lanes_per_row 0
   bool lanesExplicit = false;  // lanes_per_row set by the caller: the cooperative kernel always runs
   int32_t xcdRemap = 0;
   // every chunk of a ChunkedEllpack matrix keeps its padding after its elements (true for
   // storage built here from CSR; checked when storage is read from a file)
   bool trailingPadding = true;      // row distribution of the padded kernels: 0 grid-stride, 1 XCD-remapped blocks
   int32_t pairRows = 0;      // padded kernels: two adjacent rows per lane with 8 / 16-byte loads
   int32_t blocksPerCu = 16;  // workgroups per CU the grid of the padded kernels is sized for
   int32_t streamY = -1;      // y stores: 1 non-temporal, 0 through L2, -1 the kernel's default
   bool symmetric = false;    // SymmetricMatrix: the lower triangle is stored
   DeviceArray< uint8_t > diagonal;  // cached Jacobi diagonal (vector type)
   bool diagonalValid = false;
   std::unique_ptr< StripePlan > stripes;  // column-striped plan of a CSR matrix, when built

   void uploadMetadata();
};

// Device launches implemented in spmv.hip / convert.hip
void spmvLaunch( const Matrix& A, const void* x, void* y, double alpha, double beta, int32_t first, int32_t last,
                 const void* diag, hipStream_t stream );
// Gated form used inside solvers: skipped on the device when gate[0] != 0 (gateMode 1) or,
// with gateMode 2, unless gate[1] != 0.  last < 0 means all rows.
void spmvLaunchGated( const Matrix& A, const void* x, void* y, double alpha, double beta, const void* diag,
                      const int32_t* gate, int gateMode, hipStream_t stream, int32_t first = 0, int32_t last = -1 );
// Y = alpha A X + beta Y for k <= kMaxMultiRhs right-hand sides, X and Y row-major (entry
// row * k + j) over rows [first, last) (multi.hip)
constexpr int32_t kMaxMultiRhs = 64;
void spmvMultiLaunch( const Matrix& A, int32_t k, const void* X, void* Y, double alpha, double beta, int32_t first,
                      int32_t last, hipStream_t stream );
// One Jacobi step of the splitting of P: znew = z + ( rhs - P z ) / diag over all rows, gated
// as above (general real-valued P only; znew aliases neither z nor rhs)
void spmvJacobiStepGated( const Matrix& P, const void* z, void* znew, const void* rhs, const void* diag,
                          const int32_t* gate, int gateMode, hipStream_t stream );
// SymmetricMatrix product (symmetric.hip), used by spmvLaunchGated for symmetric matrices
void spmvSymmetric( const Matrix& A, const void* x, void* y, double alpha, double beta, int32_t first, int32_t last,
                    const void* diag, Gate gate, hipStream_t stream );
// Binary matrix product (symmetric.hip): int32 vectors, alpha / beta truncated to int32
void spmvBinary( const Matrix& A, const void* x, void* y, double alpha, double beta, int32_t first, int32_t last,
                 Gate gate, hipStream_t stream );
// getNonzeroElementsCount (SparseMatrixView.hpp:180-215): stored entries, or for a symmetric
// matrix the entries of the whole matrix they stand for (capi.hip, traverse.hip)
int64_t countNonzeros( const Matrix& M, hipStream_t stream );
int64_t symmetricNonzeros( const Matrix& M, hipStream_t stream );
// Builds (S > 0) or drops (S == 0) the column-striped plan of a CSR matrix (stripes.hip).
void buildStripePlan( Matrix& M, int S, hipStream_t stream );
void convertFromCsrDevice( Matrix& A, const int32_t* rowPtr, const int32_t* cols, const double* vals,
                           hipStream_t stream,
                           bool keepZeros = false );
void gatherToCsr( const Matrix& A, const int32_t* rpDevice, int32_t* colsDevice, double* valsDevice,
                  hipStream_t stream,
                  bool keepZeros = false );
void extractDiagonal( const Matrix& A, void* diag, hipStream_t stream );
// segment size of every row (the slots a row owns)
void capacitiesOf( const Matrix& A, int32_t* capsDevice, hipStream_t stream );
// getRowCapacities: stored (non-padding) entries of every row of a matrix (traverse.hip) or of a
// device CSR (convert.hip)
void storedCounts( const Matrix& A, int32_t* countsDevice, hipStream_t stream );
void csrStoredCounts( int32_t n, const int32_t* rpDevice, const int32_t* colsDevice, int32_t* capsDevice,
                      hipStream_t stream );

// Matrix Market files with TNL's MatrixReader / MatrixWriter semantics (mtx.cpp).  The CSR
// keeps TNL's reading result: row capacities from the file, padding -1, explicit zeros kept.
struct MtxCsr
{
   int32_t rows = 0, columns = 0;
   std::vector< int32_t > rowPointers, columnIndexes;
   std::vector< double > values;
};
MtxCsr readMtxCsr( const std::string& path );
void writeMtx( const std::string& path, int32_t rows, int32_t columns, const std::vector< int32_t >& rowPointers,
               const std::vector< int32_t >& columnIndexes, const std::vector< double >& values );

// TNL binary files (File / Object format) of SparseMatrix, tnlfile.cpp.  host = the row-major
// organization and Devices::Host kernel names of TNL's Host matrices; otherwise column-major
// with Devices::Cuda names, the layout the device arrays already have.
struct TnlFileMatrix
{
   std::string type;
   int format = TNLHIP_CSR_SCALAR, dtype = TNLHIP_F64;
   bool symmetric = false;
   int32_t rows = 0, columns = 0;
   int64_t fileStorage = 0;               // length of the stored arrays in the file
   HostSegments meta;                     // checked, column-major
   std::vector< int32_t > columnIndexes;  // meta.allocated, device order
   std::vector< uint8_t > values;         // meta.allocated elements of dtype, device order
};
TnlFileMatrix readTnlFile( const std::string& path, bool host );
void writeTnlFile( const std::string& path, bool host, int format, int dtype, bool symmetric, int32_t rows,
                   int32_t columns, const HostSegments& meta, const int32_t* columnIndexes, const void* values );

}  // namespace tnlhip
