// extern "C" entry points of include/tnlhip.h: argument checks, matrix construction and
// conversion, and the stand-alone SpMV / BLAS-1 calls.  Solver entry points live in
// solvers.hip.
#include "blas.hpp"
#include "matrix.hpp"

#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>

// Opaque handle type of the C ABI.
struct tnlhip_matrix_s : tnlhip::Matrix
{};

namespace tnlhip {

namespace {
thread_local std::string g_lastError;
}

void setLastError( const std::string& msg )
{
   g_lastError = msg;
}

const DeviceProps& deviceProps()
{
   static DeviceProps props;
   static std::once_flag once;
   std::call_once( once, [] {
      int dev = 0;
      if( hipGetDevice( &dev ) == hipSuccess ) {
         hipDeviceProp_t p;
         if( hipGetDeviceProperties( &p, dev ) == hipSuccess ) {
            props.cus = p.multiProcessorCount > 0 ? p.multiProcessorCount : 256;
            props.memory = (int64_t) p.totalGlobalMem;
            props.name = p.gcnArchName;
         }
      }
   } );
   return props;
}

Workspace& threadWorkspace()
{
   thread_local Workspace ws;
   ws.ensure( 4, kMaxMulti, 64 );
   return ws;
}

void Matrix::uploadMetadata()
{
   switch( meta.kind ) {
      case KIND_CSR: offsets.upload( meta.offsets ); break;
      case KIND_SLICED:
         sliceOffsets.upload( meta.sliceOffsets );
         sliceSizes.upload( meta.sliceSegmentSizes );
         break;
      case KIND_CHUNKED:
         rowToChunk.upload( meta.rowToChunk );
         rowToSlice.upload( meta.rowToSlice );
         slices.upload( meta.slices );
         break;
      case KIND_BIELLPACK: {
         rowPerm.upload( meta.rowPerm );
         groupPointers.upload( meta.groupPointers );
         // the product needs only each row's rank in its strip: one byte instead of four
         std::vector< uint8_t > rank( meta.rowPerm.size() );
         for( size_t r = 0; r < rank.size(); r++ )
            rank[ r ] = (uint8_t) ( meta.rowPerm[ r ] - (int32_t) ( r / kStrip ) * kStrip );
         biRank.upload( rank );
         break;
      }
      default: break;
   }
   if( format == TNLHIP_CSR_ADAPTIVE ) {
      // device descriptor: {type | warpIdx << 8 | warpsCount << 16, first, size, segments}
      const int32_t nb = (int32_t) ( adaptiveHost.size() / 6 ) - 1;  // drop the closing block
      std::vector< int32_t > d( (size_t) 4 * std::max( nb, 0 ) );
      for( int32_t i = 0; i < nb; i++ ) {
         const int32_t* b = &adaptiveHost[ 6 * i ];
         d[ 4 * i + 0 ] = b[ 0 ] | ( b[ 4 ] << 8 ) | ( b[ 5 ] << 16 );
         d[ 4 * i + 1 ] = b[ 1 ];
         d[ 4 * i + 2 ] = b[ 2 ];
         d[ 4 * i + 3 ] = b[ 3 ];
      }
      adaptive.upload( d );
      adaptiveBlocks = std::max( nb, 0 );
   }
   TNLHIP_CHECK( hipDeviceSynchronize() );
}

// Wave64 parameters of the adaptive split: 256-lane workgroups (4 waves), 4 KiB of LDS per
// wave for STREAM blocks, VECTOR rows up to 2048 elements.
constexpr int kAdaptiveStreamBytes = 4 * kAdaptiveStreamBytesPerWave;
constexpr int kAdaptiveMaxElements = 1024;

namespace {

int lightLanes( int64_t nnz, int32_t rows )
{
   // CSRLightKernel::init (Kernels/CSRLightKernel.hpp:475-500), extended to wave64
   if( rows <= 0 )
      return 2;
   const int64_t per = ( nnz + rows - 1 ) / rows;
   if( per <= 2 )
      return 2;
   if( per <= 4 )
      return 4;
   if( per <= 8 )
      return 8;
   if( per <= 16 )
      return 16;
   if( per <= 32 )
      return 32;
   return 64;
}

int hybridLanes( int64_t nnz, int32_t rows )
{
   // CSRHybridKernel::init (Kernels/CSRHybridKernel.hpp:159-175): 2^ceil(log2(avg)), capped
   // at the wavefront width
   if( rows <= 0 )
      return 1;
   const int64_t per = (int64_t) std::ceil( (double) nnz / (double) rows );
   int g = 1;
   while( g < per && g < kWave )
      g *= 2;
   return g;
}

// A matrix of the given metadata with uninitialized storage; capTotal = the sum of the row
// capacities, which sets the lanes per row of the CSR kernels.
tnlhip_matrix_s* buildFromSegments( int format, int dtype, int32_t rows, int32_t cols, HostSegments meta,
                                    int64_t capTotal )
{
   auto M = std::make_unique< tnlhip_matrix_s >();
   M->format = format;
   M->dtype = dtype;
   M->rows = rows;
   M->cols = cols;
   M->meta = std::move( meta );
   if( format == TNLHIP_CSR_ADAPTIVE ) {
      const int vsize = dtype == TNLHIP_F64 ? 8 : 4;  // LDS holds products in the compute type
      M->adaptiveHost = buildAdaptiveBlocks( rows, M->meta.offsets.data(), vsize, kWave, 256, kAdaptiveStreamBytes,
                                             kAdaptiveMaxElements );
   }
   // ChunkedEllpack and SlicedEllpack with 4-byte values: slices / row blocks in XCD order (each
   // XCD a contiguous eighth) keep the x they gather in that XCD's L2: at 160^3 FP32 ChunkedEllpack
   // 45.8 -> 42.3 us, SlicedEllpack 47.5 -> 45.7 us; FP64 (twice the x) measured ChunkedEllpack
   // 70 -> 72 us and FP32 Ellpack 49.4 -> 52.1 us, so those keep the plain order
   if( ( format == TNLHIP_CHUNKED_ELLPACK || format == TNLHIP_SLICED_ELLPACK ) && dtype != TNLHIP_F64 )
      M->xcdRemap = 1;
   // FP64 Ellpack: each XCD strides over its own eighth of the row blocks (distribution 3):
   // 160^3, three alternated runs on one box, 78.4 -> 76.1 us; the other FP64 formats measured
   // slower with an XCD split (SlicedEllpack 75.9 -> 79.1, CSRScalar 79.7 -> 82.9, CSRAdaptive
   // 82.4 -> 84.9, ChunkedEllpack 72.0 -> 75.1 us) and keep the plain grid-stride
   if( format == TNLHIP_ELLPACK && ( dtype == TNLHIP_F64 || dtype == TNLHIP_F32 ) )
      M->xcdRemap = 3;  // FP32 with 32 workgroups per CU, below: 51.5 -> 50.3 us
   // FP64 launch shapes from a sweep of workgroups per CU (8-64) and one or two rows per lane at
   // 160^3 (two alternated runs): Ellpack 32 workgroups per CU, 76.6 -> 73.7 us; SlicedEllpack two
   // rows per lane and 32 workgroups per CU, 76.1 -> 73.2 us (each row still summed left to right by
   // one lane, so the bits do not change)
   if( format == TNLHIP_ELLPACK && ( dtype == TNLHIP_F64 || dtype == TNLHIP_F32 ) )
      M->blocksPerCu = 32;
   if( format == TNLHIP_SLICED_ELLPACK && dtype == TNLHIP_F64 ) {
      M->pairRows = 1;
      M->blocksPerCu = 32;
   }
   if( format == TNLHIP_CSR_LIGHT )
      M->lanes = lightLanes( capTotal, rows );
   else if( format == TNLHIP_CSR_HYBRID || format == TNLHIP_CSR_ADAPTIVE )
      M->lanes = hybridLanes( capTotal, rows );
   else if( format == TNLHIP_CSR_VECTOR )
      M->lanes = kWave;
   M->defaultLanes = M->lanes;
   M->values.resize( M->meta.allocated * (int64_t) dtypeBytes( dtype ) );
   M->columns.resize( M->meta.allocated );
   M->uploadMetadata();
   return M.release();
}

tnlhip_matrix_s* buildEmpty( int format, int dtype, int32_t rows, int32_t cols, const std::vector< int32_t >& caps,
                             int32_t sliceSize = kSliceSize )
{
   TNLHIP_REQUIRE( format >= TNLHIP_CSR_SCALAR && format <= TNLHIP_BIELLPACK, "unknown format" );
   TNLHIP_REQUIRE( dtype >= TNLHIP_F64 && dtype <= TNLHIP_BINARY, "unknown value type" );
   TNLHIP_REQUIRE( rows >= 0 && cols >= 0, "negative dimensions" );
   HostSegments meta;
   meta.sliceSize = sliceSize;
   buildSegments( meta, kindOf( format ), rows, caps.data() );
   int64_t capTotal = 0;
   for( int32_t c : caps )
      capTotal += c;
   return buildFromSegments( format, dtype, rows, cols, std::move( meta ), capTotal );
}

__global__ __launch_bounds__( kReduceBlock ) void countKernel( int64_t n, const int32_t* __restrict__ cols,
                                                               ReduceSlot slot )
{
   double acc[ 1 ] = { 0.0 };
   for( int64_t i = (int64_t) blockIdx.x * kReduceBlock + threadIdx.x; i < n; i += (int64_t) gridDim.x * kReduceBlock )
      acc[ 0 ] += cols[ i ] != -1 ? 1.0 : 0.0;
   storeBlockPartials< 1 >( acc, slot );
}

__global__ __launch_bounds__( kReduceBlock ) void countFinalKernel( ReduceSlot slot, int grid, int64_t* result )
{
   __shared__ double tot[ 1 ];
   sumPartials< 1 >( slot, grid, tot );
   if( threadIdx.x == 0 )
      *result = (int64_t) tot[ 0 ];
}

// Smallest and largest entry of a device int32 array (column indexes handed over on the device).
__global__ __launch_bounds__( 256 ) void rangeKernel( int64_t n, const int32_t* __restrict__ v, int32_t* range )
{
   int32_t lo = INT32_MAX, hi = INT32_MIN;
   for( int64_t i = (int64_t) blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t) gridDim.x * 256 ) {
      lo = min( lo, v[ i ] );
      hi = max( hi, v[ i ] );
   }
   for( int o = kWave / 2; o > 0; o /= 2 ) {
      lo = min( lo, __shfl_xor( lo, o, kWave ) );
      hi = max( hi, __shfl_xor( hi, o, kWave ) );
   }
   if( threadIdx.x % kWave == 0 ) {
      atomicMin( range, lo );
      atomicMax( range + 1, hi );
   }
}

void columnRange( const int32_t* cols, int64_t n, int32_t* lo, int32_t* hi, hipStream_t st )
{
   DeviceArray< int32_t > r( 2 );
   const int32_t init[ 2 ] = { INT32_MAX, INT32_MIN };
   TNLHIP_CHECK( hipMemcpyAsync( r.ptr, init, sizeof( init ), hipMemcpyHostToDevice, st ) );
   rangeKernel<<< gridFor( n, 256 ), 256, 0, st >>>( n, cols, r.ptr );
   TNLHIP_CHECK( hipGetLastError() );
   int32_t out[ 2 ];
   TNLHIP_CHECK( hipMemcpyAsync( out, r.ptr, sizeof( out ), hipMemcpyDeviceToHost, st ) );
   TNLHIP_CHECK( hipStreamSynchronize( st ) );
   *lo = out[ 0 ];
   *hi = out[ 1 ];
}

int64_t countStored( const Matrix& M, hipStream_t st )
{
   if( M.meta.allocated == 0 )
      return 0;
   Workspace& ws = threadWorkspace();
   int64_t* r = reinterpret_cast< int64_t* >( ws.scalars.ptr );
   countKernel<<< kReduceGrid, kReduceBlock, 0, st >>>( M.meta.allocated, M.columns.ptr, ws.slot( 0 ) );
   countFinalKernel<<< 1, kReduceBlock, 0, st >>>( ws.slot( 0 ), kReduceGrid, r );
   TNLHIP_CHECK( hipGetLastError() );
   int64_t h = 0;
   TNLHIP_CHECK( hipMemcpyAsync( &h, r, sizeof( h ), hipMemcpyDeviceToHost, st ) );
   TNLHIP_CHECK( hipStreamSynchronize( st ) );
   return h;
}

}  // namespace

int64_t countNonzeros( const Matrix& M, hipStream_t st )
{
   return M.symmetric ? symmetricNonzeros( M, st ) : countStored( M, st );
}

namespace {

inline float bf16ToFloat( uint16_t b )
{
   uint32_t u = (uint32_t) b << 16;
   float f;
   std::memcpy( &f, &u, 4 );
   return f;
}

}  // namespace

}  // namespace tnlhip

using namespace tnlhip;

// Partial sums of tnlhip_multidot_noalloc: one column pass of up to kMaxMulti columns writes
// kReduceGrid partials per column.  Each GPU holds its own copy for the life of the process.
constexpr int kNoAllocColumns = kMaxMulti;
__device__ __align__( 256 ) double scratch[ (size_t) kNoAllocColumns * kReduceGrid ];

namespace {

// The device address of the current GPU's `scratch`, looked up once per device (with that
// device current) on its first use; host code passes this address, never the array's name.
double* noAllocScratch()
{
   constexpr int kMaxDevices = 64;
   static std::mutex lock;
   static double* address[ kMaxDevices ] = {};
   int dev = 0;
   TNLHIP_CHECK( hipGetDevice( &dev ) );
   TNLHIP_REQUIRE( dev >= 0 && dev < kMaxDevices, "device ordinal out of range" );
   std::lock_guard< std::mutex > guard( lock );
   if( ! address[ dev ] ) {
      void* p = nullptr;
      TNLHIP_CHECK( hipGetSymbolAddress( &p, HIP_SYMBOL( scratch ) ) );
      TNLHIP_REQUIRE( p, "no device address for the scratch array" );
      address[ dev ] = static_cast< double* >( p );
   }
   return address[ dev ];
}

}  // namespace

extern "C" {

const char* tnlhip_status_string( int s )
{
   switch( s ) {
      case TNLHIP_SUCCESS: return "success";
      case TNLHIP_ERROR_INVALID_ARGUMENT: return "invalid argument";
      case TNLHIP_ERROR_RUNTIME: return "HIP runtime error";
      case TNLHIP_ERROR_OUT_OF_MEMORY: return "out of memory";
      case TNLHIP_ERROR_CAPACITY: return "row capacity exceeded";
      case TNLHIP_ERROR_UNSUPPORTED: return "unsupported";
      case TNLHIP_ERROR_BREAKDOWN: return "solver breakdown";
   }
   return "unknown status";
}

const char* tnlhip_last_error( void )
{
   return g_lastError.c_str();
}

int tnlhip_device_info( char* name, int name_len, int* compute_units, int64_t* global_memory )
{
   return guarded( [ & ] {
      int count = 0;
      TNLHIP_CHECK( hipGetDeviceCount( &count ) );
      TNLHIP_REQUIRE( count > 0, "no HIP device" );
      const DeviceProps& p = deviceProps();
      if( name && name_len > 0 ) {
         std::strncpy( name, p.name.c_str(), (size_t) name_len - 1 );
         name[ name_len - 1 ] = 0;
      }
      if( compute_units )
         *compute_units = p.cus;
      if( global_memory )
         *global_memory = p.memory;
   } );
}

}  // extern "C"

namespace tnlhip {
namespace {

// A matrix holding a CSR given by host or device arrays.
//  staged == false: SparseMatrix::operator=( csr ) (SparseMatrix.hpp:1070-1107): the row
//    capacities are the source's getRowCapacities -- its stored (non-padding) entries per row,
//    explicit zeros included (SparseMatrixView.hpp:143-160) -- and the entries are copied in
//    source order with explicit zeros dropped, so a row keeps trailing padding for its zeros.
//  staged == true: setRowCapacities( row lengths ) followed by setElement of every stored entry
//    (SparseMatrix.hpp:284-305, SparseMatrixView.hpp:270-345): the capacities are the CSR row
//    lengths, padding slots (-1) stay padding and explicit zeros are stored.
tnlhip_matrix_s* createFromCsr( int format, int value_dtype, int32_t rows, int32_t columns, const int32_t* row_pointers,
                                const int32_t* column_indexes, const double* values, bool host_pointers, bool staged )
{
   TNLHIP_REQUIRE( rows >= 0 && columns >= 0, "negative dimensions" );
   TNLHIP_REQUIRE( row_pointers != nullptr || rows == 0, "row_pointers is NULL" );
   std::vector< int32_t > rp( (size_t) rows + 1, 0 );
   if( rows > 0 ) {
      if( host_pointers )
         std::memcpy( rp.data(), row_pointers, sizeof( int32_t ) * rp.size() );
      else
         TNLHIP_CHECK( hipMemcpy( rp.data(), row_pointers, sizeof( int32_t ) * rp.size(), hipMemcpyDeviceToHost ) );
   }
   TNLHIP_REQUIRE( rp[ 0 ] == 0, "row_pointers[0] must be 0" );
   for( int32_t i = 0; i < rows; i++ )
      TNLHIP_REQUIRE( rp[ i + 1 ] >= rp[ i ], "row_pointers must be non-decreasing" );
   const int64_t nnzIn = rp[ rows ];
   TNLHIP_REQUIRE( nnzIn == 0 || column_indexes != nullptr, "column_indexes is NULL" );
   if( host_pointers && nnzIn > 0 ) {
      for( int64_t k = 0; k < nnzIn; k++ )
         TNLHIP_REQUIRE( column_indexes[ k ] >= -1 && column_indexes[ k ] < columns, "column index out of range" );
   }
   else if( nnzIn > 0 ) {
      // device arrays: the range of the column indexes is checked on the device before any
      // kernel gathers x through them
      int32_t lo = 0, hi = 0;
      columnRange( column_indexes, nnzIn, &lo, &hi, nullptr );
      TNLHIP_REQUIRE( lo >= -1 && hi < columns, "column index out of range" );
   }
   DeviceArray< int32_t > drp, dcol;
   DeviceArray< double > dval;
   const int32_t* prp = row_pointers;
   const int32_t* pcol = column_indexes;
   // a binary matrix may come without values: every entry is then a one
   std::vector< double > ones;
   if( value_dtype == TNLHIP_BINARY && values == nullptr ) {
      ones.assign( (size_t) std::max< int64_t >( nnzIn, 0 ), 1.0 );
      values = ones.data();
      if( ! host_pointers ) {
         dval.upload( ones );
         values = dval.ptr;
      }
   }
   TNLHIP_REQUIRE( nnzIn == 0 || values != nullptr, "values is NULL" );
   const double* pval = values;
   if( host_pointers ) {
      drp.upload( rp );
      dcol.upload( column_indexes, nnzIn );
      dval.upload( values, nnzIn );
      prp = drp.ptr;
      pcol = dcol.ptr;
      pval = dval.ptr;
   }
   std::vector< int32_t > caps( rows );
   if( staged ) {
      for( int32_t i = 0; i < rows; i++ )
         caps[ i ] = rp[ i + 1 ] - rp[ i ];
   }
   else if( rows > 0 ) {
      DeviceArray< int32_t > dcaps( rows );
      csrStoredCounts( rows, prp, pcol, dcaps.ptr, nullptr );
      caps = dcaps.download();
   }
   std::unique_ptr< tnlhip_matrix_s > M( buildEmpty( format, value_dtype, rows, columns, caps ) );
   convertFromCsrDevice( *M, prp, pcol, pval, nullptr, staged );
   M->nnz = countStored( *M, nullptr );
   return M.release();
}

}  // namespace
}  // namespace tnlhip

extern "C" {

int tnlhip_matrix_create_csr( int format, int value_dtype, int32_t rows, int32_t columns, const int32_t* row_pointers,
                              const int32_t* column_indexes, const double* values, int host_pointers,
                              tnlhip_matrix* out )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( out != nullptr, "out is NULL" );
      *out = createFromCsr( format, value_dtype, rows, columns, row_pointers, column_indexes, values, host_pointers != 0,
                            false );
   } );
}

int tnlhip_matrix_create_staged( int format, int value_dtype, int symmetric, int32_t rows, int32_t columns,
                                 const int32_t* row_pointers, const int32_t* column_indexes, const double* values,
                                 tnlhip_matrix* out )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( out != nullptr, "out is NULL" );
      if( symmetric ) {
         TNLHIP_REQUIRE( rows == columns, "a symmetric matrix is square" );
         if( value_dtype == TNLHIP_BINARY )
            throw Error( TNLHIP_ERROR_UNSUPPORTED, "symmetric binary matrices are not supported" );
         for( int32_t i = 0; i < rows; i++ )
            for( int32_t k = row_pointers[ i ]; k < row_pointers[ i + 1 ]; k++ )
               if( column_indexes[ k ] > i )
                  throw Error( TNLHIP_ERROR_INVALID_ARGUMENT,
                               "a symmetric matrix stores its lower triangle: entry (" + std::to_string( i ) + ", " +
                                  std::to_string( column_indexes[ k ] ) + ") lies above the diagonal" );
      }
      std::unique_ptr< tnlhip_matrix_s > M(
         createFromCsr( format, value_dtype, rows, columns, row_pointers, column_indexes, values, true, true ) );
      M->symmetric = symmetric != 0;
      M->nnz = countNonzeros( *M, nullptr );
      *out = M.release();
   } );
}

int tnlhip_matrix_create_symmetric_csr( int format, int value_dtype, int32_t rows, const int32_t* row_pointers,
                                        const int32_t* column_indexes, const double* values, int host_pointers,
                                        tnlhip_matrix* out )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( out != nullptr && rows >= 0, "bad argument" );
      TNLHIP_REQUIRE( row_pointers != nullptr || rows == 0, "row_pointers is NULL" );
      if( value_dtype == TNLHIP_BINARY )
         throw Error( TNLHIP_ERROR_UNSUPPORTED, "symmetric binary matrices are not supported" );
      std::vector< int32_t > rp( (size_t) rows + 1, 0 );
      if( rows > 0 ) {
         if( host_pointers )
            std::memcpy( rp.data(), row_pointers, sizeof( int32_t ) * rp.size() );
         else
            TNLHIP_CHECK( hipMemcpy( rp.data(), row_pointers, sizeof( int32_t ) * rp.size(), hipMemcpyDeviceToHost ) );
      }
      const int64_t nnz = rp[ rows ];
      std::vector< int32_t > ci( (size_t) std::max< int64_t >( nnz, 0 ) );
      if( nnz > 0 ) {
         if( host_pointers )
            std::memcpy( ci.data(), column_indexes, sizeof( int32_t ) * ci.size() );
         else
            TNLHIP_CHECK( hipMemcpy( ci.data(), column_indexes, sizeof( int32_t ) * ci.size(), hipMemcpyDeviceToHost ) );
      }
      for( int32_t i = 0; i < rows; i++ )
         for( int32_t k = rp[ i ]; k < rp[ i + 1 ] && k < nnz; k++ )
            if( ci[ k ] > i )
               throw Error( TNLHIP_ERROR_INVALID_ARGUMENT,
                            "a symmetric matrix stores its lower triangle: entry (" + std::to_string( i ) + ", " +
                               std::to_string( ci[ k ] ) + ") lies above the diagonal" );
      tnlhip_matrix M = nullptr;
      TNLHIP_CHECK_STATUS( tnlhip_matrix_create_csr( format, value_dtype, rows, rows, row_pointers, column_indexes,
                                                     values, host_pointers, &M ) );
      M->symmetric = true;
      M->nnz = countNonzeros( *M, nullptr );
      *out = M;
   } );
}

}  // extern "C"

namespace {

// SlicedEllpack with another slice size (TNL's SliceSize template parameter): the same stored
// entries, every one of them (explicit zeros included) in its row's order, laid out for slices of
// S rows; row capacities become the stored entries per row, as operator= makes them.
void reslice( Matrix& A, int32_t S )
{
   DeviceArray< int32_t > dcaps( A.rows );
   storedCounts( A, dcaps.ptr, nullptr );
   std::vector< int32_t > caps = dcaps.download();
   std::unique_ptr< tnlhip_matrix_s > M( buildEmpty( A.format, A.dtype, A.rows, A.cols, caps, S ) );
   std::vector< int32_t > rp( (size_t) A.rows + 1, 0 );
   for( int32_t i = 0; i < A.rows; i++ )
      rp[ i + 1 ] = rp[ i ] + caps[ i ];
   DeviceArray< int32_t > drp, dcol( rp[ A.rows ] );
   DeviceArray< double > dval( rp[ A.rows ] );
   drp.upload( rp );
   gatherToCsr( A, drp.ptr, dcol.ptr, dval.ptr, nullptr, true );
   convertFromCsrDevice( *M, drp.ptr, dcol.ptr, dval.ptr, nullptr, true );
   A.meta = std::move( M->meta );
   A.values = std::move( M->values );
   A.columns = std::move( M->columns );
   A.sliceOffsets = std::move( M->sliceOffsets );
   A.sliceSizes = std::move( M->sliceSizes );
   A.diagonalValid = false;
}

}  // namespace

extern "C" {

int tnlhip_matrix_convert( tnlhip_matrix source, int format, int value_dtype, tnlhip_matrix* out )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( source && out, "NULL argument" );
      const Matrix& S = *source;
      // operator= takes the source's getRowCapacities: its stored entries per row
      DeviceArray< int32_t > dcaps( S.rows );
      storedCounts( S, dcaps.ptr, nullptr );
      std::vector< int32_t > caps = dcaps.download();
      std::unique_ptr< tnlhip_matrix_s > M( buildEmpty( format, value_dtype, S.rows, S.cols, caps ) );
      std::vector< int32_t > rp( (size_t) S.rows + 1, 0 );
      for( int32_t i = 0; i < S.rows; i++ )
         rp[ i + 1 ] = rp[ i ] + caps[ i ];
      DeviceArray< int32_t > drp, dcol( rp[ S.rows ] );
      DeviceArray< double > dval( rp[ S.rows ] );
      drp.upload( rp );
      gatherToCsr( S, drp.ptr, dcol.ptr, dval.ptr, nullptr );
      convertFromCsrDevice( *M, drp.ptr, dcol.ptr, dval.ptr, nullptr );
      M->symmetric = S.symmetric;
      M->nnz = countNonzeros( *M, nullptr );
      *out = M.release();
   } );
}

int tnlhip_mtx_read_csr( const char* path, int32_t* rows, int32_t* columns, int64_t* storage, int32_t* row_pointers,
                         int32_t* column_indexes, double* values )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( path, "NULL path" );
      const MtxCsr m = readMtxCsr( path );
      if( rows )
         *rows = m.rows;
      if( columns )
         *columns = m.columns;
      if( storage )
         *storage = (int64_t) m.columnIndexes.size();
      if( row_pointers )
         std::memcpy( row_pointers, m.rowPointers.data(), sizeof( int32_t ) * m.rowPointers.size() );
      if( column_indexes )
         std::memcpy( column_indexes, m.columnIndexes.data(), sizeof( int32_t ) * m.columnIndexes.size() );
      if( values )
         std::memcpy( values, m.values.data(), sizeof( double ) * m.values.size() );
   } );
}

int tnlhip_mtx_read( const char* path, int format, int value_dtype, tnlhip_matrix* out )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( path && out, "NULL argument" );
      // MatrixReader< Matrix, Device >::readMtx (MatrixReader.hpp:19-26): read into a host matrix
      // (padding where entries repeat, explicit zeros stored), then matrix = hostMatrix, which
      // takes its getRowCapacities and drops the explicit zeros
      const MtxCsr m = readMtxCsr( path );
      *out = createFromCsr( format, value_dtype, m.rows, m.columns, m.rowPointers.data(),
                            m.columnIndexes.empty() ? nullptr : m.columnIndexes.data(),
                            m.values.empty() ? nullptr : m.values.data(), true, false );
   } );
}

int tnlhip_mtx_write_csr( const char* path, int32_t rows, int32_t columns, const int32_t* row_pointers,
                          const int32_t* column_indexes, const double* values )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( path && rows >= 0 && columns >= 0 && ( rows == 0 || row_pointers ), "bad argument" );
      std::vector< int32_t > rp( row_pointers, row_pointers + rows + 1 );
      const int64_t nnz = rows > 0 ? rp[ rows ] : 0;
      std::vector< int32_t > ci( column_indexes, column_indexes + nnz );
      std::vector< double > v( values, values + nnz );
      writeMtx( path, rows, columns, rp, ci, v );
   } );
}

int tnlhip_mtx_write( tnlhip_matrix A, const char* path )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( A && path, "NULL argument" );
      const Matrix& S = *A;
      // writeMtx of a device matrix writes a host copy, hostMatrix = matrix (MatrixWriter.hpp:35-40):
      // capacities from getRowCapacities, explicit zeros dropped
      DeviceArray< int32_t > dcaps( S.rows );
      storedCounts( S, dcaps.ptr, nullptr );
      std::vector< int32_t > caps = dcaps.download();
      std::vector< int32_t > rp( (size_t) S.rows + 1, 0 );
      for( int32_t i = 0; i < S.rows; i++ )
         rp[ i + 1 ] = rp[ i ] + caps[ i ];
      DeviceArray< int32_t > drp, dcol( rp[ S.rows ] );
      DeviceArray< double > dval( rp[ S.rows ] );
      drp.upload( rp );
      gatherToCsr( S, drp.ptr, dcol.ptr, dval.ptr, nullptr );
      writeMtx( path, S.rows, S.cols, rp, dcol.download(), dval.download() );
   } );
}

int tnlhip_matrix_save( tnlhip_matrix A, const char* path, int layout )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( A && path, "NULL argument" );
      TNLHIP_REQUIRE( layout == TNLHIP_FILE_CUDA_LAYOUT || layout == TNLHIP_FILE_HOST_LAYOUT, "unknown file layout" );
      if( A->dtype == TNLHIP_BF16 || A->dtype == TNLHIP_BINARY )
         throw Error( TNLHIP_ERROR_UNSUPPORTED, "TNL files are written here for double or float values only" );
      // the type string of a SlicedEllpack file does not name its slice size: TNL reads it with 32
      if( kindOf( A->format ) == KIND_SLICED && A->meta.sliceSize != kSliceSize )
         throw Error( TNLHIP_ERROR_UNSUPPORTED, "SlicedEllpack files hold slices of 32 rows (TNL's default SliceSize)" );
      const int64_t n = A->meta.allocated;
      std::vector< int32_t > cols( (size_t) n );
      std::vector< uint8_t > vals( (size_t) n * dtypeBytes( A->dtype ) );
      if( n > 0 ) {
         TNLHIP_CHECK( hipMemcpy( cols.data(), A->columns.ptr, sizeof( int32_t ) * n, hipMemcpyDeviceToHost ) );
         TNLHIP_CHECK( hipMemcpy( vals.data(), A->values.ptr, vals.size(), hipMemcpyDeviceToHost ) );
      }
      writeTnlFile( path, layout == TNLHIP_FILE_HOST_LAYOUT, A->format, A->dtype, A->symmetric, A->rows, A->cols,
                    A->meta, cols.data(), vals.data() );
   } );
}

// Whether every chunk of a ChunkedEllpack storage (device order: slot k of chunk c at
// pointer + c + 256 k) keeps its padding after its elements.  Storage read from a file may not.
bool chunkedPaddingTrails( const HostSegments& meta, const std::vector< int32_t >& cols )
{
   for( int32_t s = 0; s < meta.numberOfSlices; s++ ) {
      const int32_t chunkSize = meta.slices[ 4 * s + 1 ], pointer = meta.slices[ 4 * s + 3 ];
      for( int32_t c = 0; c < kChunksInSlice; c++ ) {
         bool padding = false;
         for( int32_t k = 0; k < chunkSize; k++ ) {
            const int64_t g = (int64_t) pointer + c + (int64_t) kChunksInSlice * k;
            if( g >= (int64_t) cols.size() )
               return false;
            if( cols[ g ] < 0 )
               padding = true;
            else if( padding )
               return false;
         }
      }
   }
   return true;
}

int tnlhip_matrix_load( const char* path, int format, int value_dtype, int layout, tnlhip_matrix* out )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( path && out, "NULL argument" );
      TNLHIP_REQUIRE( layout == TNLHIP_FILE_CUDA_LAYOUT || layout == TNLHIP_FILE_HOST_LAYOUT, "unknown file layout" );
      TNLHIP_REQUIRE( format >= -1 && format <= TNLHIP_BIELLPACK, "unknown format" );
      TNLHIP_REQUIRE( value_dtype >= -1 && value_dtype <= TNLHIP_BF16, "unknown value type" );
      TnlFileMatrix f = readTnlFile( path, layout == TNLHIP_FILE_HOST_LAYOUT );
      const int fmt = format < 0 ? f.format : format;
      const int dtype = value_dtype < 0 ? f.dtype : value_dtype;
      // SparseMatrix::load requires the file's type (Object.hpp:42-50); the CSR kernel is free
      if( kindOf( fmt ) != kindOf( f.format ) || dtype != f.dtype )
         throw Error( TNLHIP_ERROR_INVALID_ARGUMENT,
                      "Failed to deserialize an object from the file '" + std::string( path ) +
                         "': object type does not match (expected format " + std::to_string( fmt ) + ", value type " +
                         std::to_string( dtype ) + ", found " + f.type + ")." );
      const int64_t capTotal = kindOf( fmt ) == KIND_CSR ? f.meta.allocated : 0;
      std::unique_ptr< tnlhip_matrix_s > M(
         buildFromSegments( fmt, dtype, f.rows, f.columns, std::move( f.meta ), capTotal ) );
      if( fmt == TNLHIP_CHUNKED_ELLPACK )
         M->trailingPadding = chunkedPaddingTrails( M->meta, f.columnIndexes );
      M->values.upload( f.values.data(), (int64_t) f.values.size() );
      M->columns.upload( f.columnIndexes );
      M->symmetric = f.symmetric;
      M->nnz = countNonzeros( *M, nullptr );
      *out = M.release();
   } );
}

int tnlhip_file_inspect( const char* path, int layout, tnlhip_file_info* info, char* type, int type_len )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( path, "NULL path" );
      TNLHIP_REQUIRE( layout == TNLHIP_FILE_CUDA_LAYOUT || layout == TNLHIP_FILE_HOST_LAYOUT, "unknown file layout" );
      const TnlFileMatrix f = readTnlFile( path, layout == TNLHIP_FILE_HOST_LAYOUT );
      if( info ) {
         info->format = f.format;
         info->value_dtype = f.dtype;
         info->rows = f.rows;
         info->columns = f.columns;
         info->storage_size = f.fileStorage;
         info->symmetric = f.symmetric ? 1 : 0;
         info->reserved = 0;
      }
      if( type && type_len > 0 )
         std::snprintf( type, (size_t) type_len, "%s", f.type.c_str() );
   } );
}

int tnlhip_matrix_destroy( tnlhip_matrix A )
{
   return guarded( [ & ] { delete A; } );
}

int tnlhip_matrix_get_info( tnlhip_matrix A, tnlhip_matrix_info* info )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( A && info, "NULL argument" );
      info->rows = A->rows;
      info->columns = A->cols;
      info->format = A->format;
      info->value_dtype = A->dtype;
      info->storage_size = A->meta.storage;
      info->allocated = A->meta.allocated;
      info->nonzeros = A->nnz;
      info->lanes_per_row = A->lanes;
      info->symmetric = A->symmetric ? 1 : 0;
   } );
}

int64_t tnlhip_matrix_get_metadata( tnlhip_matrix A, const char* name_c, int32_t* out )
{
   if( ! A || ! name_c )
      return -1;
   const HostSegments& s = A->meta;
   const std::string name( name_c );
   auto put = [ & ]( const std::vector< int32_t >& v ) -> int64_t
   {
      if( out && ! v.empty() )
         std::memcpy( out, v.data(), v.size() * sizeof( int32_t ) );
      return (int64_t) v.size();
   };
   switch( s.kind ) {
      case KIND_CSR:
         if( name == "offsets" )
            return put( s.offsets );
         if( name == "adaptiveBlocks" && A->format == TNLHIP_CSR_ADAPTIVE )
            return put( A->adaptiveHost );
         break;
      case KIND_ELLPACK:
         if( name == "ellpack" )
            return put( { s.segmentSize, s.n, s.alignedSize } );
         break;
      case KIND_SLICED:
         if( name == "sliceOffsets" )
            return put( s.sliceOffsets );
         if( name == "sliceSegmentSizes" )
            return put( s.sliceSegmentSizes );
         if( name == "sliced" )
            return put( { s.slicedSize, (int32_t) s.storage } );
         if( name == "sliceSize" )
            return put( { s.sliceSize } );
         break;
      case KIND_CHUNKED:
         if( name == "rowToChunkMapping" )
            return put( s.rowToChunk );
         if( name == "rowToSliceMapping" )
            return put( s.rowToSlice );
         if( name == "chunksToSegmentsMapping" )
            return put( s.chunksToSegments );
         if( name == "rowPointers" )
            return put( s.rowPointers );
         if( name == "slices" )
            return put( s.slices );
         if( name == "chunked" )
            return put( { s.numberOfSlices, (int32_t) s.storage } );
         break;
      case KIND_BIELLPACK:
         if( name == "rowPermArray" )
            return put( s.rowPerm );
         if( name == "groupPointers" )
            return put( s.groupPointers );
         if( name == "biellpack" )
            return put( { s.virtualRows, (int32_t) ( s.storage & 0xffffffff ), (int32_t) ( s.storage >> 32 ) } );
         break;
   }
   return -1;
}

int64_t tnlhip_segments_metadata( int format, int32_t n, const int32_t* capacities, const char* name, int32_t* out )
{
   if( format < TNLHIP_CSR_SCALAR || format > TNLHIP_BIELLPACK || n < 0 || ( n > 0 && ! capacities ) || ! name )
      return -1;
   try {
      tnlhip_matrix_s M;
      M.format = format;
      M.rows = n;
      buildSegments( M.meta, kindOf( format ), n, capacities );
      if( format == TNLHIP_CSR_ADAPTIVE )
         M.adaptiveHost = buildAdaptiveBlocks( n, M.meta.offsets.data(), 8, kWave, 256, kAdaptiveStreamBytes,
                                               kAdaptiveMaxElements );
      if( std::string( name ) == "storage" ) {
         if( out ) {
            out[ 0 ] = (int32_t) ( M.meta.storage & 0xffffffff );
            out[ 1 ] = (int32_t) ( M.meta.storage >> 32 );
         }
         return 2;
      }
      return tnlhip_matrix_get_metadata( &M, name, out );
   }
   catch( const std::exception& e ) {
      setLastError( e.what() );
      return -1;
   }
}

int tnlhip_matrix_get_storage( tnlhip_matrix A, double* values, int32_t* columns )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( A, "NULL matrix" );
      const int64_t n = A->meta.allocated;
      if( n == 0 )
         return;
      if( columns )
         TNLHIP_CHECK( hipMemcpy( columns, A->columns.ptr, sizeof( int32_t ) * n, hipMemcpyDeviceToHost ) );
      if( values ) {
         if( A->dtype == TNLHIP_BINARY ) {
            std::vector< int32_t > c( n );
            TNLHIP_CHECK( hipMemcpy( c.data(), A->columns.ptr, sizeof( int32_t ) * n, hipMemcpyDeviceToHost ) );
            for( int64_t i = 0; i < n; i++ )
               values[ i ] = c[ i ] != -1 ? 1.0 : 0.0;
         }
         else if( A->dtype == TNLHIP_F64 )
            TNLHIP_CHECK( hipMemcpy( values, A->values.ptr, sizeof( double ) * n, hipMemcpyDeviceToHost ) );
         else if( A->dtype == TNLHIP_F32 ) {
            std::vector< float > t( n );
            TNLHIP_CHECK( hipMemcpy( t.data(), A->values.ptr, sizeof( float ) * n, hipMemcpyDeviceToHost ) );
            for( int64_t i = 0; i < n; i++ )
               values[ i ] = t[ i ];
         }
         else {
            std::vector< uint16_t > t( n );
            TNLHIP_CHECK( hipMemcpy( t.data(), A->values.ptr, sizeof( uint16_t ) * n, hipMemcpyDeviceToHost ) );
            for( int64_t i = 0; i < n; i++ )
               values[ i ] = bf16ToFloat( t[ i ] );
         }
      }
   } );
}

int tnlhip_matrix_get_diagonal( tnlhip_matrix A, void* diagonal, void* stream )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( A && diagonal, "NULL argument" );
      extractDiagonal( *A, diagonal, (hipStream_t) stream );
   } );
}

int tnlhip_matrix_set_param( tnlhip_matrix A, const char* name_c, int32_t value )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( A && name_c, "NULL argument" );
      const std::string name( name_c );
      if( name == "lanes_per_row" ) {
         TNLHIP_REQUIRE( value >= 0 && value <= kWave && ( value & ( value - 1 ) ) == 0,
                         "lanes_per_row must be 0 (the format's default) or a power of two in [1, 64]" );
         TNLHIP_REQUIRE( kindOf( A->format ) == KIND_CSR && A->format != TNLHIP_CSR_SCALAR,
                         "lanes_per_row applies to CSR light / hybrid / vector / adaptive fallback" );
         A->lanes = value == 0 ? A->defaultLanes : value;
         A->lanesExplicit = value != 0;
      }
      else if( name == "xcd_remap" ) {
         // 2 (contiguous blocks without the XCD remap) is switched off: a run of it ended in a
         // GPU memory fault whose cause is not established
         TNLHIP_REQUIRE( value == 0 || value == 1 || value == 3,
                         "xcd_remap is 0 (grid-stride), 1 (XCD-remapped blocks) or 3 (per-XCD grid-stride)" );
         A->xcdRemap = value;
      }
      else if( name == "stream_y" ) {
         TNLHIP_REQUIRE( value >= -1 && value <= 1, "stream_y is -1 (kernel default), 0 or 1" );
         A->streamY = value;
      }
      else if( name == "column_stripes" )
         buildStripePlan( *A, value, nullptr );
      else if( name == "slice_size" ) {
         TNLHIP_REQUIRE( kindOf( A->format ) == KIND_SLICED, "slice_size applies to SlicedEllpack matrices" );
         TNLHIP_REQUIRE( value >= 1 && value <= 256 && ( value & ( value - 1 ) ) == 0,
                         "slice_size is a power of two in [1, 256]" );
         if( value != A->meta.sliceSize )
            reslice( *A, value );
      }
      else if( name == "pair_rows" )
         A->pairRows = value != 0;
      else if( name == "blocks_per_cu" ) {
         TNLHIP_REQUIRE( value >= 1 && value <= 64, "blocks_per_cu must be in [1, 64]" );
         A->blocksPerCu = value;
      }
      else
         throw Error( TNLHIP_ERROR_INVALID_ARGUMENT, "unknown parameter " + name );
   } );
}

int tnlhip_spmv( tnlhip_matrix A, const void* x, void* y, double alpha, double beta, int32_t first, int32_t last,
                 void* stream )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( A, "NULL matrix" );
      // x holds A->cols entries and y A->rows: either may be NULL when it holds none
      TNLHIP_REQUIRE( ( x || A->cols == 0 ) && ( y || A->rows == 0 ), "NULL vector" );
      if( last == 0 )
         last = A->rows;
      TNLHIP_REQUIRE( first >= 0 && first <= last && last <= A->rows, "row range out of bounds" );
      spmvLaunch( *A, x, y, alpha, beta, first, last, nullptr, (hipStream_t) stream );
   } );
}

int tnlhip_spmv_multi( tnlhip_matrix A, int32_t k, const void* X, void* Y, double alpha, double beta, int32_t first,
                       int32_t last, void* stream )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( A, "NULL matrix" );
      TNLHIP_REQUIRE( ( X || A->cols == 0 ) && ( Y || A->rows == 0 ), "NULL block of vectors" );
      if( last == 0 )
         last = A->rows;
      TNLHIP_REQUIRE( first >= 0 && first <= last && last <= A->rows, "row range out of bounds" );
      spmvMultiLaunch( *A, k, X, Y, alpha, beta, first, last, (hipStream_t) stream );
   } );
}

int tnlhip_dot( int dtype, int64_t n, const void* a, const void* b, double* result, void* stream )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( result, "NULL result" );
      TNLHIP_REQUIRE( dtype == TNLHIP_F64 || dtype == TNLHIP_F32, "vectors are F64 or F32" );
      if( n <= 0 ) {
         *result = 0.0;
         return;
      }
      Workspace& ws = threadWorkspace();
      deviceDot( dtype, n, a, b, ws.scalars.ptr, false, ws, (hipStream_t) stream );
      TNLHIP_CHECK( hipMemcpyAsync( result, ws.scalars.ptr, sizeof( double ), hipMemcpyDeviceToHost, (hipStream_t) stream ) );
      TNLHIP_CHECK( hipStreamSynchronize( (hipStream_t) stream ) );
   } );
}

int tnlhip_norm2( int dtype, int64_t n, const void* a, double* result, void* stream )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( result, "NULL result" );
      TNLHIP_REQUIRE( dtype == TNLHIP_F64 || dtype == TNLHIP_F32, "vectors are F64 or F32" );
      if( n <= 0 ) {
         *result = 0.0;
         return;
      }
      Workspace& ws = threadWorkspace();
      deviceDot( dtype, n, a, a, ws.scalars.ptr, true, ws, (hipStream_t) stream );
      TNLHIP_CHECK( hipMemcpyAsync( result, ws.scalars.ptr, sizeof( double ), hipMemcpyDeviceToHost, (hipStream_t) stream ) );
      TNLHIP_CHECK( hipStreamSynchronize( (hipStream_t) stream ) );
   } );
}

int tnlhip_axpby( int dtype, int64_t n, double alpha, const void* x, double beta, void* y, void* stream )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( dtype == TNLHIP_F64 || dtype == TNLHIP_F32, "vectors are F64 or F32" );
      deviceAxpby( dtype, n, alpha, x, beta, y, (hipStream_t) stream );
   } );
}

int tnlhip_multidot( int dtype, int64_t n, int32_t count, const void* V, int64_t ld, const void* w, double* result,
                     void* stream )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( dtype == TNLHIP_F64 || dtype == TNLHIP_F32, "vectors are F64 or F32" );
      TNLHIP_REQUIRE( n > 0 && count > 0 && ld >= n, "bad multidot shape" );
      Workspace& ws = threadWorkspace();
      ws.ensure( 4, kMaxMulti, count );
      deviceMultidot( dtype, n, count, V, ld, w, ws.scalars.ptr, ws, (hipStream_t) stream );
      TNLHIP_CHECK( hipMemcpyAsync( result, ws.scalars.ptr, sizeof( double ) * count, hipMemcpyDeviceToHost,
                                    (hipStream_t) stream ) );
      TNLHIP_CHECK( hipStreamSynchronize( (hipStream_t) stream ) );
   } );
}

int tnlhip_multidot_noalloc( int dtype, int64_t n, int32_t count, const void* V, int64_t ld, const void* w,
                             double* result, void* stream )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( dtype == TNLHIP_F64 || dtype == TNLHIP_F32, "vectors are F64 or F32" );
      TNLHIP_REQUIRE( n > 0 && ld >= n && V && w && result, "bad multidot shape or NULL argument" );
      if( count < 1 || count > kNoAllocColumns )
         throw Error( TNLHIP_ERROR_INVALID_ARGUMENT,
                      "tnlhip_multidot_noalloc takes 1 to " + std::to_string( kNoAllocColumns )
                         + " columns (its partial sums live in a fixed device array); got "
                         + std::to_string( count ) );
      // every partial the final pass reads is written by the first pass of the same call, so the
      // array needs no clearing
      multidotSlot( dtype, n, count, V, ld, w, result, ReduceSlot{ noAllocScratch() }, (hipStream_t) stream );
   } );
}

int tnlhip_gemv( int dtype, int64_t n, int32_t count, double alpha, const void* V, int64_t ld, const double* x,
                 double beta, void* y, void* stream )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( dtype == TNLHIP_F64 || dtype == TNLHIP_F32, "vectors are F64 or F32" );
      TNLHIP_REQUIRE( n > 0 && count > 0 && count <= 64 && ld >= n, "bad gemv shape" );
      Workspace& ws = threadWorkspace();
      ws.ensure( 4, kMaxMulti, 64 );
      TNLHIP_CHECK( hipMemcpyAsync( ws.scalars.ptr, x, sizeof( double ) * count, hipMemcpyHostToDevice,
                                    (hipStream_t) stream ) );
      deviceGemv( dtype, n, count, alpha, V, ld, ws.scalars.ptr, beta, y, (hipStream_t) stream );
      TNLHIP_CHECK( hipStreamSynchronize( (hipStream_t) stream ) );
   } );
}

int tnlhip_exclusive_scan_i32( int64_t n, const int32_t* in, int32_t* out, int total_out, void* stream )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( n >= 0, "negative size" );
      Workspace& ws = threadWorkspace();
      deviceExclusiveScan( n, in, out, total_out ? out + n : nullptr, ws, (hipStream_t) stream );
   } );
}

}  // extern "C"
