// Row traversal of a sparse matrix on the device: SparseMatrixView::reduceRows with a fixed
// menu of fetch / reduce operations (Matrices/SparseMatrixView.hpp:495-595), and what TNL
// builds on it -- getRowCapacities and getCompressedRowLengths (SparseMatrixView.hpp:116-160) --
// plus getRowCapacity, a host copy of one row (getRow, SparseMatrixView.hpp:222-250) and the
// device view that the C++ front's lambda traversals (forElements, forRows, reduceRows with
// user lambdas) run over.
//
// reduceRows follows TNL's Host traversal: one row at a time, local indexes in increasing
// order, a padding slot (column -1) contributes the identity, the reduction folds left to
// right starting from the identity.  One lane per row here, in that order, so the results are
// TNL's bit for bit for every fetch / reduce pair.
#include "views.hpp"

#include <cmath>
#include <cstring>

namespace tnlhip {

namespace {

// internal fetch: the entries a stored entry of a SymmetricMatrix stands for (its mirror too)
constexpr int kFetchSymmetricCount = 100;

__device__ __forceinline__ double valueAt( const double* v, int64_t g ) { return v[ g ]; }
__device__ __forceinline__ double valueAt( const float* v, int64_t g ) { return (double) v[ g ]; }
__device__ __forceinline__ double valueAt( const bf16* v, int64_t g ) { return (double) __bfloat162float( v[ g ] ); }
__device__ __forceinline__ double valueAt( const uint8_t*, int64_t ) { return 1.0; }  // binary: every entry is a one

template< typename View, typename V, typename Out >
__global__ __launch_bounds__( 256 ) void reduceRowsKernel( int32_t first, int32_t last, View view,
                                                           const int32_t* __restrict__ col, const V* __restrict__ val,
                                                           int fetch, int reduce, double identity, Out* __restrict__ out )
{
   const int64_t stride = (int64_t) gridDim.x * blockDim.x;
   for( int64_t r = first + (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < last; r += stride ) {
      const int32_t size = view.size( (int32_t) r );
      double acc = identity;
      for( int32_t l = 0; l < size; l++ ) {
         const int64_t g = view.global( (int32_t) r, l );
         const int32_t c = col[ g ];
         double f = identity;  // padding slots fetch the identity (SparseMatrixView.hpp:518-527)
         if( c != -1 ) {
            const double a = valueAt( val, g );
            switch( fetch ) {
               case TNLHIP_FETCH_VALUE: f = a; break;
               case TNLHIP_FETCH_ABS_VALUE: f = fabs( a ); break;
               case TNLHIP_FETCH_ONE: f = 1.0; break;
               case TNLHIP_FETCH_NONZERO: f = a != 0.0 ? 1.0 : 0.0; break;
               case TNLHIP_FETCH_COLUMN: f = (double) c; break;
               default: f = c != r ? 2.0 : 1.0; break;  // kFetchSymmetricCount
            }
         }
         switch( reduce ) {
            case TNLHIP_REDUCE_SUM: acc = acc + f; break;
            case TNLHIP_REDUCE_MIN: acc = f < acc ? f : acc; break;
            default: acc = f > acc ? f : acc; break;  // TNLHIP_REDUCE_MAX
         }
      }
      out[ r ] = (Out) acc;
   }
}

// Slots of one row in local order (padding included), for the host copy of getRow.
template< typename View, typename V >
__global__ __launch_bounds__( 256 ) void rowSlotsKernel( int32_t row, int32_t size, View view,
                                                         const int32_t* __restrict__ col, const V* __restrict__ val,
                                                         int32_t* __restrict__ ocol, double* __restrict__ oval )
{
   for( int32_t l = threadIdx.x; l < size; l += blockDim.x ) {
      const int64_t g = view.global( row, l );
      ocol[ l ] = col[ g ];
      oval[ l ] = col[ g ] == -1 && ! val ? 0.0 : valueAt( val, g );
   }
}

template< typename View >
__global__ void rowSizeKernel( int32_t row, View view, int32_t* out )
{
   *out = view.size( row );
}

void reduceRows( const Matrix& A, int32_t first, int32_t last, int fetch, int reduce, double identity, void* out,
                 int outDtype, hipStream_t st )
{
   if( last <= first )
      return;
   withValueType( A.dtype, [ & ]( auto* tag ) {
      using V = std::remove_pointer_t< decltype( tag ) >;
      const V* val = reinterpret_cast< const V* >( A.values.ptr );
      withView( A, [ & ]( auto view ) {
         const unsigned grid = gridFor( last - first, 256 );
         switch( outDtype ) {
            case TNLHIP_F64:
               reduceRowsKernel<<< grid, 256, 0, st >>>( first, last, view, A.columns.ptr, val, fetch, reduce, identity,
                                                         (double*) out );
               break;
            case TNLHIP_F32:
               reduceRowsKernel<<< grid, 256, 0, st >>>( first, last, view, A.columns.ptr, val, fetch, reduce, identity,
                                                         (float*) out );
               break;
            default:
               reduceRowsKernel<<< grid, 256, 0, st >>>( first, last, view, A.columns.ptr, val, fetch, reduce, identity,
                                                         (int32_t*) out );
               break;
         }
      } );
   } );
   TNLHIP_CHECK( hipGetLastError() );
}

int32_t rowSize( const Matrix& A, int32_t row )
{
   const HostSegments& m = A.meta;
   switch( m.kind ) {
      case KIND_CSR: return m.offsets[ row + 1 ] - m.offsets[ row ];
      case KIND_ELLPACK: return m.segmentSize;
      case KIND_SLICED: return m.sliceSegmentSizes[ row / m.sliceSize ];
      default: {
         // ChunkedEllpack / BiEllpack: the device index map's own count
         DeviceArray< int32_t > d( 1 );
         withView( A, [ & ]( auto view ) { rowSizeKernel<<< 1, 1 >>>( row, view, d.ptr ); } );
         TNLHIP_CHECK( hipGetLastError() );
         return d.download()[ 0 ];
      }
   }
}

Matrix& matrixOf( tnlhip_matrix A ) { return *reinterpret_cast< Matrix* >( A ); }

}  // namespace

// getRowCapacities: the stored (non-padding) entries of every row (SparseMatrixView.hpp:143-160)
void storedCounts( const Matrix& A, int32_t* countsDevice, hipStream_t st )
{
   reduceRows( A, 0, A.rows, TNLHIP_FETCH_ONE, TNLHIP_REDUCE_SUM, 0.0, countsDevice, TNLHIP_INT32, st );
}

// getNonzeroElementsCount of a SymmetricMatrix: every stored entry off the diagonal stands for
// two (SparseMatrixView.hpp:195-214)
int64_t symmetricNonzeros( const Matrix& A, hipStream_t st )
{
   if( A.rows == 0 )
      return 0;
   DeviceArray< int32_t > counts( A.rows );
   reduceRows( A, 0, A.rows, kFetchSymmetricCount, TNLHIP_REDUCE_SUM, 0.0, counts.ptr, TNLHIP_INT32, st );
   TNLHIP_CHECK( hipStreamSynchronize( st ) );
   int64_t total = 0;
   for( int32_t c : counts.download() )
      total += c;
   return total;
}

}  // namespace tnlhip

using namespace tnlhip;

extern "C" {

int tnlhip_matrix_reduce_rows( tnlhip_matrix A, int32_t first, int32_t last, int fetch, int reduce, double identity,
                               void* out, int out_dtype, void* stream )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( A, "NULL matrix" );
      const Matrix& M = matrixOf( A );
      if( last == 0 )
         last = M.rows;
      TNLHIP_REQUIRE( first >= 0 && first <= last && last <= M.rows, "row range out of bounds" );
      TNLHIP_REQUIRE( out || first == last, "NULL output" );
      TNLHIP_REQUIRE( fetch >= TNLHIP_FETCH_VALUE && fetch <= TNLHIP_FETCH_COLUMN, "unknown fetch" );
      TNLHIP_REQUIRE( reduce >= TNLHIP_REDUCE_SUM && reduce <= TNLHIP_REDUCE_MAX, "unknown reduction" );
      TNLHIP_REQUIRE( out_dtype == TNLHIP_F64 || out_dtype == TNLHIP_F32 || out_dtype == TNLHIP_INT32,
                      "output type is F64, F32 or INT32" );
      reduceRows( M, first, last, fetch, reduce, identity, out, out_dtype, (hipStream_t) stream );
   } );
}

int tnlhip_matrix_get_row_capacities( tnlhip_matrix A, int32_t* capacities, void* stream )
{
   return tnlhip_matrix_reduce_rows( A, 0, 0, TNLHIP_FETCH_ONE, TNLHIP_REDUCE_SUM, 0.0, capacities, TNLHIP_INT32,
                                     stream );
}

int tnlhip_matrix_get_compressed_row_lengths( tnlhip_matrix A, int32_t* lengths, void* stream )
{
   return tnlhip_matrix_reduce_rows( A, 0, 0, TNLHIP_FETCH_NONZERO, TNLHIP_REDUCE_SUM, 0.0, lengths, TNLHIP_INT32,
                                     stream );
}

int tnlhip_matrix_get_row_capacity( tnlhip_matrix A, int32_t row, int32_t* capacity )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( A && capacity, "NULL argument" );
      const Matrix& M = matrixOf( A );
      TNLHIP_REQUIRE( row >= 0 && row < M.rows, "row out of bounds" );
      *capacity = rowSize( M, row );
   } );
}

int tnlhip_matrix_get_row( tnlhip_matrix A, int32_t row, int32_t* columns, double* values, int32_t capacity,
                           int32_t* size )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( A && size, "NULL argument" );
      const Matrix& M = matrixOf( A );
      TNLHIP_REQUIRE( row >= 0 && row < M.rows, "row out of bounds" );
      const int32_t n = rowSize( M, row );
      *size = n;
      if( n == 0 || ( ! columns && ! values ) )
         return;
      TNLHIP_REQUIRE( capacity >= n, "row buffers are shorter than the row" );
      DeviceArray< int32_t > dc( n );
      DeviceArray< double > dv( n );
      withValueType( M.dtype, [ & ]( auto* tag ) {
         using V = std::remove_pointer_t< decltype( tag ) >;
         withView( M, [ & ]( auto view ) {
            rowSlotsKernel<<< 1, 256 >>>( row, n, view, M.columns.ptr, reinterpret_cast< const V* >( M.values.ptr ),
                                          dc.ptr, dv.ptr );
         } );
      } );
      TNLHIP_CHECK( hipGetLastError() );
      if( columns )
         TNLHIP_CHECK( hipMemcpy( columns, dc.ptr, sizeof( int32_t ) * n, hipMemcpyDeviceToHost ) );
      if( values )
         TNLHIP_CHECK( hipMemcpy( values, dv.ptr, sizeof( double ) * n, hipMemcpyDeviceToHost ) );
   } );
}

int tnlhip_matrix_get_view( tnlhip_matrix A, tnlhip_matrix_view* view )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( A && view, "NULL argument" );
      const Matrix& M = matrixOf( A );
      std::memset( view, 0, sizeof( *view ) );
      view->format = M.format;
      view->value_dtype = M.dtype;
      view->rows = M.rows;
      view->columns = M.cols;
      view->values = M.values.ptr;
      view->column_indexes = M.columns.ptr;
      view->offsets = M.offsets.ptr;
      view->segment_size = M.meta.segmentSize;
      view->aligned_size = M.meta.alignedSize;
      view->slice_offsets = M.sliceOffsets.ptr;
      view->slice_segment_sizes = M.sliceSizes.ptr;
      view->row_to_chunk = M.rowToChunk.ptr;
      view->row_to_slice = M.rowToSlice.ptr;
      view->slices = M.slices.ptr;
      view->row_perm = M.rowPerm.ptr;
      view->group_pointers = M.groupPointers.ptr;
      view->symmetric = M.symmetric ? 1 : 0;
   } );
}

int tnlhip_matrix_values_changed( tnlhip_matrix A, void* stream )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( A, "NULL matrix" );
      Matrix& M = matrixOf( A );
      TNLHIP_CHECK( hipStreamSynchronize( (hipStream_t) stream ) );
      M.diagonalValid = false;
      if( M.stripes )
         buildStripePlan( M, M.stripes->stripes, nullptr );  // the plan holds a copy of the entries
      M.nnz = countNonzeros( M, nullptr );  // forElements may have changed column indexes
   } );
}

}  // extern "C"
