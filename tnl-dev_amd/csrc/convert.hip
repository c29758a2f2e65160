// Format conversion on the device: SparseMatrix::operator= (Matrices/SparseMatrix.hpp:1070-1180).
// The target gets the source's row capacities (getRowCapacities, i.e. the source segment
// sizes), every slot starts as padding (column -1, value 0) and the source elements are
// copied in source order with explicit zeros and padding skipped.  Also: diagonal
// extraction for the Jacobi preconditioner and row capacities of a matrix.
#include "views.hpp"

namespace tnlhip {

__device__ __forceinline__ double toDouble( double v ) { return v; }
__device__ __forceinline__ double toDouble( float v ) { return (double) v; }
__device__ __forceinline__ double toDouble( bf16 v ) { return (double) __bfloat162float( v ); }
__device__ __forceinline__ double toDouble( uint8_t v ) { return (double) v; }

template< typename V >
__device__ __forceinline__ V fromDouble( double v );
template<>
__device__ __forceinline__ double fromDouble< double >( double v ) { return v; }
template<>
__device__ __forceinline__ float fromDouble< float >( double v ) { return (float) v; }
template<>
__device__ __forceinline__ bf16 fromDouble< bf16 >( double v ) { return __float2bfloat16( (float) v ); }
template<>
__device__ __forceinline__ uint8_t fromDouble< uint8_t >( double v ) { return v != 0.0; }

template< typename V >
__global__ void fillPaddingKernel( int64_t n, int32_t* cols, V* vals )
{
   const int64_t stride = (int64_t) gridDim.x * blockDim.x;
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride ) {
      cols[ i ] = -1;
      if( vals )
         vals[ i ] = fromDouble< V >( 0.0 );
   }
}

// One lane per row, sequential over the row: local indexes must be assigned in source order.
template< typename View, typename V >
__global__ void scatterFromCsrKernel( int32_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ scol,
                                      const double* __restrict__ sval, View dst, int32_t* __restrict__ dcol,
                                      V* __restrict__ dval, int32_t* __restrict__ overflow, int keepZeros )
{
   const int64_t stride = (int64_t) gridDim.x * blockDim.x;
   for( int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride ) {
      int32_t local = 0;
      const int32_t cap = dst.size( (int32_t) r );
      for( int32_t k = rp[ r ]; k < rp[ r + 1 ]; k++ ) {
         const double v = sval[ k ];
         const int32_t c = scol[ k ];
         if( ( keepZeros || v != 0.0 ) && c != -1 ) {
            if( local >= cap ) {
               *overflow = 1;
               break;
            }
            const int64_t g = dst.global( (int32_t) r, local++ );
            dcol[ g ] = c;
            if( dval )
               dval[ g ] = fromDouble< V >( v );
         }
      }
   }
}

// Source row r, local order -> compressed CSR row (values as double) of capacity size(r).
template< typename View, typename V >
__global__ void gatherToCsrKernel( int32_t n, View src, const int32_t* __restrict__ scol, const V* __restrict__ sval,
                                   const int32_t* __restrict__ rp, int32_t* __restrict__ ocol,
                                   double* __restrict__ oval, int keepZeros )
{
   const int64_t stride = (int64_t) gridDim.x * blockDim.x;
   for( int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride ) {
      const int32_t cap = src.size( (int32_t) r );
      int32_t o = rp[ r ];
      for( int32_t l = 0; l < cap; l++ ) {
         const int64_t g = src.global( (int32_t) r, l );
         const int32_t c = scol[ g ];
         const double v = sval ? toDouble( sval[ g ] ) : 1.0;  // a binary matrix stores ones
         if( c != -1 && ( keepZeros || v != 0.0 ) ) {
            ocol[ o ] = c;
            oval[ o ] = v;
            o++;
         }
      }
      for( ; o < rp[ r + 1 ]; o++ ) {
         ocol[ o ] = -1;
         oval[ o ] = 0.0;
      }
   }
}

template< typename View >
__global__ void capacitiesKernel( int32_t n, View v, int32_t* caps )
{
   const int64_t stride = (int64_t) gridDim.x * blockDim.x;
   for( int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride )
      caps[ r ] = v.size( (int32_t) r );
}

// a_ii by a search of row i (SparseMatrixView::getElement), Preconditioners/Diagonal.hpp:22-41.
template< typename View, typename V, typename Y >
__global__ void diagonalKernel( int32_t n, View v, const int32_t* __restrict__ col, const V* __restrict__ val,
                                Y* __restrict__ diag )
{
   const int64_t stride = (int64_t) gridDim.x * blockDim.x;
   for( int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride ) {
      const int32_t cap = v.size( (int32_t) r );
      double d = 0.0;
      for( int32_t l = 0; l < cap; l++ ) {
         const int64_t g = v.global( (int32_t) r, l );
         const int32_t c = col[ g ];
         if( c == r ) {
            d = toDouble( val[ g ] );
            break;
         }
      }
      diag[ r ] = (Y) d;
   }
}

void convertFromCsrDevice( Matrix& A, const int32_t* rowPtr, const int32_t* cols, const double* vals,
                           hipStream_t stream, bool keepZeros )
{
   DeviceArray< int32_t > overflow( 1 );
   TNLHIP_CHECK( hipMemsetAsync( overflow.ptr, 0, sizeof( int32_t ), stream ) );
   withValueType( A.dtype, [ & ]( auto* tag ) {
      using V = std::remove_pointer_t< decltype( tag ) >;
      V* dval = reinterpret_cast< V* >( A.values.ptr );
      if( A.meta.allocated > 0 )
         fillPaddingKernel< V ><<< gridFor( A.meta.allocated, 256 ), 256, 0, stream >>>( A.meta.allocated,
                                                                                          A.columns.ptr, dval );
      if( A.rows > 0 )
         withView( A, [ & ]( auto view ) {
            scatterFromCsrKernel<<< gridFor( A.rows, 256 ), 256, 0, stream >>>(
               A.rows, rowPtr, cols, vals, view, A.columns.ptr, dval, overflow.ptr, keepZeros ? 1 : 0 );
         } );
   } );
   TNLHIP_CHECK( hipGetLastError() );
   int32_t ov = 0;
   TNLHIP_CHECK( hipMemcpyAsync( &ov, overflow.ptr, sizeof( int32_t ), hipMemcpyDeviceToHost, stream ) );
   TNLHIP_CHECK( hipStreamSynchronize( stream ) );
   if( ov )
      throw Error( TNLHIP_ERROR_CAPACITY, "row capacity exceeded during conversion" );
}

__global__ void csrStoredCountsKernel( int32_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ cols,
                                       int32_t* __restrict__ caps )
{
   const int64_t stride = (int64_t) gridDim.x * blockDim.x;
   for( int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride ) {
      int32_t k = 0;
      for( int32_t g = rp[ r ]; g < rp[ r + 1 ]; g++ )
         k += cols[ g ] != -1;
      caps[ r ] = k;
   }
}

void csrStoredCounts( int32_t n, const int32_t* rpDevice, const int32_t* colsDevice, int32_t* capsDevice,
                      hipStream_t stream )
{
   if( n == 0 )
      return;
   csrStoredCountsKernel<<< gridFor( n, 256 ), 256, 0, stream >>>( n, rpDevice, colsDevice, capsDevice );
   TNLHIP_CHECK( hipGetLastError() );
}

void capacitiesOf( const Matrix& A, int32_t* capsDevice, hipStream_t stream )
{
   if( A.rows == 0 )
      return;
   withView( A, [ & ]( auto view ) {
      capacitiesKernel<<< gridFor( A.rows, 256 ), 256, 0, stream >>>( A.rows, view, capsDevice );
   } );
   TNLHIP_CHECK( hipGetLastError() );
}

void gatherToCsr( const Matrix& A, const int32_t* rpDevice, int32_t* colsDevice, double* valsDevice,
                  hipStream_t stream, bool keepZeros )
{
   if( A.rows == 0 )
      return;
   withValueType( A.dtype, [ & ]( auto* tag ) {
      using V = std::remove_pointer_t< decltype( tag ) >;
      const V* sval = reinterpret_cast< const V* >( A.values.ptr );
      withView( A, [ & ]( auto view ) {
         gatherToCsrKernel<<< gridFor( A.rows, 256 ), 256, 0, stream >>>( A.rows, view, A.columns.ptr, sval,
                                                                           rpDevice, colsDevice, valsDevice,
                                                                           keepZeros ? 1 : 0 );
      } );
   } );
   TNLHIP_CHECK( hipGetLastError() );
}

void extractDiagonal( const Matrix& A, void* diag, hipStream_t stream )
{
   if( A.dtype == TNLHIP_BINARY )
      throw Error( TNLHIP_ERROR_UNSUPPORTED, "the diagonal of a binary matrix is not extracted" );
   if( A.rows == 0 )
      return;
   withValueType( A.dtype, [ & ]( auto* tag ) {
      using V = std::remove_pointer_t< decltype( tag ) >;
      const V* val = reinterpret_cast< const V* >( A.values.ptr );
      withView( A, [ & ]( auto view ) {
         if( A.dtype == TNLHIP_F64 )
            diagonalKernel<<< gridFor( A.rows, 256 ), 256, 0, stream >>>( A.rows, view, A.columns.ptr, val,
                                                                           (double*) diag );
         else
            diagonalKernel<<< gridFor( A.rows, 256 ), 256, 0, stream >>>( A.rows, view, A.columns.ptr, val,
                                                                           (float*) diag );
      } );
   } );
   TNLHIP_CHECK( hipGetLastError() );
}

}  // namespace tnlhip
