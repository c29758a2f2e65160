// Product with several right-hand sides at once: Y = alpha A X + beta Y for k <= 64 columns.
//
// SparseMatrixView::vectorProduct (Matrices/SparseMatrixView.hpp:390-491) takes one vector; a
// caller with k of them (block Krylov methods, many time levels, the solve server's requests)
// would read the matrix k times.  Here the matrix is read once: X and Y are row-major (entry
// row * k + j), a row is taken by KP / VW lanes (KP = k rounded up to a power of two), each lane
// keeping VW adjacent right-hand sides as one 16-byte load and store (two FP64 or four FP32
// columns when k is a multiple of VW and X, Y are 16-byte aligned; VW = 1 otherwise); the lanes
// of a row load the same stored element (one transaction) while their gathers of x at that column
// are contiguous, and each lane keeps 8 slots' loads in flight.  Each column is summed left to right
// over the row's slots, rounding each product (TNL's Host order: CSRScalarKernel.hpp:70-107 and
// the padded formats' per-segment loops), so column j of Y equals TNL's product of column j of X
// bit for bit; padding slots (column -1) are skipped.  Keepers as SparseMatrixView.hpp:447-471.
#include "spmv.hpp"

namespace tnlhip {

namespace {

constexpr int kMultiChunk = 8;  // slots of a row whose loads are in flight together
constexpr int kMultiXcds = 8;   // workgroups are dispatched round-robin over the 8 XCDs

// VW-wide vector of the compute type (VW = 1: the scalar itself)
template< typename A, int VW >
struct Lanes
{
   typedef A type __attribute__( ( ext_vector_type( VW ) ) );
};
template< typename A >
struct Lanes< A, 1 >
{
   typedef A type;
};

template< typename A, typename T >
__device__ __forceinline__ A lane( const T& v, int i )
{
   if constexpr( sizeof( T ) == sizeof( A ) )
      return v;
   else
      return v[ i ];
}

// A row is taken by L = KP / VW lanes; lane jq holds columns VW jq .. VW jq + VW - 1 of X and Y,
// loaded and stored VW at a time (VW > 1 only when k % VW == 0 and X, Y are 16-byte aligned).
// Workgroup b works on logical row block (b % 8) * (grid / 8) + b / 8, so each XCD takes a
// contiguous eighth of the rows and the x rows a block gathers stay in that XCD's L2 (the grid is
// a multiple of 8; blocks past the last row exit).
template< int KP, int VW, typename V, typename A, typename Y, typename View >
__global__ __launch_bounds__( 256 ) void multiProductKernel( View view, const int32_t* __restrict__ col,
                                                             const V* __restrict__ val, int32_t k,
                                                             const Y* __restrict__ X, Y* __restrict__ Yout, A alpha,
                                                             A beta, int mode, int32_t first, int32_t last,
                                                             int32_t ncols )
{
   constexpr int L = KP / VW;
   constexpr int kRowsPerBlock = 256 / L;
   using VA = typename Lanes< A, VW >::type;
   using VY = typename Lanes< Y, VW >::type;
   const int jq = threadIdx.x % L;
   const int64_t block = (int64_t) ( blockIdx.x % kMultiXcds ) * ( gridDim.x / kMultiXcds ) + blockIdx.x / kMultiXcds;
   const int64_t r = first + block * kRowsPerBlock + threadIdx.x / L;
   if( r >= last )
      return;
   // no columns: every slot is padding and X may be empty
   const int32_t len = ncols > 0 ? view.size( (int32_t) r ) : 0;
   const int j0 = VW * jq;
   const bool mine = j0 < k;
   const Y* xs = X + ( mine ? j0 : 0 );
   VA sum = (VA) 0;
   for( int32_t l0 = 0; l0 < len; l0 += kMultiChunk ) {
      // the chunk's loads go out together: slots past the row re-read its last slot, padding and
      // lanes past k read row 0 of X; then the slots are summed left to right, each column on its own
      int32_t c[ kMultiChunk ];
      A a[ kMultiChunk ];
      VY xv[ kMultiChunk ];
#pragma unroll
      for( int u = 0; u < kMultiChunk; u++ ) {
         const int64_t g = view.global( (int32_t) r, min( l0 + u, len - 1 ) );
         c[ u ] = l0 + u < len ? col[ g ] : -1;
         a[ u ] = toCompute( val[ g ], A() );
      }
#pragma unroll
      for( int u = 0; u < kMultiChunk; u++ )
         xv[ u ] = *reinterpret_cast< const VY* >( xs + (int64_t) ( c[ u ] >= 0 ? c[ u ] : 0 ) * k );
#pragma unroll
      for( int u = 0; u < kMultiChunk; u++ )
         if( c[ u ] >= 0 ) {  // padding skipped
            if constexpr( VW == 1 )
               sum = mulAddExact( sum, a[ u ], (A) xv[ u ] );
            else
#pragma unroll
               for( int i = 0; i < VW; i++ )
                  sum[ i ] = mulAddExact( lane< A >( sum, i ), a[ u ], (A) lane< Y >( xv[ u ], i ) );
         }
   }
   if( ! mine )
      return;
   Y* out = Yout + r * k + j0;
   VY o;
   if constexpr( VW == 1 )
      o = mode >= 2 ? *out : (Y) 0;
   else
      o = mode >= 2 ? *reinterpret_cast< const VY* >( out ) : (VY) 0;
#pragma unroll
   for( int i = 0; i < VW; i++ ) {
      const A sv = lane< A >( sum, i );
      A v;
      switch( mode ) {
         case 0: v = sv; break;
         case 1: v = alpha * sv; break;
         case 2: v = beta * (A) lane< Y >( o, i ) + sv; break;
         default: v = beta * (A) lane< Y >( o, i ) + alpha * sv; break;
      }
      if constexpr( VW == 1 )
         o = (Y) v;
      else
         o[ i ] = (Y) v;
   }
   if constexpr( VW == 1 )
      *out = o;
   else
      *reinterpret_cast< VY* >( out ) = o;
}

template< int KP, int VW, typename V, typename A, typename Y >
void launchMultiW( const Matrix& M, int32_t k, const Y* X, Y* Yv, double alpha, double beta, int32_t first,
                   int32_t last, hipStream_t st )
{
   int mode;
   if( beta == 0.0 )
      mode = alpha == 1.0 ? 0 : 1;
   else
      mode = alpha == 1.0 ? 2 : 3;
   constexpr int kRowsPerBlock = 256 / ( KP / VW );
   const int64_t rows = last - first;
   int64_t blocks = ( rows + kRowsPerBlock - 1 ) / kRowsPerBlock;
   blocks = ( blocks + kMultiXcds - 1 ) / kMultiXcds * kMultiXcds;
   TNLHIP_REQUIRE( blocks < ( 1ll << 31 ), "too many rows for one launch" );
   withView( M, [ & ]( auto view ) {
      multiProductKernel< KP, VW, V, A, Y ><<< (unsigned) blocks, 256, 0, st >>>(
         view, M.columns.ptr, reinterpret_cast< const V* >( M.values.ptr ), k, X, Yv, (A) alpha, (A) beta, mode, first,
         last, M.cols );
   } );
   TNLHIP_CHECK( hipGetLastError() );
}

// 16-byte lanes (two FP64 or four FP32 columns) when k allows them and X, Y are 16-byte aligned
template< int KP, typename V, typename A, typename Y >
void launchMulti( const Matrix& M, int32_t k, const Y* X, Y* Yv, double alpha, double beta, int32_t first,
                  int32_t last, hipStream_t st )
{
   constexpr int VW = 16 / (int) sizeof( Y ) < KP ? 16 / (int) sizeof( Y ) : KP;
   const bool aligned = ( ( reinterpret_cast< uintptr_t >( X ) | reinterpret_cast< uintptr_t >( Yv ) ) % 16 ) == 0;
   if constexpr( VW > 1 ) {
      if( k % VW == 0 && aligned ) {
         launchMultiW< KP, VW, V, A, Y >( M, k, X, Yv, alpha, beta, first, last, st );
         return;
      }
   }
   launchMultiW< KP, 1, V, A, Y >( M, k, X, Yv, alpha, beta, first, last, st );
}

template< typename V, typename A, typename Y >
void dispatchK( const Matrix& M, int32_t k, const Y* X, Y* Yv, double alpha, double beta, int32_t first, int32_t last,
                hipStream_t st )
{
   if( k <= 1 )
      launchMulti< 1, V, A, Y >( M, k, X, Yv, alpha, beta, first, last, st );
   else if( k <= 2 )
      launchMulti< 2, V, A, Y >( M, k, X, Yv, alpha, beta, first, last, st );
   else if( k <= 4 )
      launchMulti< 4, V, A, Y >( M, k, X, Yv, alpha, beta, first, last, st );
   else if( k <= 8 )
      launchMulti< 8, V, A, Y >( M, k, X, Yv, alpha, beta, first, last, st );
   else if( k <= 16 )
      launchMulti< 16, V, A, Y >( M, k, X, Yv, alpha, beta, first, last, st );
   else if( k <= 32 )
      launchMulti< 32, V, A, Y >( M, k, X, Yv, alpha, beta, first, last, st );
   else
      launchMulti< 64, V, A, Y >( M, k, X, Yv, alpha, beta, first, last, st );
}

}  // namespace

void spmvMultiLaunch( const Matrix& M, int32_t k, const void* X, void* Yv, double alpha, double beta, int32_t first,
                      int32_t last, hipStream_t st )
{
   if( M.symmetric || M.dtype == TNLHIP_BINARY )
      throw Error( TNLHIP_ERROR_UNSUPPORTED, "products with several right-hand sides take general real-valued matrices" );
   TNLHIP_REQUIRE( k >= 1 && k <= kMaxMultiRhs, "the number of right-hand sides must be in [1, 64]" );
   if( last <= first )
      return;
   switch( M.dtype ) {
      case TNLHIP_F64:
         dispatchK< double, double, double >( M, k, (const double*) X, (double*) Yv, alpha, beta, first, last, st );
         break;
      case TNLHIP_F32:
         dispatchK< float, float, float >( M, k, (const float*) X, (float*) Yv, alpha, beta, first, last, st );
         break;
      default:
         dispatchK< bf16, float, float >( M, k, (const float*) X, (float*) Yv, alpha, beta, first, last, st );
         break;
   }
}

}  // namespace tnlhip
