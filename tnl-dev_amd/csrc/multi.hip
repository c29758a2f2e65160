// Product with several right-hand sides at once: Y = alpha A X + beta Y for k <= 64 columns.
//
// SparseMatrixView::vectorProduct (Matrices/SparseMatrixView.hpp:390-491) takes one vector; a
// caller with k of them (block Krylov methods, many time levels, the solve server's requests)
// would read the matrix k times.  Here the matrix is read once: X and Y are row-major (entry
// row * k + j), a row is taken by KP = k rounded up to a power of two lanes, lane j < k keeps
// right-hand side j, and every lane of the group loads the same stored element (one transaction)
// while the KP gathers of x at that column are contiguous.  Each column is summed left to right
// over the row's slots, rounding each product (TNL's Host order: CSRScalarKernel.hpp:70-107 and
// the padded formats' per-segment loops), so column j of Y equals TNL's product of column j of X
// bit for bit; padding slots (column -1) are skipped.  Keepers as SparseMatrixView.hpp:447-471.
#include "spmv.hpp"

namespace tnlhip {

namespace {

template< int KP, typename V, typename A, typename Y, typename View >
__global__ __launch_bounds__( 256 ) void multiProductKernel( View view, const int32_t* __restrict__ col,
                                                             const V* __restrict__ val, int32_t k,
                                                             const Y* __restrict__ X, Y* __restrict__ Yout, A alpha,
                                                             A beta, int mode, int32_t first, int32_t last )
{
   constexpr int kRowsPerBlock = 256 / KP;
   const int j = threadIdx.x % KP;
   const int64_t r = first + (int64_t) blockIdx.x * kRowsPerBlock + threadIdx.x / KP;
   if( r >= last )
      return;
   const int32_t len = view.size( (int32_t) r );
   const bool mine = j < k;
   A sum = 0;
   for( int32_t l = 0; l < len; l++ ) {
      const int64_t g = view.global( (int32_t) r, l );
      const int32_t c = col[ g ];
      if( c < 0 )
         continue;  // padding
      const A a = toCompute( val[ g ], A() );
      const A xv = mine ? (A) X[ (int64_t) c * k + j ] : A( 0 );
      sum = mulAddExact( sum, a, xv );
   }
   if( ! mine )
      return;
   Y* out = Yout + r * k + j;
   A v;
   switch( mode ) {
      case 0: v = sum; break;
      case 1: v = alpha * sum; break;
      case 2: v = beta * (A) *out + sum; break;
      default: v = beta * (A) *out + alpha * sum; break;
   }
   *out = (Y) v;
}

template< int KP, typename V, typename A, typename Y >
void launchMulti( const Matrix& M, int32_t k, const Y* X, Y* Yv, double alpha, double beta, int32_t first,
                  int32_t last, hipStream_t st )
{
   int mode;
   if( beta == 0.0 )
      mode = alpha == 1.0 ? 0 : 1;
   else
      mode = alpha == 1.0 ? 2 : 3;
   const int64_t rows = last - first;
   const unsigned grid = (unsigned) ( ( rows + 256 / KP - 1 ) / ( 256 / KP ) );
   withView( M, [ & ]( auto view ) {
      multiProductKernel< KP, V, A, Y ><<< grid, 256, 0, st >>>( view, M.columns.ptr,
                                                                 reinterpret_cast< const V* >( M.values.ptr ), k, X,
                                                                 Yv, (A) alpha, (A) beta, mode, first, last );
   } );
   TNLHIP_CHECK( hipGetLastError() );
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
