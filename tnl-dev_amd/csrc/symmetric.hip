// Products of symmetric and of binary matrices, one lane per row over the format's index map.
//
// Symmetric: SparseMatrix< ..., SymmetricMatrix, ... >::vectorProduct
// (Matrices/SparseMatrixView.hpp:390-491 with MatrixType::isSymmetric()).  The matrix stores its
// lower triangle, diagonal included.  When beta != 1 the whole output vector is scaled by beta
// first (TNL multiplies rather than clears, so a NaN in y survives beta = 0).  Then every stored
// a_rc (c != -1) of a row r in [first, last) adds (alpha a_rc) x_r to y_c when c < r -- the
// transposed entry -- and the row sum alpha * sum_c a_rc x_c is added to y_r.  Both additions are
// device-scope floating-point atomics, as on TNL's GPU path; the row sum keeps TNL's Host order
// (each product rounded, then added), the order in which the atomics land is not fixed.  One
// lane per row over the format's index map; the solvers' Jacobi keeper divides by the diagonal
// afterwards.
#include "views.hpp"

namespace tnlhip {

namespace {

template< typename View, typename V, typename A, typename Y >
__global__ __launch_bounds__( 256 ) void symmetricKernel( int32_t first, int32_t last, View view,
                                                          const int32_t* __restrict__ col, const V* __restrict__ val,
                                                          const Y* __restrict__ x, Y* y, A alpha, Gate gate )
{
   if( ! gate.open() )
      return;
   const int64_t stride = (int64_t) gridDim.x * blockDim.x;
   for( int64_t r = first + (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < last; r += stride ) {
      const A xr = (A) x[ r ];
      const int32_t size = view.size( (int32_t) r );
      A s = 0;
      for( int32_t l = 0; l < size; l++ ) {
         const int64_t g = view.global( (int32_t) r, l );
         const int32_t c = col[ g ];
         if( c < 0 )
            continue;
         const A a = toCompute( val[ g ], A() );
         if( c < r )
            unsafeAtomicAdd( y + c, (Y) ( ( alpha * a ) * xr ) );
         s = mulAddExact( s, a, (A) x[ c ] );
      }
      unsafeAtomicAdd( y + r, (Y) ( alpha * s ) );
   }
}

// y *= beta, or y = 0 when clear
template< typename Y >
__global__ __launch_bounds__( 256 ) void scaleKernel( int64_t n, Y* y, Y beta, int clear, Gate gate )
{
   if( ! gate.open() )
      return;
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x )
      y[ i ] = clear ? (Y) 0 : beta * y[ i ];
}

template< typename Y >
__global__ __launch_bounds__( 256 ) void divideKernel( int32_t first, int32_t last, Y* y, const Y* diag, Gate gate )
{
   if( ! gate.open() )
      return;
   for( int64_t i = first + (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < last;
        i += (int64_t) gridDim.x * blockDim.x )
      y[ i ] = y[ i ] / diag[ i ];
}

template< typename V, typename A, typename Y >
void launchSymmetric( const Matrix& M, const Y* x, Y* y, double alpha, double beta, int32_t first, int32_t last,
                      const Y* diag, Gate gate, hipStream_t st )
{
   // the solvers' Jacobi keeper (y = (A x) / diag) starts from y = 0
   if( ( diag || beta != 1.0 ) && M.rows > 0 )
      scaleKernel< Y ><<< gridFor( M.rows, 256 ), 256, 0, st >>>( M.rows, y, (Y) beta, diag ? 1 : 0, gate );
   if( last > first )
      withView( M, [ & ]( auto view ) {
         symmetricKernel<<< gridFor( last - first, 256 ), 256, 0, st >>>(
            first, last, view, M.columns.ptr, reinterpret_cast< const V* >( M.values.ptr ), x, y, (A) alpha, gate );
      } );
   if( diag && last > first )
      divideKernel< Y ><<< gridFor( last - first, 256 ), 256, 0, st >>>( first, last, y, diag, gate );
   TNLHIP_CHECK( hipGetLastError() );
}

// Binary matrices (Real = bool, ComputeReal = Index): the fetch of a stored entry is x_c itself
// (SparseMatrixView.hpp:413-437), summed in int32; the keepers are those of a general matrix with
// alpha and beta as int32 (SparseMatrixView.hpp:439-491).
template< typename View >
__global__ __launch_bounds__( 256 ) void binaryKernel( int32_t first, int32_t last, View view,
                                                       const int32_t* __restrict__ col, const int32_t* __restrict__ x,
                                                       int32_t* y, int32_t alpha, int32_t beta, Gate gate )
{
   if( ! gate.open() )
      return;
   const int64_t stride = (int64_t) gridDim.x * blockDim.x;
   for( int64_t r = first + (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < last; r += stride ) {
      const int32_t size = view.size( (int32_t) r );
      uint32_t s = 0;  // int32 sum with wrap-around
      for( int32_t l = 0; l < size; l++ ) {
         const int32_t c = col[ view.global( (int32_t) r, l ) ];
         if( c >= 0 )
            s += (uint32_t) x[ c ];
      }
      const int32_t v = (int32_t) s;
      if( beta == 0 )
         y[ r ] = alpha == 1 ? v : (int32_t) ( (uint32_t) alpha * (uint32_t) v );
      else
         y[ r ] = (int32_t) ( (uint32_t) beta * (uint32_t) y[ r ] + ( alpha == 1 ? (uint32_t) v
                                                                                  : (uint32_t) alpha * (uint32_t) v ) );
   }
}

}  // namespace

void spmvBinary( const Matrix& M, const void* x, void* y, double alpha, double beta, int32_t first, int32_t last,
                 Gate gate, hipStream_t st )
{
   if( last <= first )
      return;
   withView( M, [ & ]( auto view ) {
      binaryKernel<<< gridFor( last - first, 256 ), 256, 0, st >>>( first, last, view, M.columns.ptr,
                                                                     (const int32_t*) x, (int32_t*) y,
                                                                     (int32_t) alpha, (int32_t) beta, gate );
   } );
   TNLHIP_CHECK( hipGetLastError() );
}

void spmvSymmetric( const Matrix& M, const void* x, void* y, double alpha, double beta, int32_t first, int32_t last,
                    const void* diag, Gate gate, hipStream_t st )
{
   switch( M.dtype ) {
      case TNLHIP_F64:
         launchSymmetric< double, double, double >( M, (const double*) x, (double*) y, alpha, beta, first, last,
                                                    (const double*) diag, gate, st );
         break;
      case TNLHIP_F32:
         launchSymmetric< float, float, float >( M, (const float*) x, (float*) y, alpha, beta, first, last,
                                                 (const float*) diag, gate, st );
         break;
      case TNLHIP_BF16:
         launchSymmetric< bf16, float, float >( M, (const float*) x, (float*) y, alpha, beta, first, last,
                                                (const float*) diag, gate, st );
         break;
      default: throw Error( TNLHIP_ERROR_INVALID_ARGUMENT, "unknown value type" );
   }
}

}  // namespace tnlhip
