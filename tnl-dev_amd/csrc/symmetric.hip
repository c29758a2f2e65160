// Products of symmetric and of binary matrices, one lane per row over the format's index map.
//
// Symmetric: SparseMatrix< ..., SymmetricMatrix, ... >::vectorProduct
// (Matrices/SparseMatrixView.hpp:390-491 with MatrixType::isSymmetric()).  The matrix stores its
// lower triangle, diagonal included.  When beta != 1 the whole output vector is scaled by beta
// first (TNL multiplies rather than clears, so a NaN in y survives beta = 0).  Then every stored
// a_rc (c != -1) of a row r in [first, last) adds (alpha a_rc) x_r to y_c when c < r -- the
// transposed entry -- and the row sum alpha * sum_c a_rc x_c is added to y_r.  The additions are
// floating-point atomics, as on TNL's GPU path (in LDS for rows of the wave's own block, then
// one device atomic per row; device atomics for the rest); the row sum keeps TNL's Host order
// (each product rounded, then added), the order in which the atomics land is not fixed.  One
// lane per row over the format's index map; the solvers' Jacobi keeper divides by the diagonal
// afterwards.
#include "views.hpp"

namespace tnlhip {

namespace {

// Each wavefront takes kSymRows consecutive rows (kSymRows / 64 per lane).  Additions that land
// on a row of the wave's own block (the row sums, and transposed entries a_rc with c in the
// block, e.g. c = r - 1 and c = r - 160 of a 3D stencil) meet in the wave's LDS slice through
// LDS atomics; after a wave barrier every row of the block takes one device atomic.  Transposed
// entries that leave the block go to y directly.  At 160^3 that is about 2.6 device atomics per
// row instead of 4.  Wave barriers only (no workgroup barrier): the loop over blocks is uniform
// per wave.
constexpr int kSymRows = 256;
constexpr int kSlotChunk = 8;  // slots of a row whose loads are in flight together

__device__ __forceinline__ void waveBarrier()
{
   __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
   __builtin_amdgcn_wave_barrier();
   __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
}

template< typename View, typename V, typename A, typename Y >
__global__ __launch_bounds__( 256 ) void symmetricKernel( int32_t first, int32_t last, View view,
                                                          const int32_t* __restrict__ col, const V* __restrict__ val,
                                                          const Y* __restrict__ x, Y* y, A alpha, Gate gate )
{
   if( ! gate.open() )
      return;
   constexpr int kPerLane = kSymRows / kWave;
   __shared__ A acc[ 4 ][ kSymRows ];
   const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
   const int64_t step = (int64_t) gridDim.x * 4 * kSymRows;
   for( int64_t w0 = first + ( (int64_t) blockIdx.x * 4 + wave ) * kSymRows; w0 < last; w0 += step ) {
#pragma unroll
      for( int k = 0; k < kPerLane; k++ )
         acc[ wave ][ lane + k * kWave ] = 0;
      waveBarrier();
#pragma unroll
      for( int k = 0; k < kPerLane; k++ ) {
         const int64_t r = w0 + lane + k * kWave;
         if( r >= last )
            continue;
         const A xr = (A) x[ r ];
         const int32_t size = view.size( (int32_t) r );
         A s = 0;
         for( int32_t l0 = 0; l0 < size; l0 += kSlotChunk ) {
            // the chunk's column, value and x loads go out together (slots past the row re-read
            // its last slot, padding reads x[0]: a symmetric matrix with rows has columns), then
            // the slots are taken left to right
            int32_t c[ kSlotChunk ];
            A a[ kSlotChunk ], xc[ kSlotChunk ];
#pragma unroll
            for( int k = 0; k < kSlotChunk; k++ ) {
               const int64_t g = view.global( (int32_t) r, min( l0 + k, size - 1 ) );
               c[ k ] = col[ g ];
               a[ k ] = toCompute( val[ g ], A() );
            }
#pragma unroll
            for( int k = 0; k < kSlotChunk; k++ ) {
               c[ k ] = l0 + k < size ? c[ k ] : -1;
               xc[ k ] = (A) x[ c[ k ] >= 0 ? c[ k ] : 0 ];
            }
#pragma unroll
            for( int k = 0; k < kSlotChunk; k++ ) {
               if( c[ k ] < 0 )
                  continue;
               if( c[ k ] < r ) {
                  const A t = ( alpha * a[ k ] ) * xr;
                  if( c[ k ] >= w0 )
                     atomicAdd( &acc[ wave ][ c[ k ] - w0 ], t );
                  else
                     unsafeAtomicAdd( y + c[ k ], (Y) t );
               }
               s = mulAddExact( s, a[ k ], xc[ k ] );
            }
         }
         atomicAdd( &acc[ wave ][ r - w0 ], alpha * s );
      }
      waveBarrier();
#pragma unroll
      for( int k = 0; k < kPerLane; k++ ) {
         const int64_t r = w0 + lane + k * kWave;
         if( r < last )
            unsafeAtomicAdd( y + r, (Y) acc[ wave ][ lane + k * kWave ] );
      }
      waveBarrier();  // the block's sums are read before the next block clears them
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
         symmetricKernel<<< gridFor( ( (int64_t) last - first + kSymRows - 1 ) / kSymRows, 4 ), 256, 0, st >>>(
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
                                                       int32_t* y, int32_t alpha, int32_t beta, int32_t ncols,
                                                       Gate gate )
{
   if( ! gate.open() )
      return;
   const int64_t stride = (int64_t) gridDim.x * blockDim.x;
   for( int64_t r = first + (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < last; r += stride ) {
      // no columns: every slot is padding and x may be empty
      const int32_t size = ncols > 0 ? view.size( (int32_t) r ) : 0;
      uint32_t s = 0;  // int32 sum with wrap-around (any order gives the same bits)
      for( int32_t l0 = 0; l0 < size; l0 += kSlotChunk ) {
         // all loads of a chunk issued together: slots past the row re-read its last slot and
         // padding reads x[0]; the select drops both
         int32_t c[ kSlotChunk ];
#pragma unroll
         for( int k = 0; k < kSlotChunk; k++ )
            c[ k ] = col[ view.global( (int32_t) r, min( l0 + k, size - 1 ) ) ];
#pragma unroll
         for( int k = 0; k < kSlotChunk; k++ ) {
            const bool on = l0 + k < size && c[ k ] >= 0;
            const uint32_t xv = (uint32_t) x[ on ? c[ k ] : 0 ];
            s += on ? xv : 0u;
         }
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
                                                                     (int32_t) alpha, (int32_t) beta, M.cols, gate );
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
