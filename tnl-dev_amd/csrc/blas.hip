// Reductions and BLAS-1 kernels of the Krylov path: (a, b), lpNorm(a, 2), y = ax + by,
// GMRES's multi-reduction V^T w and tall-skinny gemv, and the int32 exclusive scan used for
// CSR offsets.  All HBM-bound; every reduction is deterministic (reduce.hpp).
#include "blas.hpp"

namespace tnlhip {

// ------------------------------------------------------------------ dot / norm

template< typename T, bool SQUARE >
__global__ __launch_bounds__( kReduceBlock ) void dotKernel( int64_t n, const T* __restrict__ a,
                                                             const T* __restrict__ b, ReduceSlot slot )
{
   // grid-stride, two elements in flight per lane
   double acc[ 1 ] = { 0.0 };
   const int64_t stride = (int64_t) gridDim.x * kReduceBlock;
   int64_t i = (int64_t) blockIdx.x * kReduceBlock + threadIdx.x;
   for( ; i + stride < n; i += 2 * stride ) {
      const T a0 = a[ i ], a1 = a[ i + stride ];
      const T b0 = SQUARE ? a0 : b[ i ], b1 = SQUARE ? a1 : b[ i + stride ];
      acc[ 0 ] += (double) a0 * (double) b0;
      acc[ 0 ] += (double) a1 * (double) b1;
   }
   if( i < n ) {
      const double x = (double) a[ i ];
      acc[ 0 ] += SQUARE ? x * x : x * (double) b[ i ];
   }
   storeBlockPartials< 1 >( acc, slot );
}

// result = sum of the partials (or its square root)
__global__ __launch_bounds__( kReduceBlock ) void dotFinalKernel( ReduceSlot slot, int grid, double* __restrict__ result,
                                                                  int sqrtResult )
{
   __shared__ double tot[ 1 ];
   sumPartials< 1 >( slot, grid, tot );
   if( threadIdx.x == 0 )
      *result = sqrtResult ? sqrt( tot[ 0 ] ) : tot[ 0 ];
}

void deviceDot( int dtype, int64_t n, const void* a, const void* b, double* resultDevice, bool sqrtResult,
                Workspace& ws, hipStream_t st )
{
   ReduceSlot slot = ws.slot( 0 );
   const unsigned grid = reduceGridFor( 1, n );
   const bool sq = a == b;
   if( dtype == TNLHIP_F64 ) {
      if( sq )
         dotKernel< double, true ><<< grid, kReduceBlock, 0, st >>>( n, (const double*) a, (const double*) b, slot );
      else
         dotKernel< double, false ><<< grid, kReduceBlock, 0, st >>>( n, (const double*) a, (const double*) b, slot );
   }
   else {
      if( sq )
         dotKernel< float, true ><<< grid, kReduceBlock, 0, st >>>( n, (const float*) a, (const float*) b, slot );
      else
         dotKernel< float, false ><<< grid, kReduceBlock, 0, st >>>( n, (const float*) a, (const float*) b, slot );
   }
   dotFinalKernel<<< 1, kReduceBlock, 0, st >>>( slot, grid, resultDevice, sqrtResult ? 1 : 0 );
   TNLHIP_CHECK( hipGetLastError() );
}

// ------------------------------------------------------------------ axpby

template< typename T >
__global__ __launch_bounds__( 256 ) void axpbyKernel( int64_t n, double alpha, const T* __restrict__ x, double beta,
                                                      T* __restrict__ y )
{
   const int64_t stride = (int64_t) gridDim.x * blockDim.x;
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride ) {
      if( beta == 0.0 )
         y[ i ] = (T) ( alpha * (double) x[ i ] );
      else
         y[ i ] = (T) ( alpha * (double) x[ i ] + beta * (double) y[ i ] );
   }
}

void deviceAxpby( int dtype, int64_t n, double alpha, const void* x, double beta, void* y, hipStream_t st )
{
   if( n <= 0 )
      return;
   if( dtype == TNLHIP_F64 )
      axpbyKernel< double ><<< gridFor( n, 256 ), 256, 0, st >>>( n, alpha, (const double*) x, beta, (double*) y );
   else
      axpbyKernel< float ><<< gridFor( n, 256 ), 256, 0, st >>>( n, alpha, (const float*) x, beta, (float*) y );
   TNLHIP_CHECK( hipGetLastError() );
}

// ------------------------------------------------------------------ multi-reduction V^T w

// Every row i is read once: w[i] and the `count` entries V[i + k*ld]; each lane keeps
// `count` running sums (count <= 32).  Partials per workgroup and column, fixed-order finish.
template< typename T, int C >
__global__ __launch_bounds__( kReduceBlock ) void multidotKernel( int64_t n, int32_t count, const T* __restrict__ V,
                                                                  int64_t ld, const T* __restrict__ w,
                                                                  ReduceSlot slot )
{
   // all C column loads of a row issue without branches; columns past `count` repeat the
   // last one (a cache hit) and are multiplied by zero
   double acc[ C ];
#pragma unroll
   for( int k = 0; k < C; k++ )
      acc[ k ] = 0.0;
   const int64_t stride = (int64_t) gridDim.x * kReduceBlock;
   for( int64_t i = (int64_t) blockIdx.x * kReduceBlock + threadIdx.x; i < n; i += stride ) {
      const double wi = (double) w[ i ];
      T v[ C ];
#pragma unroll
      for( int k = 0; k < C; k++ )
         v[ k ] = V[ i + (int64_t) min( k, count - 1 ) * ld ];
#pragma unroll
      for( int k = 0; k < C; k++ )
         acc[ k ] += ( k < count ? (double) v[ k ] : 0.0 ) * wi;
   }
   storeBlockPartials< C >( acc, slot );
}

template< int C >
__global__ __launch_bounds__( kReduceBlock ) void multidotFinalKernel( ReduceSlot slot, int grid, int32_t count,
                                                                       double* __restrict__ result )
{
   __shared__ double tot[ C ];
   sumPartials< C >( slot, grid, tot );
   if( threadIdx.x < count )
      result[ threadIdx.x ] = tot[ threadIdx.x ];
}

template< typename T, int C >
void launchMultidot( int64_t n, int32_t c, const T* V, int64_t ld, const T* w, ReduceSlot slot, double* result,
                     hipStream_t st )
{
   const int grid = reduceGridFor( C, n );
   multidotKernel< T, C ><<< grid, kReduceBlock, 0, st >>>( n, c, V, ld, w, slot );
   multidotFinalKernel< C ><<< 1, kReduceBlock, 0, st >>>( slot, grid, c, result );
}

template< typename T >
void launchMultidotBucket( int64_t n, int32_t c, const T* V, int64_t ld, const T* w, ReduceSlot slot, double* result,
                           hipStream_t st )
{
   if( c <= 8 )
      launchMultidot< T, 8 >( n, c, V, ld, w, slot, result, st );
   else if( c <= 16 )
      launchMultidot< T, 16 >( n, c, V, ld, w, slot, result, st );
   else
      launchMultidot< T, 32 >( n, c, V, ld, w, slot, result, st );
}

// count <= kMaxMulti columns into resultDevice with the partials in `slot` (kMaxMulti * kReduceGrid doubles)
void multidotSlot( int dtype, int64_t n, int32_t c, const void* V, int64_t ld, const void* w, double* resultDevice,
                   ReduceSlot slot, hipStream_t st )
{
   if( dtype == TNLHIP_F64 )
      launchMultidotBucket< double >( n, c, (const double*) V, ld, (const double*) w, slot, resultDevice, st );
   else
      launchMultidotBucket< float >( n, c, (const float*) V, ld, (const float*) w, slot, resultDevice, st );
   TNLHIP_CHECK( hipGetLastError() );
}

void deviceMultidot( int dtype, int64_t n, int32_t count, const void* V, int64_t ld, const void* w,
                     double* resultDevice, Workspace& ws, hipStream_t st )
{
   for( int32_t k0 = 0; k0 < count; k0 += kMaxMulti ) {
      const int32_t c = std::min( kMaxMulti, count - k0 );
      ReduceSlot slot = ws.slot( 0 );
      if( dtype == TNLHIP_F64 )
         launchMultidotBucket< double >( n, c, (const double*) V + k0 * ld, ld, (const double*) w, slot,
                                         resultDevice + k0, st );
      else
         launchMultidotBucket< float >( n, c, (const float*) V + k0 * ld, ld, (const float*) w, slot,
                                        resultDevice + k0, st );
      TNLHIP_CHECK( hipGetLastError() );
   }
}

// ------------------------------------------------------------------ gemv

// y = alpha * V c + beta * y, V column-major n x count, c on the device (count <= 64).
template< typename T >
__global__ __launch_bounds__( 256 ) void gemvKernel( int64_t n, int32_t count, double alpha, const T* __restrict__ V,
                                                     int64_t ld, const double* __restrict__ c, double beta,
                                                     T* __restrict__ y )
{
   __shared__ double cs[ 64 ];
   if( threadIdx.x < count )
      cs[ threadIdx.x ] = alpha * c[ threadIdx.x ];
   __syncthreads();
   const int64_t stride = (int64_t) gridDim.x * blockDim.x;
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride ) {
      double s = 0.0;
      for( int32_t k = 0; k < count; k++ )
         s += (double) V[ i + k * ld ] * cs[ k ];
      y[ i ] = (T) ( beta == 0.0 ? s : s + beta * (double) y[ i ] );
   }
}

void deviceGemv( int dtype, int64_t n, int32_t count, double alpha, const void* V, int64_t ld,
                 const double* coefDevice, double beta, void* y, hipStream_t st )
{
   if( n <= 0 || count <= 0 )
      return;
   TNLHIP_REQUIRE( count <= 64, "gemv supports at most 64 columns per call" );
   if( dtype == TNLHIP_F64 )
      gemvKernel< double ><<< gridFor( n, 256 ), 256, 0, st >>>( n, count, alpha, (const double*) V, ld, coefDevice,
                                                                 beta, (double*) y );
   else
      gemvKernel< float ><<< gridFor( n, 256 ), 256, 0, st >>>( n, count, alpha, (const float*) V, ld, coefDevice,
                                                                beta, (float*) y );
   TNLHIP_CHECK( hipGetLastError() );
}

// ------------------------------------------------------------------ exclusive scan (int32)

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

// Block-wide exclusive scan of one value per thread; returns the block total in *total.
__device__ int32_t blockExclusiveScan( int32_t v, int32_t* total )
{
   __shared__ int32_t waveSums[ kScanBlock / kWave ];
   const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
   int32_t incl = v;
#pragma unroll
   for( int o = 1; o < kWave; o <<= 1 ) {
      const int32_t t = __shfl_up( incl, o, kWave );
      if( lane >= o )
         incl += t;
   }
   if( lane == kWave - 1 )
      waveSums[ wave ] = incl;
   __syncthreads();
   int32_t before = 0, all = 0;
   for( int w = 0; w < kScanBlock / kWave; w++ ) {
      if( w < wave )
         before += waveSums[ w ];
      all += waveSums[ w ];
   }
   __syncthreads();
   *total = all;
   return before + incl - v;
}

__global__ __launch_bounds__( kScanBlock ) void scanTileSums( int64_t n, const int32_t* __restrict__ in,
                                                              int32_t* __restrict__ sums )
{
   const int64_t base = (int64_t) blockIdx.x * kScanTile + threadIdx.x * kScanItems;
   int32_t s = 0;
   for( int j = 0; j < kScanItems; j++ )
      if( base + j < n )
         s += in[ base + j ];
   int32_t total;
   blockExclusiveScan( s, &total );
   if( threadIdx.x == 0 )
      sums[ blockIdx.x ] = total;
}

// one workgroup: exclusive scan of the tile sums, carried over chunks of kScanBlock
__global__ __launch_bounds__( kScanBlock ) void scanSums( int64_t tiles, int32_t* sums, int32_t* grand )
{
   int32_t carry = 0;
   for( int64_t b = 0; b < tiles; b += kScanBlock ) {
      const int64_t i = b + threadIdx.x;
      const int32_t v = i < tiles ? sums[ i ] : 0;
      int32_t total;
      const int32_t ex = blockExclusiveScan( v, &total );
      if( i < tiles )
         sums[ i ] = carry + ex;
      carry += total;
   }
   if( threadIdx.x == 0 && grand )
      *grand = carry;
}

__global__ __launch_bounds__( kScanBlock ) void scanApply( int64_t n, const int32_t* __restrict__ in,
                                                           int32_t* __restrict__ out,
                                                           const int32_t* __restrict__ sums )
{
   const int64_t base = (int64_t) blockIdx.x * kScanTile + threadIdx.x * kScanItems;
   int32_t v[ kScanItems ];
   int32_t s = 0;
   for( int j = 0; j < kScanItems; j++ ) {
      v[ j ] = base + j < n ? in[ base + j ] : 0;
      s += v[ j ];
   }
   int32_t total;
   int32_t run = blockExclusiveScan( s, &total ) + sums[ blockIdx.x ];
   for( int j = 0; j < kScanItems; j++ )
      if( base + j < n ) {
         out[ base + j ] = run;
         run += v[ j ];
      }
}

void deviceExclusiveScan( int64_t n, const int32_t* in, int32_t* out, int32_t* totalDevice, Workspace& ws,
                          hipStream_t st )
{
   if( n <= 0 ) {
      if( totalDevice )
         TNLHIP_CHECK( hipMemsetAsync( totalDevice, 0, sizeof( int32_t ), st ) );
      return;
   }
   const int64_t tiles = ( n + kScanTile - 1 ) / kScanTile;
   int32_t* sums = ws.scanScratch( tiles );
   scanTileSums<<< (unsigned) tiles, kScanBlock, 0, st >>>( n, in, sums );
   scanSums<<< 1, kScanBlock, 0, st >>>( tiles, sums, totalDevice );
   scanApply<<< (unsigned) tiles, kScanBlock, 0, st >>>( n, in, out, sums );
   TNLHIP_CHECK( hipGetLastError() );
}

}  // namespace tnlhip
