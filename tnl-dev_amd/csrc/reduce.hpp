// Deterministic reductions in two launches.  The first kernel (a fixed grid of kReduceGrid
// workgroups) folds each workgroup's grid-stride share into one partial per column; the
// second, one workgroup, sums the partials in index order.  The result depends only on the
// grid size, never on timing, so the same inputs give the same bits on every run and in
// every batch mix.  Partial sums are kept in float64 whatever the vector type.
//
// Two launches rather than a last-workgroup ticket: on this GPU the eight XCDs have private
// L2 caches, so a ticket needs a device-scope release (an L2 write-back) in every workgroup,
// which costs more than the whole pass over a 2-4 M vector; the kernel boundary gives the
// same ordering once per launch.
#pragma once

#include "common.hpp"

#include <algorithm>

namespace tnlhip {

constexpr int kReduceBlock = 256;
constexpr int kReduceGrid = 1024;  // partials per column (4 workgroups per CU on 256 CUs)
constexpr int kMaxMulti = 32;      // columns per multi-reduction pass

// Workgroups of a pass keeping W running sums per lane over n entries.  The 32-column passes
// need about 140 VGPRs (3 resident workgroups per CU), so their grid is one resident round
// instead of 1024 workgroups in a full and a one-third round.  One- and two-column passes over
// short vectors (n < 1024 x 256) launch only the ceil(n / 256) workgroups that hold an entry
// (one per thread) instead of 1024 mostly empty ones: the partials that remain are the same
// and sumPartials adds them in the same order, so the sums keep their bits.  The grid (hence
// the sum order) depends only on W and n.
inline int reduceGridFor( int W, int64_t n )
{
   const int full = W >= 32 ? std::min( kReduceGrid, 3 * deviceProps().cus ) : kReduceGrid;
   if( W > 2 )
      return full;
   const int64_t need = std::max< int64_t >( 1, ( n + kReduceBlock - 1 ) / kReduceBlock );
   return (int) std::min< int64_t >( full, need );
}

// Scratch of one reduction slot: kReduceGrid partials per column, column-major.
struct ReduceSlot
{
   double* partials;
};

// Called by every thread of a kReduceBlock workgroup with its own values v[0..W); the
// workgroup's W sums are stored at partials[k * gridDim.x + blockIdx.x].
template< int W >
__device__ __forceinline__ void storeBlockPartials( const double ( &v )[ W ], ReduceSlot slot )
{
   __shared__ double lds[ W ][ kReduceBlock / kWave ];
   const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
#pragma unroll
   for( int k = 0; k < W; k++ ) {
      const double s = groupSum< kWave >( v[ k ] );
      if( lane == 0 )
         lds[ k ][ wave ] = s;
   }
   __syncthreads();
   if( threadIdx.x < W ) {
      const int k = threadIdx.x;
      double s = 0;
      for( int w = 0; w < kReduceBlock / kWave; w++ )
         s += lds[ k ][ w ];
      slot.partials[ (int64_t) k * gridDim.x + blockIdx.x ] = s;
   }
}

// One kReduceBlock workgroup: out[k] = sum over the `grid` partials of column k, in a fixed
// order; out (W doubles, usually LDS) is valid in every thread on return.  Up to 64 columns
// get kReduceBlock / W lanes each, one column gets the block.  The partials were written by
// the previous launch, possibly through another XCD's L2, so every load is a trip to memory:
// each lane keeps kSumStreams independent loads in flight (a 32-column sum of 768 partials is
// 6 dependent trips instead of 24) and folds its accumulators in a fixed tree.
constexpr int kSumStreams = 16;

template< int W >
__device__ __forceinline__ void sumPartials( ReduceSlot slot, int grid, double* out )
{
   if constexpr( W >= 4 ) {
      static_assert( ( W & ( W - 1 ) ) == 0 && W <= kReduceBlock, "column count must be a power of two" );
      constexpr int G = kReduceBlock / W;  // lanes per column, a power of two <= 64
      const int k = threadIdx.x / G, g = threadIdx.x % G;
      const double* p = slot.partials + (int64_t) k * grid;
      double a[ kSumStreams ];
#pragma unroll
      for( int u = 0; u < kSumStreams; u++ )
         a[ u ] = 0;
      int j = g;
      for( ; j + ( kSumStreams - 1 ) * G < grid; j += kSumStreams * G ) {
#pragma unroll
         for( int u = 0; u < kSumStreams; u++ )
            a[ u ] += p[ j + u * G ];
      }
      for( ; j < grid; j += G )
         a[ 0 ] += p[ j ];
#pragma unroll
      for( int w = kSumStreams / 2; w >= 1; w /= 2 ) {
#pragma unroll
         for( int u = 0; u < w; u++ )
            a[ u ] += a[ u + w ];
      }
      double s = a[ 0 ];
      s = groupSum< G >( s );
      if( g == 0 )
         out[ k ] = s;
      __syncthreads();
   }
   else {
      __shared__ double lds[ W ][ kReduceBlock / kWave ];
      const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
#pragma unroll
      for( int k = 0; k < W; k++ ) {
         const double* p = slot.partials + (int64_t) k * grid;
         double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
         int j = threadIdx.x;
         for( ; j + 3 * kReduceBlock < grid; j += 4 * kReduceBlock ) {
            a0 += p[ j ];
            a1 += p[ j + kReduceBlock ];
            a2 += p[ j + 2 * kReduceBlock ];
            a3 += p[ j + 3 * kReduceBlock ];
         }
         for( ; j < grid; j += kReduceBlock )
            a0 += p[ j ];
         double s = ( a0 + a1 ) + ( a2 + a3 );
         s = groupSum< kWave >( s );
         if( lane == 0 )
            lds[ k ][ wave ] = s;
      }
      __syncthreads();
      if( threadIdx.x < W ) {
         const int k = threadIdx.x;
         double s = 0;
         for( int w = 0; w < kReduceBlock / kWave; w++ )
            s += lds[ k ][ w ];
         out[ k ] = s;
      }
      __syncthreads();
   }
}

}  // namespace tnlhip
