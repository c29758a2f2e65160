// Deterministic single-launch reductions.  Each workgroup of a fixed-size grid folds its
// grid-stride share into one partial, the last workgroup to finish (atomic ticket) sums
// the partials in index order.  The result depends only on the grid size, never on
// timing, so the same inputs give the same bits on every run and in every batch mix.
// Partial sums are kept in float64 whatever the vector type.
#pragma once

#include "common.hpp"

namespace tnlhip {

constexpr int kReduceBlock = 256;
constexpr int kReduceGrid = 2048;  // partials per reduction (8 workgroups per CU on 256 CUs)
constexpr int kMaxMulti = 32;      // columns per multi-reduction pass

// Scratch of one reduction slot: partials, ticket counter.
struct ReduceSlot
{
   double* partials;   // kReduceGrid * width
   unsigned* counter;  // one ticket counter, zero between launches
};

// Called by every thread of the block at the end of a reduction kernel with its own
// partial values v[0..W).  Returns true in the workgroup that performs the final sum;
// in that workgroup thread 0 gets the totals in out[0..W).
template< int W >
__device__ __forceinline__ bool finishReduction( const double ( &v )[ W ], ReduceSlot slot, double ( &out )[ W ] )
{
   __shared__ double lds[ W ][ kReduceBlock / kWave ];
   __shared__ bool last;
   const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
#pragma unroll
   for( int k = 0; k < W; k++ ) {
      const double s = groupSum< kWave >( v[ k ] );
      if( lane == 0 )
         lds[ k ][ wave ] = s;
   }
   __syncthreads();
   if( threadIdx.x == 0 ) {
#pragma unroll
      for( int k = 0; k < W; k++ ) {
         double s = 0;
         for( int w = 0; w < kReduceBlock / kWave; w++ )
            s += lds[ k ][ w ];
         slot.partials[ (int64_t) k * gridDim.x + blockIdx.x ] = s;
      }
      __threadfence();
      const unsigned t = atomicAdd( slot.counter, 1u );
      last = t == gridDim.x - 1;
   }
   __syncthreads();
   if( ! last )
      return false;
   __threadfence();
#pragma unroll
   for( int k = 0; k < W; k++ ) {
      double s = 0;
      for( int j = threadIdx.x; j < (int) gridDim.x; j += kReduceBlock )
         s += slot.partials[ (int64_t) k * gridDim.x + j ];
      s = groupSum< kWave >( s );
      if( lane == 0 )
         lds[ k ][ wave ] = s;
   }
   __syncthreads();
   if( threadIdx.x == 0 ) {
#pragma unroll
      for( int k = 0; k < W; k++ ) {
         double s = 0;
         for( int w = 0; w < kReduceBlock / kWave; w++ )
            s += lds[ k ][ w ];
         out[ k ] = s;
      }
      *slot.counter = 0;
   }
   return true;
}

}  // namespace tnlhip
