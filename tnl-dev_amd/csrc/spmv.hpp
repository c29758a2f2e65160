// Pieces of the sparse matrix-vector product shared by the kernels of spmv.hip and the
// column-striped plan (stripes.hip): the result keeper and the wave-local LDS handoff.
#pragma once

#include "views.hpp"

namespace tnlhip {

// ------------------------------------------------------------------ result keeper

// SparseMatrixView.hpp:447-490: direct / alpha / beta / general keepers; mode 4 divides by
// a diagonal, which fuses Preconditioners::Diagonal::solve (x = b / diag) into the product;
// mode 5 is one step of the Jacobi splitting, y = prev + (rhs - v) / diag (the Jacobi-polynomial
// preconditioner; y must not alias prev, which other rows gather).
template< typename A, typename Y >
struct Keeper
{
   Y* y;
   A alpha, beta;
   const Y* diag;
   int mode;
   const int32_t* gate;  // solver flags: the launch is skipped when gate[0] != 0 (and, with
   int gateMode;         // gateMode 2, unless gate[1] != 0)
   // y stored with non-temporal stores (the default: y is written once, L2 stays for x); kernels
   // whose waves end rows inside a cache line store through L2, which merges the partial lines
   bool streamY = true;
   const Y* rhs = nullptr;   // mode 5
   const Y* prev = nullptr;  // mode 5
   __device__ __forceinline__ bool closed() const
   {
      if( gateMode == 0 )
         return false;
      if( __builtin_nontemporal_load( gate ) )
         return true;
      return gateMode == 2 && ! __builtin_nontemporal_load( gate + 1 );
   }
   __device__ __forceinline__ void operator()( int64_t row, A v ) const
   {
      A out;
      switch( mode ) {
         case 0: out = v; break;
         case 1: out = alpha * v; break;
         case 2: out = beta * (A) y[ row ] + v; break;
         case 3: out = beta * (A) y[ row ] + alpha * v; break;
         case 4: out = v / (A) diag[ row ]; break;
         default: out = (A) prev[ row ] + ( (A) rhs[ row ] - v ) / (A) diag[ row ]; break;
      }
      if( streamY )
         __builtin_nontemporal_store( (Y) out, y + row );  // y is written once: keep L2 for x
      else
         y[ row ] = (Y) out;
   }
};

// Wave-local LDS handoff between lanes of one wavefront (no workgroup barrier).
__device__ __forceinline__ void waveSync()
{
   __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
   __builtin_amdgcn_wave_barrier();
   __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
}

// Column-striped product of a CSR matrix with a plan (stripes.hip).
template< typename V, typename A, typename Y >
void stripeLaunch( const Matrix& M, const Y* x, Keeper< A, Y > keep, int32_t first, int32_t last, hipStream_t st );

}  // namespace tnlhip
