// Row distribution of the padded-format kernels (Ellpack, SlicedEllpack): the launch geometry
// chosen on the host and the rows each lane of that launch takes.  Written for host and device
// alike, so that the kernels and a host program that checks every row is covered exactly once
// and every slot stays inside the storage (tests/cpp/test_rowdist.cpp) run the same code.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

namespace tnlhip {

constexpr int kXcds = 8;

// blockIdx.x, gridDim.x, blockDim.x, threadIdx.x of one lane
struct LaunchIndex
{
   int64_t block, grid, threads, thread;
};

// Workgroup -> row block mapping.  Workgroups are dispatched round-robin over the 8 XCDs;
// with `remap` the logical index puts consecutive logical blocks on the same XCD, so each
// XCD streams one contiguous eighth of the rows and x's neighbourhoods stay in its own L2.
// The grid must be a multiple of kXcds when remap is set.
__host__ __device__ inline int64_t logicalBlockOf( int64_t block, int64_t grid, int remap )
{
   if( ! remap )
      return block;
   const int64_t per = grid / kXcds;
   return ( block % kXcds ) * per + block / kXcds;
}

// Rows start, start + step, ... below end, P consecutive rows at a time.
struct RowSeq
{
   int64_t start, end, step;
};

// The row sequence of a lane.  dist 0: grid-stride over the whole range; 1: each workgroup owns
// `per` consecutive rows, logical blocks remapped so an XCD owns a contiguous eighth; 3: each XCD
// grid-strides over its own contiguous eighth.  No other value reaches a kernel (the host checks).
template< int P >
__host__ __device__ inline RowSeq rowSequenceAt( int64_t base, int64_t last, int64_t per, int dist, LaunchIndex li )
{
   if( dist == 0 )
      return RowSeq{ base + ( li.block * li.threads + li.thread ) * P, last, li.grid * li.threads * P };
   if( dist == 3 ) {
      // XCD x (block % 8) sweeps the x-th eighth of the rows with a grid-stride loop of its own,
      // so the x entries a row band gathers stay in that XCD's L2 while the sweep passes
      const int64_t perXcd = li.grid / kXcds;
      const int64_t pos = li.block / kXcds;
      if( pos >= perXcd )
         return RowSeq{ last, last, 1 };
      const int64_t chunk = ( ( last - base + kXcds - 1 ) / kXcds + P - 1 ) / P * P;
      const int64_t r0 = base + ( li.block % kXcds ) * chunk;
      const int64_t e = r0 + chunk < last ? r0 + chunk : last;
      return RowSeq{ r0 + ( pos * li.threads + li.thread ) * P, e, perXcd * li.threads * P };
   }
   const int64_t r0 = base + logicalBlockOf( li.block, li.grid, 1 ) * per;
   const int64_t e = r0 + per < last ? r0 + per : last;
   return RowSeq{ r0 + li.thread * P, e, li.threads * P };
}

// The distributions the padded kernels implement.
inline bool knownDistribution( int dist ) { return dist == 0 || dist == 1 || dist == 3; }

// Launch of the padded kernels over rows [first, last): 256-lane workgroups, one or two (pair)
// rows per lane, enough workgroups for blocksPerCu per CU (no more than the rows need), rounded
// up to a multiple of the XCD count; `per` rows per workgroup for distribution 1.
struct PaddedLaunch
{
   int64_t grid, per, base;
};

inline PaddedLaunch paddedLaunch( int64_t first, int64_t last, bool pair, int blocksPerCu, int cus )
{
   const int64_t rowsPerThread = pair ? 2 : 1;
   const int64_t unit = 256 * rowsPerThread;
   const int64_t rows = last - first;
   const int64_t need = ( rows + unit - 1 ) / unit;
   int64_t grid = std::max< int64_t >( 1, std::min< int64_t >( need, (int64_t) cus * blocksPerCu ) );
   grid = ( grid + kXcds - 1 ) / kXcds * kXcds;
   const int64_t base = pair ? ( first & ~(int64_t) 1 ) : first;
   const int64_t span = last - base;
   const int64_t per = ( ( span + grid - 1 ) / grid + unit - 1 ) / unit * unit;
   return PaddedLaunch{ grid, per, base };
}

}  // namespace tnlhip
