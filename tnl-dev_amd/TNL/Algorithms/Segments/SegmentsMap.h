// Device-side index map of a Devices::Hip matrix's segments, for traversal kernels compiled into
// user code (the C++ front's forElements / forRows / reduceRows with device lambdas).  It reads the
// arrays tnlhip_matrix_get_view exports and answers what TNL's segment views answer:
// getSegmentSize( segment ) and getGlobalIndex( segment, localIdx ), column-major organization
// (Algorithms/Segments/{CSRView.hpp, EllpackView.hpp:246-252, SlicedEllpackView.hpp:108-132,
// detail/ChunkedEllpack.h:110-145, detail/BiEllpack.h:59-78, 150-185}).  Needs hipcc.
#pragma once

#ifdef __HIPCC__

#include <hip/hip_runtime.h>

#include "../../Devices/Hip.h"

namespace TNL {
namespace Algorithms {
namespace Segments {

struct SegmentsMap
{
   tnlhip_matrix_view v;

   __device__ int getSegmentSize( int r ) const
   {
      switch( v.format ) {
         case TNLHIP_ELLPACK: return v.segment_size;
         case TNLHIP_SLICED_ELLPACK: return v.slice_segment_sizes[ r / 32 ];
         case TNLHIP_CHUNKED_ELLPACK: {
            int first, chunks, chunkSize, pointer;
            chunk( r, first, chunks, chunkSize, pointer );
            return chunks * chunkSize;
         }
         case TNLHIP_BIELLPACK: {
            const int strip = r / 32, rank = v.row_perm[ r ] - strip * 32;
            const int* gp = v.group_pointers + strip * 6;
            int s = 0;
            for( int g = 0; g < groups( rank ); g++ )
               s += ( gp[ g + 1 ] - gp[ g ] ) >> ( 5 - g );
            return s;
         }
         default: return v.offsets[ r + 1 ] - v.offsets[ r ];
      }
   }

   __device__ long long getGlobalIndex( int r, int l ) const
   {
      switch( v.format ) {
         case TNLHIP_ELLPACK: return r + (long long) v.aligned_size * l;
         case TNLHIP_SLICED_ELLPACK: return v.slice_offsets[ r / 32 ] + r % 32 + 32LL * l;
         case TNLHIP_CHUNKED_ELLPACK: {
            int first, chunks, chunkSize, pointer;
            chunk( r, first, chunks, chunkSize, pointer );
            return pointer + (long long) ( l % chunkSize ) * 256 + first + l / chunkSize;
         }
         case TNLHIP_BIELLPACK: {
            const int strip = r / 32, rank = v.row_perm[ r ] - strip * 32;
            const int* gp = v.group_pointers + strip * 6;
            for( int g = 0; g < groups( rank ); g++ ) {
               const int width = ( gp[ g + 1 ] - gp[ g ] ) >> ( 5 - g );
               if( l < width )
                  return gp[ g ] + rank + (long long) l * ( 32 >> g );
               l -= width;
            }
            return -1;
         }
         default: return (long long) v.offsets[ r ] + l;
      }
   }

private:
   // groups a BiEllpack row of this rank spans (detail/BiEllpack.h:59-78)
   __device__ static int groups( int rank ) { return 6 - ( 32 - __clz( rank ) ); }
   __device__ void chunk( int r, int& first, int& chunks, int& chunkSize, int& pointer ) const
   {
      const int sl = v.row_to_slice[ r ];
      const int* s = v.slices + 4 * sl;  // size, chunkSize, firstSegment, pointer
      first = r != s[ 2 ] ? v.row_to_chunk[ r - 1 ] : 0;
      chunks = v.row_to_chunk[ r ] - first;
      chunkSize = s[ 1 ];
      pointer = s[ 3 ];
   }
};

}  // namespace Segments
}  // namespace Algorithms
}  // namespace TNL

#endif  // __HIPCC__
