// Segment types usable as the Segments parameter of Matrices::SparseMatrix on Devices::Hip.
// Names follow src/TNL/Algorithms/Segments/{CSR.h:577-605, Ellpack.h, SlicedEllpack.h,
// ChunkedEllpack.h, BiEllpack.h}; each carries the tnlhip format it selects.  Padded formats
// use TNL's GPU defaults: column-major organization, Ellpack alignment 32, slice size 32,
// 256 chunks per slice, BiEllpack strips of 32.
#pragma once

#include "../../Devices/Hip.h"

namespace TNL {
namespace Algorithms {
namespace Segments {

#define TNLHIP_SEGMENTS_TAG( Name, Format )                                             \
   template< typename Device, typename Index, typename IndexAllocator = void >         \
   struct Name                                                                         \
   {                                                                                   \
      static constexpr int format = Format;                                            \
      static constexpr const char* name() { return #Name; }                            \
   };

TNLHIP_SEGMENTS_TAG( CSRScalar, TNLHIP_CSR_SCALAR )
TNLHIP_SEGMENTS_TAG( CSRVector, TNLHIP_CSR_VECTOR )
TNLHIP_SEGMENTS_TAG( CSRLight, TNLHIP_CSR_LIGHT )
TNLHIP_SEGMENTS_TAG( CSRHybrid, TNLHIP_CSR_HYBRID )
TNLHIP_SEGMENTS_TAG( CSRAdaptive, TNLHIP_CSR_ADAPTIVE )
TNLHIP_SEGMENTS_TAG( Ellpack, TNLHIP_ELLPACK )
TNLHIP_SEGMENTS_TAG( SlicedEllpack, TNLHIP_SLICED_ELLPACK )
TNLHIP_SEGMENTS_TAG( ChunkedEllpack, TNLHIP_CHUNKED_ELLPACK )
TNLHIP_SEGMENTS_TAG( BiEllpack, TNLHIP_BIELLPACK )

#undef TNLHIP_SEGMENTS_TAG

// CSR.h:605: the default CSR kernel is the scalar one
template< typename Device, typename Index, typename IndexAllocator = void >
using CSRDefault = CSRScalar< Device, Index, IndexAllocator >;

}  // namespace Segments
}  // namespace Algorithms
}  // namespace TNL
