// Device-side index maps of the segment formats: for a segment (row) and a local index,
// where its element lives in the values / column-index arrays, and how many slots the
// segment has.  Column-major organization throughout (TNL's GPU default).
//   CSR            offsets[row] + local                            (CSRView::getGlobalIndex)
//   Ellpack        row + alignedSize * local                       (EllpackView.hpp:246-252)
//   SlicedEllpack  sliceOffsets[row/S] + row%S + S*local, S the slice size (32 by default)
//                                                                  (SlicedEllpackView.hpp:108-132)
//   ChunkedEllpack pointer + (local % chunkSize)*256 + firstChunk + local / chunkSize
//                                                                  (detail/ChunkedEllpack.h:110-145)
//   BiEllpack      groupPointer + rank + local'*height, walking the strip's groups
//                                                                  (detail/BiEllpack.h:150-185)
#pragma once

#include "matrix.hpp"

namespace tnlhip {

struct CsrView
{
   const int32_t* offsets;
   __device__ int32_t size( int32_t r ) const { return offsets[ r + 1 ] - offsets[ r ]; }
   __device__ int64_t global( int32_t r, int32_t l ) const { return (int64_t) offsets[ r ] + l; }
};

struct EllpackView
{
   int32_t segmentSize, alignedSize;
   __device__ int32_t size( int32_t ) const { return segmentSize; }
   __device__ int64_t global( int32_t r, int32_t l ) const { return r + (int64_t) alignedSize * l; }
};

struct SlicedView
{
   const int32_t* sliceOffsets;
   const int32_t* sliceSizes;
   int32_t logSlice;  // log2 of the slice size
   __device__ int32_t size( int32_t r ) const { return sliceSizes[ r >> logSlice ]; }
   __device__ int64_t global( int32_t r, int32_t l ) const
   {
      return sliceOffsets[ r >> logSlice ] + ( r & ( ( 1 << logSlice ) - 1 ) ) + ( (int64_t) l << logSlice );
   }
};

inline int32_t log2Exact( int32_t v )
{
   int32_t k = 0;
   while( ( 1 << k ) < v )
      k++;
   return k;
}

struct ChunkedView
{
   const int32_t* rowToChunk;
   const int32_t* rowToSlice;
   const int32_t* slices;  // size, chunkSize, firstSegment, pointer
   __device__ void locate( int32_t r, int32_t& firstChunk, int32_t& chunks, int32_t& chunkSize, int32_t& ptr ) const
   {
      const int32_t sl = rowToSlice[ r ];
      const int4 s = reinterpret_cast< const int4* >( slices )[ sl ];
      firstChunk = r != s.z ? rowToChunk[ r - 1 ] : 0;
      chunks = rowToChunk[ r ] - firstChunk;
      chunkSize = s.y;
      ptr = s.w;
   }
   __device__ int32_t size( int32_t r ) const
   {
      int32_t f, c, cs, p;
      locate( r, f, c, cs, p );
      return c * cs;
   }
   __device__ int64_t global( int32_t r, int32_t l ) const
   {
      int32_t f, c, cs, p;
      locate( r, f, c, cs, p );
      return p + (int64_t) ( l % cs ) * kChunksInSlice + f + l / cs;
   }
};

struct BiView
{
   const int32_t* rowPerm;
   const int32_t* groupPointers;
   // number of groups segment r spans (detail/BiEllpack.h:59-78): ranks 0 -> 6, 1 -> 5,
   // 2..3 -> 4, 4..7 -> 3, 8..15 -> 2, 16..31 -> 1
   __device__ static int32_t activeGroups( int32_t rank ) { return kGroupsPerStrip - ( 32 - __clz( rank ) ); }
   __device__ int32_t size( int32_t r ) const
   {
      const int32_t strip = r / kStrip, rank = rowPerm[ r ] - strip * kStrip;
      const int32_t* gp = groupPointers + strip * kGroupsPerStrip;
      int32_t s = 0;
      for( int32_t g = 0; g < activeGroups( rank ); g++ )
         s += ( gp[ g + 1 ] - gp[ g ] ) >> ( kLogStrip - g );
      return s;
   }
   __device__ int64_t global( int32_t r, int32_t l ) const
   {
      const int32_t strip = r / kStrip, rank = rowPerm[ r ] - strip * kStrip;
      const int32_t* gp = groupPointers + strip * kGroupsPerStrip;
      const int32_t groups = activeGroups( rank );
      for( int32_t g = 0; g < groups; g++ ) {
         const int32_t gs = gp[ g + 1 ] - gp[ g ];
         const int32_t h = kStrip >> g;
         const int32_t w = gs >> ( kLogStrip - g );
         if( l < w )
            return gp[ g ] + rank + (int64_t) l * h;
         l -= w;
      }
      return -1;
   }
};

// f( view ) with the view of M's format
template< typename F >
void withView( const Matrix& M, F&& f )
{
   switch( kindOf( M.format ) ) {
      case KIND_CSR: f( CsrView{ M.offsets.ptr } ); break;
      case KIND_ELLPACK: f( EllpackView{ M.meta.segmentSize, M.meta.alignedSize } ); break;
      case KIND_SLICED: f( SlicedView{ M.sliceOffsets.ptr, M.sliceSizes.ptr, log2Exact( M.meta.sliceSize ) } ); break;
      case KIND_CHUNKED: f( ChunkedView{ M.rowToChunk.ptr, M.rowToSlice.ptr, M.slices.ptr } ); break;
      case KIND_BIELLPACK: f( BiView{ M.rowPerm.ptr, M.groupPointers.ptr } ); break;
   }
}

// f( (V*) nullptr ) with the value type of a dtype
template< typename F >
void withValueType( int dtype, F&& f )
{
   switch( dtype ) {
      case TNLHIP_F64: f( (double*) nullptr ); break;
      case TNLHIP_F32: f( (float*) nullptr ); break;
      case TNLHIP_BF16: f( (bf16*) nullptr ); break;
      case TNLHIP_BINARY: f( (uint8_t*) nullptr ); break;  // no values array: pointers are NULL
      default: throw Error( TNLHIP_ERROR_INVALID_ARGUMENT, "unknown value type" );
   }
}

}  // namespace tnlhip
