// Host construction of segment metadata from row capacities.  TNL builds ChunkedEllpack
// and BiEllpack metadata on the host even for GPU matrices and copies it over; the same
// is done here for every format, so the arrays are identical by construction.
//   CSR            Algorithms/Segments/detail/CSR.h:22-34
//   Ellpack        Algorithms/Segments/Ellpack.hpp:72-83 (column-major, alignment 32)
//   SlicedEllpack  Algorithms/Segments/SlicedEllpack.hpp:64-102
//   ChunkedEllpack Algorithms/Segments/ChunkedEllpack.hpp:84-235
//   BiEllpack      Algorithms/Segments/BiEllpack.hpp:74-279
//   CSR-Adaptive   Algorithms/Segments/Kernels/CSRAdaptiveKernel.hpp:95-164
#include "matrix.hpp"

#include <algorithm>
#include <limits>

namespace tnlhip {

namespace {

inline int32_t ceilDiv( int32_t a, int32_t b ) { return a / b + ( a % b != 0 ); }

void csrOffsets( HostSegments& s, const int32_t* caps )
{
   s.offsets.resize( (size_t) s.n + 1 );
   int64_t run = 0;
   for( int32_t i = 0; i < s.n; i++ ) {
      s.offsets[ i ] = (int32_t) run;
      run += caps[ i ];
   }
   if( run > std::numeric_limits< int32_t >::max() )
      throw Error( TNLHIP_ERROR_CAPACITY, "CSR storage exceeds the int32 index range" );
   s.offsets[ s.n ] = (int32_t) run;
   s.storage = s.allocated = run;
}

void ellpack( HostSegments& s, const int32_t* caps )
{
   int32_t mx = 0;
   for( int32_t i = 0; i < s.n; i++ )
      mx = std::max( mx, caps[ i ] );
   s.segmentSize = mx;
   s.alignedSize = ceilDiv( s.n, kEllpackAlignment ) * kEllpackAlignment;
   s.storage = s.allocated = (int64_t) s.alignedSize * s.segmentSize;
}

void sliced( HostSegments& s, const int32_t* caps )
{
   const int32_t S = s.sliceSize;
   const int32_t nslices = ceilDiv( s.n, S );
   s.sliceOffsets.assign( (size_t) nslices + 1, 0 );
   s.sliceSegmentSizes.assign( nslices, 0 );
   int64_t off = 0, total = 0;
   for( int32_t sl = 0; sl < nslices; sl++ ) {
      const int32_t b = sl * S, e = std::min( b + S, s.n );
      int32_t mx = 0;
      for( int32_t i = b; i < e; i++ ) {
         mx = std::max( mx, caps[ i ] );
         total += caps[ i ];
      }
      s.sliceSegmentSizes[ sl ] = mx;
      s.sliceOffsets[ sl ] = (int32_t) off;
      off += (int64_t) mx * S;
   }
   if( off > std::numeric_limits< int32_t >::max() )
      throw Error( TNLHIP_ERROR_CAPACITY, "SlicedEllpack storage exceeds the int32 index range" );
   s.sliceOffsets[ nslices ] = (int32_t) off;
   s.slicedSize = (int32_t) total;
   s.storage = s.allocated = off;
}

void chunked( HostSegments& s, const int32_t* caps )
{
   const int32_t n = s.n;
   const int32_t desired = kChunksInSlice * kDesiredChunkSize;
   // greedy slice closing (resolveSliceSizes): a slice ends once it holds at least
   // `desired` elements or 256 segments
   std::vector< int32_t > sliceFirst, sliceSize, sliceAlloc;
   {
      int32_t first = 0, alloc = 0;
      for( int32_t i = 0; i < n; i++ ) {
         alloc += caps[ i ];
         const int32_t size = i + 1 - first;
         if( alloc < desired && i + 1 < n && size < kChunksInSlice )
            continue;
         sliceFirst.push_back( first );
         sliceSize.push_back( size );
         sliceAlloc.push_back( alloc );
         first = i + 1;
         alloc = 0;
      }
   }
   const int32_t ns = (int32_t) sliceFirst.size();
   s.numberOfSlices = ns;
   s.rowToChunk.assign( n, 0 );
   s.rowToSlice.assign( n, 0 );
   s.rowPointers.assign( (size_t) n + 1, 0 );
   s.slices.assign( (size_t) 4 * ns, 0 );
   int64_t storage = 0;
   std::vector< int32_t > chunks;
   for( int32_t sl = 0; sl < ns; sl++ ) {
      const int32_t b = sliceFirst[ sl ], e = b + sliceSize[ sl ], alloc = sliceAlloc[ sl ];
      // chunk distribution (setSlice): one chunk per segment, then the free chunks in
      // proportion to the segment sizes (truncated double ratio), the remainder handed out
      // round-robin to the longest segments
      chunks.assign( sliceSize[ sl ], 1 );
      int32_t freeChunks = kChunksInSlice - sliceSize[ sl ];
      int32_t added = 0, longest = caps[ b ];
      for( int32_t i = b; i < e; i++ ) {
         const double ratio = alloc != 0 ? (double) caps[ i ] / (double) alloc : 0.0;
         const int32_t a = (int32_t) ( freeChunks * ratio );
         chunks[ i - b ] += a;
         added += a;
         longest = std::max( longest, caps[ i ] );
      }
      freeChunks -= added;
      while( freeChunks > 0 )
         for( int32_t i = b; i < e && freeChunks > 0; i++ )
            if( caps[ i ] == longest ) {
               chunks[ i - b ]++;
               freeChunks--;
            }
      int32_t chunkSize = 0;
      for( int32_t i = b; i < e; i++ )
         chunkSize = std::max( chunkSize, ceilDiv( caps[ i ], chunks[ i - b ] ) );
      s.slices[ 4 * sl + 0 ] = sliceSize[ sl ];
      s.slices[ 4 * sl + 1 ] = chunkSize;
      s.slices[ 4 * sl + 2 ] = b;
      s.slices[ 4 * sl + 3 ] = (int32_t) storage;
      storage += (int64_t) kChunksInSlice * chunkSize;
      int32_t inclusive = 0;
      for( int32_t i = b; i < e; i++ ) {
         s.rowToSlice[ i ] = sl;
         s.rowPointers[ i + 1 ] = chunkSize * chunks[ i - b ];
         inclusive += chunks[ i - b ];
         s.rowToChunk[ i ] = inclusive;
      }
   }
   if( storage > std::numeric_limits< int32_t >::max() )
      throw Error( TNLHIP_ERROR_CAPACITY, "ChunkedEllpack storage exceeds the int32 index range" );
   for( int32_t i = 1; i <= n; i++ )
      s.rowPointers[ i ] += s.rowPointers[ i - 1 ];
   s.chunksToSegments.assign( (size_t) ns * kChunksInSlice, 0 );
   int64_t c = 0;
   for( int32_t i = 0; i < n; i++ ) {
      const int32_t sl = s.rowToSlice[ i ];
      const int32_t firstChunk = i != s.slices[ 4 * sl + 2 ] ? s.rowToChunk[ i - 1 ] : 0;
      for( int32_t k = firstChunk; k < s.rowToChunk[ i ]; k++ )
         s.chunksToSegments[ c++ ] = i;
   }
   s.storage = s.allocated = storage;
}

void biellpack( HostSegments& s, const int32_t* caps )
{
   const int32_t n = s.n;
   s.virtualRows = ceilDiv( n, kStrip ) * kStrip;
   const int32_t strips = s.virtualRows / kStrip;
   s.rowPerm.assign( n, 0 );
   std::vector< int32_t > groups( (size_t) strips * kGroupsPerStrip + 1, 0 );
   int32_t byRank[ kStrip ];
   for( int32_t st = 0; st < strips; st++ ) {
      const int32_t b = st * kStrip, e = std::min( b + kStrip, n ), cnt = e - b;
      // position of each segment in the strip after a stable sort by decreasing size
      // (the fixed point of performRowBubbleSort, which only swaps strictly inverted pairs)
      for( int32_t i = b; i < e; i++ ) {
         int32_t rank = 0;
         for( int32_t j = b; j < e; j++ )
            rank += caps[ j ] > caps[ i ] || ( caps[ j ] == caps[ i ] && j < i );
         s.rowPerm[ i ] = b + rank;
         byRank[ rank ] = caps[ i ];
      }
      // group widths (computeColumnSizes): group g has height 32 >> g and width
      // size(rank (16 >> g) - 1) minus the widths before it; a short last strip skips
      // the groups taller than it needs
      int32_t empty = 0;
      if( st == strips - 1 )
         while( cnt <= ( 1 << ( kLogStrip - 1 - empty ) ) - 1 )
            empty++;
      int32_t allocatedColumns = 0;
      for( int32_t g = empty; g < kLogStrip; g++ ) {
         const int32_t rank = ( 1 << ( kLogStrip - 1 - g ) ) - 1;
         const int32_t width = byRank[ rank ] - allocatedColumns;
         allocatedColumns = byRank[ rank ];
         groups[ (size_t) st * kGroupsPerStrip + g ] = width * ( 1 << ( kLogStrip - g ) );
      }
   }
   int64_t run = 0;
   for( auto& g : groups ) {
      const int64_t v = g;
      g = (int32_t) run;
      run += v;
   }
   if( run * kStrip > std::numeric_limits< int32_t >::max() )
      throw Error( TNLHIP_ERROR_CAPACITY, "BiEllpack storage exceeds the int32 index range" );
   s.groupPointers = std::move( groups );
   s.allocated = run;
   s.storage = run * kStrip;  // TNL reports (and allocates) 32 x the used storage, BiEllpack.hpp:266
}

}  // namespace

void buildSegments( HostSegments& s, int kind, int32_t n, const int32_t* caps )
{
   const int32_t sliceSize = s.sliceSize;  // SlicedEllpack's slice size, set by the caller (32 by default)
   s = HostSegments();
   s.sliceSize = sliceSize;
   s.kind = kind;
   s.n = n;
   switch( kind ) {
      case KIND_CSR: csrOffsets( s, caps ); break;
      case KIND_ELLPACK: ellpack( s, caps ); break;
      case KIND_SLICED: sliced( s, caps ); break;
      case KIND_CHUNKED: chunked( s, caps ); break;
      case KIND_BIELLPACK: biellpack( s, caps ); break;
      default: throw Error( TNLHIP_ERROR_INVALID_ARGUMENT, "unknown segments kind" );
   }
}

std::vector< int32_t > buildAdaptiveBlocks( int32_t n, const int32_t* off, int value_size, int wave_size,
                                            int block_size, int streamed_bytes, int max_adaptive_elements )
{
   std::vector< int32_t > out;
   int32_t mx = 0;
   for( int32_t i = 0; i <= n; i++ )
      mx = std::max( mx, off[ i ] );
   if( mx == 0 )
      return out;
   const int waves = block_size / wave_size;
   const int64_t streamed = ( streamed_bytes / value_size ) / waves;  // elements per wave
   int32_t start = 0;
   while( start != n ) {
      // extend the block row by row until its elements exceed the streamed budget
      int type = BLOCK_STREAM;
      int32_t next = n;
      int64_t sum = 0;
      for( int32_t r = start; r < n; r++ ) {
         sum += off[ r + 1 ] - off[ r ];
         if( sum > streamed ) {
            if( r > start ) {
               next = r;
            }
            else {
               type = sum <= 2 * (int64_t) max_adaptive_elements ? BLOCK_VECTOR : BLOCK_LONG;
               next = r + 1;
            }
            break;
         }
      }
      if( type == BLOCK_LONG ) {
         // a long segment takes the waves left in the current workgroup (a whole
         // workgroup if none are left) so its partial sums meet in one LDS
         const int32_t count = (int32_t) ( out.size() / 6 );
         int32_t left = ceilDiv( count, waves ) * waves - count;
         if( left == 0 )
            left = waves;
         for( int32_t w = 0; w < left; w++ )
            out.insert( out.end(), { BLOCK_LONG, start, 0, 0, w, left } );
      }
      else
         out.insert( out.end(), { type, start, off[ next ] - off[ start ], next - start, 0, 0 } );
      start = next;
   }
   out.insert( out.end(), { BLOCK_VECTOR, start, 0, 0, 0, 0 } );  // closing block
   return out;
}

}  // namespace tnlhip
