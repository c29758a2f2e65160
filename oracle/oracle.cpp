/*
 * CPU oracle — TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * A single-threaded restatement of TNL's Devices::Host semantics for the SpMV + Krylov
 * path.  With OMP_NUM_THREADS=1 TNL's Host reductions degenerate to the Sequential
 * order restated here, so residual histories are comparable value by value.
 *
 * Reference files (paths relative to the TNL source tree):
 *   CSR offsets              src/TNL/Algorithms/Segments/detail/CSR.h:22-34
 *   Ellpack                  src/TNL/Algorithms/Segments/Ellpack.hpp:72-101, EllpackView.hpp:246-252, 331-385
 *   SlicedEllpack            src/TNL/Algorithms/Segments/SlicedEllpack.hpp:64-102, SlicedEllpackView.hpp:108-132, 237-290
 *   ChunkedEllpack           src/TNL/Algorithms/Segments/ChunkedEllpack.hpp:84-235, detail/ChunkedEllpack.h:110-145,
 *                            ChunkedEllpackView.hpp:267-318
 *   BiEllpack                src/TNL/Algorithms/Segments/BiEllpack.hpp:74-279, detail/BiEllpack.h:59-185,
 *                            BiEllpackView.hpp:234-297
 *   CSR-Adaptive blocks      src/TNL/Algorithms/Segments/Kernels/CSRAdaptiveKernel.hpp:95-164,
 *                            detail/CSRAdaptiveKernelParameters.h:16-112
 *   conversion               src/TNL/Matrices/SparseMatrix.hpp:1070-1180
 *   vectorProduct keepers    src/TNL/Matrices/SparseMatrixView.hpp:390-491 (general and symmetric)
 *   reduce (sequential)      src/TNL/Algorithms/detail/Reduction.hpp:37-80
 *   multireduce (sequential) src/TNL/Algorithms/Multireduction.hpp:30-100
 *   gemv (host)              src/TNL/Matrices/MatrixOperations.h:46-150
 *   stopping rules           src/TNL/Solvers/IterativeSolver.hpp:103-150
 *   CG                       src/TNL/Solvers/Linear/CG.hpp:17-107
 *   BiCGStab                 src/TNL/Solvers/Linear/BICGStab.hpp:39-147
 *   TFQMR                    src/TNL/Solvers/Linear/TFQMR.hpp:17-125
 *   BiCGStab(ell)            src/TNL/Solvers/Linear/BICGStabL.hpp:37-284
 *   GMRES                    src/TNL/Solvers/Linear/GMRES.hpp:79-801
 *   Jacobi                   src/TNL/Solvers/Linear/Preconditioners/Diagonal.hpp:22-47
 *   ILU(0)                   src/TNL/Solvers/Linear/Preconditioners/ILU0.hpp:17-137,
 *                            src/TNL/Solvers/Linear/Utils/TriangularSolve.h:25-111
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

namespace {

using I = int32_t;

struct SliceInfo
{
   I size, chunkSize, firstSegment, pointer;
};

}  // namespace

struct orc_segments
{
   I kind = ORC_CSR;
   bool colmajor = true;
   I n = 0;
   // CSR
   std::vector< I > offsets;
   // Ellpack
   I segmentSize = 0, alignedSize = 0;
   // SlicedEllpack
   std::vector< I > sliceOffsets, sliceSegmentSizes;
   I slicedSize = 0, slicedAligned = 0;
   I sliceSize = 32;  // the SliceSize template parameter, SlicedEllpack.h:22 (default 32)
   // ChunkedEllpack
   I chunksInSlice = 256, desiredChunkSize = 16, numberOfSlices = 0, chunkedStorage = 0;
   std::vector< I > rowToChunk, rowToSlice, chunksToSegments, rowPointers;
   std::vector< SliceInfo > slices;
   // BiEllpack
   I virtualRows = 0;
   std::vector< I > rowPerm, groupPointers;
   int64_t biStorage = 0;
};

namespace {

constexpr I kAlignment = 32;     // Ellpack.h:21
constexpr I kWarpSize = 32;      // BiEllpack.h:22
constexpr I kLogWarpSize = 5;

I roundUpDivision( I a, I b ) { return a / b + ( a % b != 0 ); }

// ---------------------------------------------------------------- metadata builders

void buildCSR( orc_segments& s, const I* caps )
{
   // detail/CSR.h:22-34 — offsets = exclusive scan of sizes, offsets[n] = storage size
   s.offsets.assign( s.n + 1, 0 );
   for( I i = 0; i < s.n; i++ )
      s.offsets[ i + 1 ] = s.offsets[ i ] + caps[ i ];
}

void buildEllpack( orc_segments& s, const I* caps )
{
   // Ellpack.hpp:72-83 — segmentSize = max(sizes); alignedSize = n (row-major) or
   // roundUp(n, Alignment) (column-major)
   I mx = std::numeric_limits< I >::lowest();
   for( I i = 0; i < s.n; i++ )
      mx = std::max( mx, caps[ i ] );
   if( s.n == 0 )
      mx = std::numeric_limits< I >::lowest();
   s.segmentSize = s.n ? mx : 0;
   s.alignedSize = s.colmajor ? roundUpDivision( s.n, kAlignment ) * kAlignment : s.n;
}

void buildSliced( orc_segments& s, const I* caps )
{
   // SlicedEllpack.hpp:64-102 — per slice of SliceSize segments: max size; offsets = exclusive
   // scan of SliceSize*max; size = sum(sizes); alignedSize = offsets[last]
   const I kSliceSize = s.sliceSize;
   const I slices = roundUpDivision( s.n, kSliceSize );
   s.sliceOffsets.assign( slices + 1, 0 );
   s.sliceSegmentSizes.assign( slices, 0 );
   for( I sl = 0; sl < slices; sl++ ) {
      I res = std::numeric_limits< I >::min();
      for( I k = 0; k < kSliceSize; k++ ) {
         const I g = sl * kSliceSize + k;
         res = std::max( res, g < s.n ? caps[ g ] : 0 );
      }
      s.sliceOffsets[ sl ] = res * kSliceSize;
      s.sliceSegmentSizes[ sl ] = res;
   }
   // exclusive scan in place (the last element was 0)
   I run = 0;
   for( I sl = 0; sl <= slices; sl++ ) {
      const I v = s.sliceOffsets[ sl ];
      s.sliceOffsets[ sl ] = run;
      run += v;
   }
   I total = 0;
   for( I i = 0; i < s.n; i++ )
      total += caps[ i ];
   s.slicedSize = total;
   s.slicedAligned = s.sliceOffsets[ slices ];
}

void buildChunked( orc_segments& s, const I* caps )
{
   // ChunkedEllpack.hpp:194-235
   const I n = s.n;
   s.slices.assign( n, SliceInfo{ 0, 0, 0, 0 } );
   s.rowToChunk.assign( n, 0 );
   s.rowToSlice.assign( n, 0 );
   s.rowPointers.assign( n + 1, 0 );

   // resolveSliceSizes, ChunkedEllpack.hpp:84-113
   const I desiredElementsInSlice = s.chunksInSlice * s.desiredChunkSize;
   I segmentIdx = 0, sliceSize = 0, allocated = 0;
   s.numberOfSlices = 0;
   while( segmentIdx < n ) {
      allocated += caps[ segmentIdx ];
      sliceSize++;
      segmentIdx++;
      if( allocated < desiredElementsInSlice )
         if( segmentIdx < n && sliceSize < s.chunksInSlice )
            continue;
      s.slices[ s.numberOfSlices ].size = sliceSize;
      s.slices[ s.numberOfSlices ].firstSegment = segmentIdx - sliceSize;
      s.slices[ s.numberOfSlices ].pointer = allocated;  // temporary, as in TNL
      sliceSize = 0;
      s.numberOfSlices++;
      allocated = 0;
   }

   // setSlice for every slice, ChunkedEllpack.hpp:118-189
   s.rowPointers[ 0 ] = 0;
   I storage = 0;
   for( I sliceIndex = 0; sliceIndex < s.numberOfSlices; sliceIndex++ ) {
      SliceInfo& sl = s.slices[ sliceIndex ];
      const I sliceBegin = sl.firstSegment;
      const I allocatedInSlice = sl.pointer;
      const I sliceEnd = sliceBegin + sl.size;
      I freeChunks = s.chunksInSlice - sl.size;
      for( I i = sliceBegin; i < sliceEnd; i++ )
         s.rowToChunk[ i ] = 1;
      int totalAddedChunks = 0;
      int maxRowLength = caps[ sliceBegin ];
      for( I i = sliceBegin; i < sliceEnd; i++ ) {
         double rowRatio = 0.0;
         if( allocatedInSlice != 0 )
            rowRatio = (double) caps[ i ] / (double) allocatedInSlice;
         const I addedChunks = freeChunks * rowRatio;  // truncation of a double, as in TNL
         totalAddedChunks += addedChunks;
         s.rowToChunk[ i ] += addedChunks;
         if( maxRowLength < caps[ i ] )
            maxRowLength = caps[ i ];
      }
      freeChunks -= totalAddedChunks;
      while( freeChunks )
         for( I i = sliceBegin; i < sliceEnd && freeChunks; i++ )
            if( caps[ i ] == maxRowLength ) {
               s.rowToChunk[ i ]++;
               freeChunks--;
            }
      I maxChunkInSlice = 0;
      for( I i = sliceBegin; i < sliceEnd; i++ )
         maxChunkInSlice = std::max( maxChunkInSlice, roundUpDivision( caps[ i ], s.rowToChunk[ i ] ) );
      sl.chunkSize = maxChunkInSlice;
      sl.pointer = storage;
      storage += s.chunksInSlice * maxChunkInSlice;
      for( I i = sliceBegin; i < sliceEnd; i++ )
         s.rowToSlice[ i ] = sliceIndex;
      for( I i = sliceBegin; i < sliceEnd; i++ )
         s.rowPointers[ i + 1 ] = maxChunkInSlice * s.rowToChunk[ i ];
      for( I j = sliceBegin + 1; j < sliceEnd; j++ )
         s.rowToChunk[ j ] += s.rowToChunk[ j - 1 ];
   }
   s.chunkedStorage = storage;
   // inplaceInclusiveScan( rowPointers )
   for( I i = 1; i <= n; i++ )
      s.rowPointers[ i ] += s.rowPointers[ i - 1 ];
   // chunksToSegmentsMapping
   s.chunksToSegments.assign( (size_t) s.numberOfSlices * s.chunksInSlice, 0 );
   I chunkIdx = 0;
   for( I seg = 0; seg < n; seg++ ) {
      const I sliceIdx = s.rowToSlice[ seg ];
      I firstChunk = 0;
      if( seg != s.slices[ sliceIdx ].firstSegment )
         firstChunk = s.rowToChunk[ seg - 1 ];
      const I cnt = s.rowToChunk[ seg ] - firstChunk;
      for( I i = 0; i < cnt; i++ )
         s.chunksToSegments[ chunkIdx++ ] = seg;
   }
}

void buildBiEllpack( orc_segments& s, const I* caps )
{
   // BiEllpack.hpp:246-279
   const I n = s.n;
   if( n % kWarpSize != 0 )
      s.virtualRows = n + kWarpSize - ( n % kWarpSize );
   else
      s.virtualRows = n;
   const I strips = s.virtualRows / kWarpSize;
   s.rowPerm.assign( n, 0 );
   s.groupPointers.assign( (size_t) strips * ( kLogWarpSize + 1 ) + 1, 0 );

   // performRowBubbleSort, BiEllpack.hpp:74-129: a cocktail sort that swaps strictly
   // inverted neighbours only, so its result is the stable descending order of the sizes
   // within each strip; rowPerm[row] = strip base + rank in that order.
   for( I st = 0; st < strips; st++ ) {
      const I begin = st * kWarpSize;
      const I end = std::min( begin + kWarpSize, n );
      for( I i = begin; i < end; i++ ) {
         I rank = 0;
         for( I j = begin; j < end; j++ )
            if( caps[ j ] > caps[ i ] || ( caps[ j ] == caps[ i ] && j < i ) )
               rank++;
         s.rowPerm[ i ] = begin + rank;
      }
   }

   // computeColumnSizes, BiEllpack.hpp:134-171
   for( I strip = 0; strip < strips; strip++ ) {
      const I firstSegment = strip * kWarpSize;
      const I groupBegin = strip * ( kLogWarpSize + 1 );
      I emptyGroups = 0;
      if( strip == strips - 1 ) {
         const I segmentsCount = n - firstSegment;
         while( segmentsCount <= ( 1 << ( kLogWarpSize - 1 - emptyGroups ) ) - 1 )
            emptyGroups++;
         for( I g = groupBegin; g < groupBegin + emptyGroups; g++ )
            s.groupPointers[ g ] = 0;
      }
      I allocatedColumns = 0;
      for( I groupIdx = emptyGroups; groupIdx < kLogWarpSize; groupIdx++ ) {
         const I segmentIdx = ( 1 << ( kLogWarpSize - 1 - groupIdx ) ) - 1;
         I permSegm = 0;
         while( s.rowPerm[ permSegm + firstSegment ] != segmentIdx + firstSegment )
            permSegm++;
         const I groupWidth = caps[ permSegm + firstSegment ] - allocatedColumns;
         const I groupHeight = 1 << ( kLogWarpSize - groupIdx );
         allocatedColumns = caps[ permSegm + firstSegment ];
         s.groupPointers[ groupIdx + groupBegin ] = groupWidth * groupHeight;
      }
   }
   // inplaceExclusiveScan( groupPointers )
   I run = 0;
   for( auto& g : s.groupPointers ) {
      const I v = g;
      g = run;
      run += v;
   }
   s.biStorage = (int64_t) kWarpSize * s.groupPointers[ (size_t) strips * ( kLogWarpSize + 1 ) ];
}

// number of groups a segment spans, detail/BiEllpack.h:59-78
I biActiveGroups( const orc_segments& s, I seg )
{
   const I strip = seg / kWarpSize;
   const I rank = s.rowPerm[ seg ] - kWarpSize * strip;
   I bisection = 1;
   for( I i = 0; i < kLogWarpSize + 1; i++ ) {
      if( rank < bisection )
         return kLogWarpSize + 1 - i;
      bisection *= 2;
   }
   return -1;
}

I biGroupSize( const orc_segments& s, I strip, I group )
{
   const I o = strip * ( kLogWarpSize + 1 ) + group;
   return s.groupPointers[ o + 1 ] - s.groupPointers[ o ];
}

}  // namespace

// ---------------------------------------------------------------- segments API

extern "C" orc_segments*
orc_segments_create( int32_t kind, int32_t n, const int32_t* caps, int32_t column_major )
{
   auto* s = new orc_segments;
   s->kind = kind;
   s->n = n;
   s->colmajor = column_major != 0;
   switch( kind ) {
      case ORC_CSR: buildCSR( *s, caps ); break;
      case ORC_ELLPACK: buildEllpack( *s, caps ); break;
      case ORC_SLICED_ELLPACK: buildSliced( *s, caps ); break;
      case ORC_CHUNKED_ELLPACK: buildChunked( *s, caps ); break;
      case ORC_BIELLPACK: buildBiEllpack( *s, caps ); break;
      default: delete s; return nullptr;
   }
   return s;
}

// SlicedEllpack< ..., SliceSize > for another slice size than TNL's default 32
extern "C" orc_segments*
orc_segments_create_sliced( int32_t n, const int32_t* caps, int32_t column_major, int32_t slice_size )
{
   if( slice_size < 1 )
      return nullptr;
   auto* s = new orc_segments;
   s->kind = ORC_SLICED_ELLPACK;
   s->n = n;
   s->colmajor = column_major != 0;
   s->sliceSize = slice_size;
   buildSliced( *s, caps );
   return s;
}

extern "C" void
orc_segments_destroy( orc_segments* s )
{
   delete s;
}

extern "C" int64_t
orc_segments_storage_size( const orc_segments* s )
{
   switch( s->kind ) {
      case ORC_CSR: return s->offsets[ s->n ];
      case ORC_ELLPACK: return (int64_t) s->alignedSize * s->segmentSize;
      case ORC_SLICED_ELLPACK: return s->slicedAligned;
      case ORC_CHUNKED_ELLPACK: return s->chunkedStorage;
      case ORC_BIELLPACK: return s->biStorage;
   }
   return -1;
}

extern "C" int32_t
orc_segments_segment_size( const orc_segments* s, int32_t seg )
{
   switch( s->kind ) {
      case ORC_CSR: return s->offsets[ seg + 1 ] - s->offsets[ seg ];
      case ORC_ELLPACK: return s->segmentSize;
      case ORC_SLICED_ELLPACK: return s->sliceSegmentSizes[ seg / s->sliceSize ];
      case ORC_CHUNKED_ELLPACK: {
         // detail/ChunkedEllpack.h:74-89
         const I sl = s->rowToSlice[ seg ];
         const I first = seg != s->slices[ sl ].firstSegment ? s->rowToChunk[ seg - 1 ] : 0;
         return s->slices[ sl ].chunkSize * ( s->rowToChunk[ seg ] - first );
      }
      case ORC_BIELLPACK: {
         // detail/BiEllpack.h:112-129
         const I strip = seg / kWarpSize;
         const I groups = biActiveGroups( *s, seg );
         I h = kWarpSize, size = 0;
         for( I g = 0; g < groups; g++ ) {
            size += biGroupSize( *s, strip, g ) / h;
            h /= 2;
         }
         return size;
      }
   }
   return -1;
}

extern "C" int64_t
orc_segments_global_index( const orc_segments* s, int32_t seg, int32_t local )
{
   switch( s->kind ) {
      case ORC_CSR: return (int64_t) s->offsets[ seg ] + local;
      case ORC_ELLPACK:
         // EllpackView.hpp:246-252
         if( ! s->colmajor )
            return (int64_t) seg * s->segmentSize + local;
         return seg + (int64_t) s->alignedSize * local;
      case ORC_SLICED_ELLPACK: {
         // SlicedEllpackView.hpp:108-132
         const I sl = seg / s->sliceSize, in = seg % s->sliceSize;
         if( ! s->colmajor )
            return s->sliceOffsets[ sl ] + (int64_t) in * s->sliceSegmentSizes[ sl ] + local;
         return s->sliceOffsets[ sl ] + in + (int64_t) s->sliceSize * local;
      }
      case ORC_CHUNKED_ELLPACK: {
         // detail/ChunkedEllpack.h:110-145
         const I sl = s->rowToSlice[ seg ];
         const I first = seg != s->slices[ sl ].firstSegment ? s->rowToChunk[ seg - 1 ] : 0;
         const I cs = s->slices[ sl ].chunkSize;
         if( ! s->colmajor )
            return s->slices[ sl ].pointer + (int64_t) first * cs + local;
         return s->slices[ sl ].pointer + (int64_t) ( local % cs ) * s->chunksInSlice + first + local / cs;
      }
      case ORC_BIELLPACK: {
         // detail/BiEllpack.h:150-185
         const I strip = seg / kWarpSize;
         const I rank = s->rowPerm[ seg ] - strip * kWarpSize;
         const I groups = biActiveGroups( *s, seg );
         int64_t g0 = s->groupPointers[ strip * ( kLogWarpSize + 1 ) ];
         I h = kWarpSize;
         for( I g = 0; g < groups; g++ ) {
            const I gs = biGroupSize( *s, strip, g );
            if( gs ) {
               const I w = gs / h;
               if( local >= w ) {
                  local -= w;
                  g0 += gs;
               }
               else
                  return s->colmajor ? g0 + rank + (int64_t) local * h : g0 + (int64_t) rank * w + local;
            }
            h /= 2;
         }
         return -1;
      }
   }
   return -1;
}

extern "C" int64_t
orc_segments_get( const orc_segments* s, const char* name_c, int32_t* out )
{
   const std::string name( name_c );
   auto put = [ & ]( const std::vector< I >& v ) -> int64_t
   {
      if( out )
         std::copy( v.begin(), v.end(), out );
      return (int64_t) v.size();
   };
   if( name == "offsets" )
      return put( s->offsets );
   if( name == "ellpack" )
      return put( { s->segmentSize, s->n, s->alignedSize } );
   if( name == "sliceOffsets" )
      return put( s->sliceOffsets );
   if( name == "sliceSegmentSizes" )
      return put( s->sliceSegmentSizes );
   if( name == "sliced" )
      return put( { s->slicedSize, s->slicedAligned } );
   if( name == "rowToChunkMapping" )
      return put( s->rowToChunk );
   if( name == "rowToSliceMapping" )
      return put( s->rowToSlice );
   if( name == "chunksToSegmentsMapping" )
      return put( s->chunksToSegments );
   if( name == "rowPointers" )
      return put( s->rowPointers );
   if( name == "slices" ) {
      std::vector< I > v;
      for( I i = 0; i < s->numberOfSlices; i++ ) {
         const auto& sl = s->slices[ i ];
         v.insert( v.end(), { sl.size, sl.chunkSize, sl.firstSegment, sl.pointer } );
      }
      return put( v );
   }
   if( name == "chunked" )
      return put( { s->numberOfSlices, s->chunkedStorage } );
   if( name == "rowPermArray" )
      return put( s->rowPerm );
   if( name == "groupPointers" )
      return put( s->groupPointers );
   if( name == "biellpack" )
      return put( { s->virtualRows, (I) ( s->biStorage & 0xffffffff ), (I) ( s->biStorage >> 32 ) } );
   return -1;
}

// ---------------------------------------------------------------- conversion

extern "C" void
orc_convert_from_csr( const orc_segments* s, const int32_t* row_ptr, const int32_t* cols, const double* vals,
                      double* out_vals, int32_t* out_cols )
{
   // SparseMatrix.hpp:1070-1107 — capacities come from the source; every slot is set to
   // the padding index (-1, values 0), then the source entries are copied in order,
   // skipping explicit zeros and source padding.
   const int64_t storage = orc_segments_storage_size( s );
   for( int64_t g = 0; g < storage; g++ ) {
      out_vals[ g ] = 0.0;
      out_cols[ g ] = -1;
   }
   for( I r = 0; r < s->n; r++ ) {
      I local = 0;
      for( I k = row_ptr[ r ]; k < row_ptr[ r + 1 ]; k++ ) {
         if( vals[ k ] != 0.0 && cols[ k ] != -1 ) {
            const int64_t g = orc_segments_global_index( s, r, local++ );
            out_cols[ g ] = cols[ k ];
            out_vals[ g ] = vals[ k ];
         }
      }
   }
}

// ---------------------------------------------------------------- SpMV

namespace {

template< typename Real >
struct Keeper
{
   // SparseMatrixView.hpp:447-490 — keeper chosen from (alpha, beta)
   Real* y;
   Real alpha, beta;
   void operator()( I row, Real v ) const
   {
      if( beta == 0.0 ) {
         if( alpha == 1.0 )
            y[ row ] = v;
         else
            y[ row ] = alpha * v;
      }
      else {
         if( alpha == 1.0 )
            y[ row ] = beta * y[ row ] + v;
         else
            y[ row ] = beta * y[ row ] + alpha * v;
      }
   }
};

template< typename Real >
void spmv( const orc_segments* s, const Real* vals, const I* cols, const Real* x, Real* y, Real alpha, Real beta, I first,
           I last )
{
   Keeper< Real > keep{ y, alpha, beta };
   if( last == 0 )
      last = s->n;
   // fetch with padding check: padding (-1) stops the row and contributes 0.
   // CSR has no padding in TNL (it would read x[-1]); here padding is skipped in every
   // format, which is the same value whenever x is finite.
   auto fetch = [ & ]( int64_t g, bool& compute ) -> Real
   {
      const I c = cols[ g ];
      compute = c != -1;
      if( ! compute )
         return 0;
      return vals[ g ] * x[ c ];
   };
   switch( s->kind ) {
      case ORC_CSR:
         // CSRScalarKernel.hpp:70-107 — row-sequential left-to-right sum from 0
         for( I r = first; r < last; r++ ) {
            Real aux = 0;
            for( I g = s->offsets[ r ]; g < s->offsets[ r + 1 ]; g++ ) {
               bool c;
               aux = aux + fetch( g, c );
            }
            keep( r, aux );
         }
         break;
      case ORC_ELLPACK:
         // EllpackView.hpp:331-385
         for( I r = first; r < last; r++ ) {
            Real aux = 0;
            bool compute = true;
            if( ! s->colmajor ) {
               const int64_t b = (int64_t) r * s->segmentSize, e = b + s->segmentSize;
               for( int64_t j = b; j < e && compute; j++ )
                  aux = aux + fetch( j, compute );
            }
            else {
               const int64_t e = (int64_t) s->alignedSize * s->segmentSize;
               for( int64_t j = r; j < e && compute; j += s->alignedSize )
                  aux = aux + fetch( j, compute );
            }
            keep( r, aux );
         }
         break;
      case ORC_SLICED_ELLPACK:
         // SlicedEllpackView.hpp:237-290 (the loop runs over padding; padding adds 0)
         for( I r = first; r < last; r++ ) {
            const I sl = r / s->sliceSize, in = r % s->sliceSize;
            Real aux = 0;
            bool compute = true;
            if( ! s->colmajor ) {
               const I sz = s->sliceSegmentSizes[ sl ];
               const int64_t b = s->sliceOffsets[ sl ] + (int64_t) in * sz, e = b + sz;
               for( int64_t g = b; g < e; g++ )
                  aux = aux + fetch( g, compute );
            }
            else {
               const int64_t b = s->sliceOffsets[ sl ] + in, e = s->sliceOffsets[ sl + 1 ];
               for( int64_t g = b; g < e; g += s->sliceSize )
                  aux = aux + fetch( g, compute );
            }
            keep( r, aux );
         }
         break;
      case ORC_CHUNKED_ELLPACK:
         // ChunkedEllpackView.hpp:267-318
         for( I r = first; r < last; r++ ) {
            const I sl = s->rowToSlice[ r ];
            const I firstChunk = r != s->slices[ sl ].firstSegment ? s->rowToChunk[ r - 1 ] : 0;
            const I chunks = s->rowToChunk[ r ] - firstChunk;
            const I off = s->slices[ sl ].pointer, cs = s->slices[ sl ].chunkSize;
            Real aux = 0;
            bool compute = true;
            if( ! s->colmajor ) {
               const int64_t b = off + (int64_t) firstChunk * cs, e = b + (int64_t) chunks * cs;
               for( int64_t g = b; g < e && compute; g++ )
                  aux = aux + fetch( g, compute );
            }
            else {
               for( I c = 0; c < chunks; c++ ) {
                  const int64_t b = off + firstChunk + c, e = b + (int64_t) s->chunksInSlice * cs;
                  for( int64_t g = b; g < e && compute; g += s->chunksInSlice )
                     aux = aux + fetch( g, compute );
               }
            }
            keep( r, aux );
         }
         break;
      case ORC_BIELLPACK:
         // BiEllpackView.hpp:234-297 — TNL's Host path visits every segment regardless of
         // [first,last); here only [first,last) is written, which is what the API promises.
         if( orc_segments_storage_size( s ) == 0 )
            return;
         for( I r = first; r < last; r++ ) {
            const I strip = r / kWarpSize;
            const I rank = s->rowPerm[ r ] - strip * kWarpSize;
            const I groups = biActiveGroups( *s, r );
            int64_t g0 = s->groupPointers[ strip * ( kLogWarpSize + 1 ) ];
            I h = kWarpSize;
            Real aux = 0;
            bool compute = true;
            for( I grp = 0; grp < groups && compute; grp++ ) {
               const I gs = biGroupSize( *s, strip, grp );
               const I w = gs / h;
               const int64_t back = g0;
               int64_t g = g0 + ( s->colmajor ? rank : (int64_t) rank * w );
               for( I j = 0; j < w && compute; j++ ) {
                  aux = aux + fetch( g, compute );
                  g += s->colmajor ? h : 1;
               }
               g0 = back + gs;
               h /= 2;
            }
            keep( r, aux );
         }
         break;
   }
}

}  // namespace

extern "C" void
orc_spmv_f64( const orc_segments* s, const double* vals, const int32_t* cols, const double* x, double* y, double alpha,
              double beta, int32_t first, int32_t last )
{
   spmv< double >( s, vals, cols, x, y, alpha, beta, first, last );
}

extern "C" void
orc_spmv_f32( const orc_segments* s, const float* vals, const int32_t* cols, const float* x, float* y, float alpha,
              float beta, int32_t first, int32_t last )
{
   spmv< float >( s, vals, cols, x, y, alpha, beta, first, last );
}

// ---------------------------------------------------------------- row reductions

// SparseMatrixView::reduceRows (SparseMatrixView.hpp:495-530): every slot of a row in local
// order through fetch_ -- a padding slot returns the identity, others fetch( row, column, value )
// -- folded by aux = reduce( aux, f ) from the identity (the Host reduceSegments loops, e.g.
// CSRScalarKernel.hpp:70-107), then keep( row, aux ).  fetch: 0 value, 1 |value|, 2 one,
// 3 value != 0 (getCompressedRowLengths, SparseMatrixView.hpp:116-135), 4 column; reduce:
// 0 std::plus, 1 TNL::min (std::min), 2 TNL::max (std::max).  getRowCapacities
// (SparseMatrixView.hpp:143-160) is fetch 2 with std::plus.
extern "C" void
orc_reduce_rows( const orc_segments* s, const double* vals, const int32_t* cols, int32_t first, int32_t last,
                 int32_t fetch, int32_t reduce, double identity, double* out )
{
   if( last == 0 )
      last = s->n;
   for( I r = first; r < last; r++ ) {
      double aux = identity;
      const I size = orc_segments_segment_size( s, r );
      for( I l = 0; l < size; l++ ) {
         const int64_t g = orc_segments_global_index( s, r, l );
         double f = identity;
         if( cols[ g ] != -1 ) {
            const double v = vals[ g ];
            f = fetch == 0 ? v : fetch == 1 ? std::abs( v ) : fetch == 2 ? 1.0 : fetch == 3 ? ( v != 0.0 ? 1.0 : 0.0 )
                                                                                           : (double) cols[ g ];
         }
         if( reduce == 0 )
            aux = aux + f;
         else if( reduce == 1 )
            aux = std::min( aux, f );
         else
            aux = std::max( aux, f );
      }
      out[ r ] = aux;
   }
}

// ---------------------------------------------------------------- CSR-Adaptive descriptors

extern "C" int64_t
orc_csr_adaptive_blocks( int32_t n, const int32_t* offsets, int32_t value_size, int32_t warp_size, int32_t block_size,
                         int32_t streamed_shared_bytes, int32_t max_adaptive_elements, int32_t* out )
{
   // CSRAdaptiveKernel.hpp:31-48 — an all-empty matrix has no blocks
   I mx = 0;
   for( I i = 0; i <= n; i++ )
      mx = std::max( mx, offsets[ i ] );
   if( mx == 0 )
      return 0;
   // CSRAdaptiveKernelParameters.h: elements per warp that fit the streamed shared memory
   const size_t warpsCount = block_size / warp_size;
   const size_t streamedPerWarp = ( streamed_shared_bytes / value_size ) / warpsCount;
   const I rows = n + 1;  // offsets.getSize()
   std::vector< I > blocks;
   auto emit = [ & ]( I type, I first, I bsize, I segs, I warpIdx, I warps )
   {
      blocks.insert( blocks.end(), { type, first, bsize, segs, warpIdx, warps } );
   };
   enum { LONG = 0, STREAM = 1, VECTOR = 2 };
   I start = 0, nextStart = 0;
   while( nextStart != rows - 1 ) {
      // findLimit, CSRAdaptiveKernel.hpp:95-124
      I type = STREAM;
      size_t sum = 0;
      nextStart = rows - 1;
      for( I current = start; current < rows - 1; current++ ) {
         sum += offsets[ current + 1 ] - offsets[ current ];
         if( sum > streamedPerWarp ) {
            if( current - start > 0 ) {
               type = STREAM;
               nextStart = current;
            }
            else {
               type = sum <= 2 * (size_t) max_adaptive_elements ? VECTOR : LONG;
               nextStart = current + 1;
            }
            break;
         }
      }
      // initValueSize, CSRAdaptiveKernel.hpp:126-164
      if( type == LONG ) {
         const I blocksCount = (I) ( blocks.size() / 6 );
         const I warpsPerBlock = block_size / warp_size;
         I warpsLeft = roundUpDivision( blocksCount, warpsPerBlock ) * warpsPerBlock - blocksCount;
         if( warpsLeft == 0 )
            warpsLeft = warpsPerBlock;
         for( I idx = 0; idx < warpsLeft; idx++ )
            emit( LONG, start, 0, 0, idx, warpsLeft );
      }
      else
         emit( type, start, offsets[ nextStart ] - offsets[ start ], nextStart - start, 0, 0 );
      start = nextStart;
   }
   emit( VECTOR, nextStart, 0, 0, 0, 0 );  // sentinel
   if( out )
      std::copy( blocks.begin(), blocks.end(), out );
   return (int64_t) blocks.size() / 6;
}

// ---------------------------------------------------------------- reductions and BLAS-1

namespace {

// Reduction< Devices::Sequential >::reduce, Reduction.hpp:37-80
template< typename Real, typename Fetch >
Real seqReduce( int64_t n, Fetch&& fetch )
{
   constexpr int block = 128;
   const int64_t blocks = n / block;
   if( blocks > 1 ) {
      Real r[ 4 ] = { 0, 0, 0, 0 };
      for( int64_t b = 0; b < blocks; b++ ) {
         const int64_t off = b * block;
         for( int i = 0; i < block; i += 4 ) {
            r[ 0 ] = r[ 0 ] + fetch( off + i );
            r[ 1 ] = r[ 1 ] + fetch( off + i + 1 );
            r[ 2 ] = r[ 2 ] + fetch( off + i + 2 );
            r[ 3 ] = r[ 3 ] + fetch( off + i + 3 );
         }
      }
      for( int64_t i = blocks * block; i < n; i++ )
         r[ 0 ] = r[ 0 ] + fetch( i );
      r[ 0 ] = r[ 0 ] + r[ 2 ];
      r[ 1 ] = r[ 1 ] + r[ 3 ];
      r[ 0 ] = r[ 0 ] + r[ 1 ];
      return r[ 0 ];
   }
   Real res = 0;
   for( int64_t i = 0; i < n; i++ )
      res = res + fetch( i );
   return res;
}

template< typename Real >
Real dot( const std::vector< Real >& a, const std::vector< Real >& b )
{
   return seqReduce< Real >( (int64_t) a.size(), [ & ]( int64_t i ) { return a[ i ] * b[ i ]; } );
}

template< typename Real >
Real norm2( const std::vector< Real >& a )
{
   return std::sqrt( seqReduce< Real >( (int64_t) a.size(), [ & ]( int64_t i ) { return a[ i ] * a[ i ]; } ) );
}

// Multireduction< Devices::Sequential >::reduce, Multireduction.hpp:30-100
template< typename Real, typename Fetch >
void seqMultiReduce( I size, int n, Fetch&& fetch, Real* result )
{
   constexpr int block = 128;
   const int blocks = size / block;
   if( blocks > 1 ) {
      std::vector< Real > r( 4 * n, 0 );
      for( int b = 0; b < blocks; b++ ) {
         const I off = b * block;
         for( int k = 0; k < n; k++ ) {
            Real* q = r.data() + 4 * k;
            for( int i = 0; i < block; i += 4 ) {
               q[ 0 ] = q[ 0 ] + fetch( off + i, k );
               q[ 1 ] = q[ 1 ] + fetch( off + i + 1, k );
               q[ 2 ] = q[ 2 ] + fetch( off + i + 2, k );
               q[ 3 ] = q[ 3 ] + fetch( off + i + 3, k );
            }
         }
      }
      for( int k = 0; k < n; k++ ) {
         Real* q = r.data() + 4 * k;
         for( I i = blocks * block; i < size; i++ )
            q[ 0 ] = q[ 0 ] + fetch( i, k );
      }
      for( int k = 0; k < n; k++ ) {
         Real* q = r.data() + 4 * k;
         q[ 0 ] = q[ 0 ] + q[ 1 ];
         q[ 0 ] = q[ 0 ] + q[ 2 ];
         q[ 0 ] = q[ 0 ] + q[ 3 ];
         result[ k ] = q[ 0 ];
      }
   }
   else {
      for( int k = 0; k < n; k++ )
         result[ k ] = 0;
      for( int b = 0; b < blocks; b++ ) {
         const I off = b * block;
         for( int k = 0; k < n; k++ )
            for( int i = 0; i < block; i++ )
               result[ k ] = result[ k ] + fetch( off + i, k );
      }
      for( int k = 0; k < n; k++ )
         for( I i = blocks * block; i < size; i++ )
            result[ k ] = result[ k ] + fetch( i, k );
   }
}

// MatrixOperations< Devices::Host >::gemv, MatrixOperations.h:46-150 (sequential order)
template< typename Real >
void gemv( I m, I n, Real alpha, const Real* A, I lda, const Real* x, Real beta, Real* y )
{
   std::vector< Real > ax( n );
   for( I k = 0; k < n; k++ )
      ax[ k ] = alpha * x[ k ];
   if( n == 1 ) {
      if( beta != 0.0 )
         for( I j = 0; j < m; j++ )
            y[ j ] = A[ j ] * ax[ 0 ] + beta * y[ j ];
      else
         for( I j = 0; j < m; j++ )
            y[ j ] = A[ j ] * ax[ 0 ];
      return;
   }
   constexpr I block = 128;
   const I blocks = m / block;
   Real aux[ block ];
   for( I b = 0; b < blocks; b++ ) {
      const I bo = b * block;
      for( I j = 0; j < block; j++ )
         aux[ j ] = 0.0;
      for( I k = 0; k < n; k++ ) {
         const int64_t o = bo + (int64_t) k * lda;
         for( I j = 0; j < block; j++ )
            aux[ j ] += A[ o + j ] * ax[ k ];
      }
      if( beta != 0.0 )
         for( I j = 0; j < block; j++ )
            y[ bo + j ] = aux[ j ] + beta * y[ bo + j ];
      else
         for( I j = 0; j < block; j++ )
            y[ bo + j ] = aux[ j ];
   }
   for( I j = blocks * block; j < m; j++ ) {
      Real tmp = 0.0;
      for( I k = 0; k < n; k++ )
         tmp += A[ j + (int64_t) k * lda ] * ax[ k ];
      y[ j ] = beta != 0.0 ? tmp + beta * y[ j ] : tmp;
   }
}

// ---------------------------------------------------------------- solvers

struct Csr
{
   I n;
   const I* rp;
   const I* c;
   const double* v;
   void spmv( const std::vector< double >& x, std::vector< double >& y ) const
   {
      for( I r = 0; r < n; r++ ) {
         double aux = 0;
         for( I g = rp[ r ]; g < rp[ r + 1 ]; g++ )
            if( c[ g ] != -1 )
               aux = aux + v[ g ] * x[ c[ g ] ];
         y[ r ] = aux;
      }
   }
   double element( I i, I j ) const
   {
      for( I g = rp[ i ]; g < rp[ i + 1 ]; g++ )
         if( c[ g ] == j )
            return v[ g ];
      return 0.0;
   }
};

// IterativeSolver< double, int > stopping rules, IterativeSolver.hpp:103-150
struct Iter
{
   I maxIt, minIt;
   double conv, div;
   I it = 0;
   double res = 0;
   std::vector< double >* hist;
   void setResidue( double r )
   {
      res = r;
      hist->push_back( r );
   }
   bool checkNext() const
   {
      return ! ( std::isnan( res ) || it > maxIt || ( res > div && it >= minIt ) || ( res < conv && it >= minIt ) );
   }
   bool next()
   {
      const bool r = checkNext();
      it++;
      return r;
   }
   bool checkConvergence() const
   {
      if( std::isnan( res ) )
         return false;
      if( res > div && it > minIt )
         return false;
      if( it >= maxIt )
         return false;
      if( res > conv )
         return false;
      return true;
   }
};

struct Jacobi
{
   std::vector< double > d;
   void update( const Csr& A )
   {
      d.resize( A.n );
      for( I i = 0; i < A.n; i++ )
         d[ i ] = A.element( i, i );
   }
   void solve( const std::vector< double >& b, std::vector< double >& x ) const
   {
      for( size_t i = 0; i < b.size(); i++ )
         x[ i ] = b[ i ] / d[ i ];
   }
};

// ILU(0), Preconditioners/ILU0.hpp:17-117 (Host): L (unit diagonal implied) and U kept as two
// CSR matrices with the sparsity of A's lower and upper parts.  Row i of A is read in storage
// order; padding (column < 0) is skipped when counting, entries with column < i go to L, the
// rest up to column N to U.  The factorization runs row by row: for each L entry k (ascending),
// L_ik /= U_kk, then every later L entry and every U entry j of the row loses L_ik * U_kj,
// where U_kj is 0 when row k of U has no column j.  TNL stores the rows of U in reverse order;
// the row order of storage changes no value and is not kept here.
// solve: triangularSolveLower< true > then triangularSolveUpper< true, true >
// (Utils/TriangularSolve.h:25-111): x_i = b_i - sum L_ij x_j in row order, then
// x_i = ( x_i - sum_{c >= 1} U_ij x_j ) / U_i[ 0 ], the first stored entry of U's row.
struct Ilu0
{
   I n = 0;
   std::vector< I > Lp, Lc, Up, Uc;
   std::vector< double > Lv, Uv;
   double uElement( I k, I j ) const
   {
      for( I c = Up[ k ]; c < Up[ k + 1 ]; c++ )
         if( Uc[ c ] == j )
            return Uv[ c ];
      return 0.0;
   }
   void update( const Csr& A )
   {
      n = A.n;
      Lp.assign( (size_t) n + 1, 0 );
      Up.assign( (size_t) n + 1, 0 );
      Lc.clear(), Uc.clear(), Lv.clear(), Uv.clear();
      for( I i = 0; i < n; i++ ) {
         // the counting pass of ILU0.hpp:36-50, then the copy of the first entries
         I nl = 0, nu = 0;
         for( I g = A.rp[ i ]; g < A.rp[ i + 1 ]; g++ ) {
            const I col = A.c[ g ];
            if( col < 0 )
               continue;
            if( col < i )
               nl++;
            else if( col < n )
               nu++;
            else
               break;
         }
         I g = A.rp[ i ];
         while( g < A.rp[ i + 1 ] && A.c[ g ] < 0 )
            g++;
         for( I c = 0; c < nl; c++, g++ ) {
            Lc.push_back( A.c[ g ] );
            Lv.push_back( A.v[ g ] );
         }
         for( I c = 0; c < nu; c++, g++ ) {
            Uc.push_back( A.c[ g ] );
            Uv.push_back( A.v[ g ] );
         }
         Lp[ i + 1 ] = (I) Lc.size();
         Up[ i + 1 ] = (I) Uc.size();
      }
      for( I i = 0; i < n; i++ )
         for( I ck = Lp[ i ]; ck < Lp[ i + 1 ]; ck++ ) {
            const I k = Lc[ ck ];
            const double lik = Lv[ ck ] / uElement( k, k );
            Lv[ ck ] = lik;
            for( I cj = ck + 1; cj < Lp[ i + 1 ]; cj++ )
               Lv[ cj ] = Lv[ cj ] - lik * uElement( k, Lc[ cj ] );
            for( I cj = Up[ i ]; cj < Up[ i + 1 ]; cj++ )
               Uv[ cj ] = Uv[ cj ] - lik * uElement( k, Uc[ cj ] );
         }
   }
   void solve( const std::vector< double >& b, std::vector< double >& x ) const
   {
      for( I i = 0; i < n; i++ ) {
         double xi = b[ i ];
         for( I c = Lp[ i ]; c < Lp[ i + 1 ]; c++ )
            xi -= Lv[ c ] * x[ Lc[ c ] ];
         x[ i ] = xi;
      }
      for( I k = 0; k < n; k++ ) {
         const I i = n - 1 - k;
         double xi = x[ i ];
         for( I c = Up[ i ] + 1; c < Up[ i + 1 ]; c++ )
            xi -= Uv[ c ] * x[ Uc[ c ] ];
         x[ i ] = xi / Uv[ Up[ i ] ];
      }
   }
};

// The solvers' preconditioner: none, Jacobi (Diagonal) or ILU(0).
struct Precond
{
   int kind = 0;
   Jacobi jacobi;
   Ilu0 ilu;
   bool on = false;
   void update( const Csr& A, int k )
   {
      kind = k;
      on = k != 0;
      if( k == 1 )
         jacobi.update( A );
      else if( k == 2 )
         ilu.update( A );
   }
   void solve( const std::vector< double >& b, std::vector< double >& x ) const
   {
      if( kind == 1 )
         jacobi.solve( b, x );
      else
         ilu.solve( b, x );
   }
};

using Vec = std::vector< double >;

bool solveCG( const Csr& A, const Precond& M, const Vec& b, Vec& x, Iter& it )
{
   // CG.hpp:17-107
   const I n = A.n;
   Vec r( n ), p( n ), Ap( n ), z( n );
   it.it = 0;
   double normb;
   if( M.on ) {
      M.solve( b, r );
      normb = norm2( r );
   }
   else
      normb = norm2( b );
   if( normb == 0.0 )
      normb = 1.0;
   A.spmv( x, r );
   for( I i = 0; i < n; i++ )
      r[ i ] = b[ i ] - r[ i ];
   double s1, s2;
   if( M.on ) {
      M.solve( r, z );
      p = z;
      s1 = dot( r, z );
   }
   else {
      p = r;
      s1 = dot( r, r );
   }
   it.setResidue( std::sqrt( s1 ) / normb );
   while( it.next() ) {
      A.spmv( p, Ap );
      s2 = dot( Ap, p );
      if( s2 == 0.0 ) {
         it.setResidue( 0.0 );
         break;
      }
      const double alpha = s1 / s2;
      for( I i = 0; i < n; i++ )
         x[ i ] += alpha * p[ i ];
      for( I i = 0; i < n; i++ )
         r[ i ] -= alpha * Ap[ i ];
      if( M.on ) {
         M.solve( r, z );
         s2 = s1;
         s1 = dot( r, z );
      }
      else {
         s2 = s1;
         s1 = dot( r, r );
      }
      const double beta = s2 == 0.0 ? 0.0 : s1 / s2;
      if( M.on )
         for( I i = 0; i < n; i++ )
            p[ i ] = z[ i ] + beta * p[ i ];
      else
         for( I i = 0; i < n; i++ )
            p[ i ] = r[ i ] + beta * p[ i ];
      it.setResidue( std::sqrt( s1 ) / normb );
   }
   return it.checkConvergence();
}

bool solveBiCGStab( const Csr& A, const Precond& M, const Vec& b, Vec& x, Iter& it, bool exact )
{
   // BICGStab.hpp:39-147
   const I n = A.n;
   Vec r( n ), r_ast( n ), p( n ), s( n ), Ap( n ), As( n ), Mt( n );
   auto compute_residue = [ & ]( Vec& out )
   {
      if( M.on ) {
         A.spmv( x, Mt );
         for( I i = 0; i < n; i++ )
            Mt[ i ] = b[ i ] - Mt[ i ];
         M.solve( Mt, out );
      }
      else {
         A.spmv( x, out );
         for( I i = 0; i < n; i++ )
            out[ i ] = b[ i ] - out[ i ];
      }
   };
   auto matvec = [ & ]( const Vec& src, Vec& dst )
   {
      if( M.on ) {
         A.spmv( src, Mt );
         M.solve( Mt, dst );
      }
      else
         A.spmv( src, dst );
   };
   double b_norm;
   if( M.on ) {
      M.solve( b, Mt );
      b_norm = norm2( Mt );
   }
   else
      b_norm = norm2( b );
   if( b_norm == 0.0 )
      b_norm = 1.0;
   compute_residue( r );
   p = r;
   r_ast = r;
   std::fill( s.begin(), s.end(), 0.0 );
   double rho = dot( r, r_ast );
   double r_ast_sqnorm = rho;
   const double eps2 = std::numeric_limits< double >::epsilon() * std::numeric_limits< double >::epsilon();
   it.it = 0;
   it.setResidue( std::sqrt( rho ) / b_norm );
   while( it.next() ) {
      matvec( p, Ap );
      const double alpha = rho / dot( Ap, r_ast );
      for( I i = 0; i < n; i++ )
         s[ i ] = r[ i ] - alpha * Ap[ i ];
      matvec( s, As );
      const double omega = dot( As, s ) / dot( As, As );
      for( I i = 0; i < n; i++ )
         x[ i ] += alpha * p[ i ] + omega * s[ i ];
      for( I i = 0; i < n; i++ )
         r[ i ] = s[ i ] - omega * As[ i ];
      const double rho_old = rho;
      rho = dot( r, r_ast );
      if( std::fabs( rho ) < eps2 * r_ast_sqnorm ) {
         compute_residue( r );
         r_ast = r;
         rho = dot( r, r_ast );
         r_ast_sqnorm = rho;
      }
      const double beta = ( rho / rho_old ) * ( alpha / omega );
      for( I i = 0; i < n; i++ )
         p[ i ] = r[ i ] + beta * p[ i ] - ( beta * omega ) * Ap[ i ];
      if( exact ) {
         compute_residue( s );
         it.setResidue( norm2( s ) / b_norm );
      }
      else
         it.setResidue( norm2( r ) / b_norm );
   }
   return it.checkConvergence();
}

// BiCGStab(ell), BICGStabL.hpp:37-253: ell Bi-CG steps on the columns of R and U, then the
// minimal-residual part (MGS of R's columns, the small triangular systems for g_0 / g_2) and three
// gemv updates.  Inside the Bi-CG part nextIteration() is called only to count (its check is
// ignored); the stopping rule is checked once per outer iteration.  Returns -1 when the initial
// residue is NaN (TNL throws).
int solveBiCGStabL( const Csr& A, const Precond& M, const Vec& b, Vec& x, Iter& it, int ell, bool exact )
{
   const I n = A.n;
   const size_t ld = n;
   Vec R( ( ell + 1 ) * ld ), U( ( ell + 1 ) * ld ), r_ast( n ), Mt( n ), res( n );
   Vec sigma( ell + 1 ), g0( ell + 1 ), g1( ell + 1 ), g2( ell + 1 ), T( ell * ell );
   auto col = [ & ]( Vec& V, int j ) { return V.data() + j * ld; };
   auto matvec = [ & ]( const double* src, double* dst )
   {
      Vec in( src, src + n ), out( n );
      if( M.on ) {
         A.spmv( in, Mt );
         M.solve( Mt, out );
      }
      else
         A.spmv( in, out );
      std::copy( out.begin(), out.end(), dst );
   };
   auto compute_residue = [ & ]( double* r )
   {
      Vec out( n );
      if( M.on ) {
         A.spmv( x, Mt );
         for( I i = 0; i < n; i++ )
            Mt[ i ] = b[ i ] - Mt[ i ];
         M.solve( Mt, out );
      }
      else {
         A.spmv( x, out );
         for( I i = 0; i < n; i++ )
            out[ i ] = b[ i ] - out[ i ];
      }
      std::copy( out.begin(), out.end(), r );
   };
   auto dotp = [ & ]( const double* a, const double* c ) { return dot( Vec( a, a + n ), Vec( c, c + n ) ); };
   double b_norm;
   if( M.on ) {
      M.solve( b, Mt );
      b_norm = norm2( Mt );
   }
   else
      b_norm = norm2( b );
   if( b_norm == 0.0 )
      b_norm = 1.0;
   double* r0 = col( R, 0 );
   double* u0 = col( U, 0 );
   compute_residue( r0 );
   sigma[ 0 ] = std::sqrt( dotp( r0, r0 ) );
   if( std::isnan( sigma[ 0 ] ) )
      return -1;
   for( I i = 0; i < n; i++ )
      r_ast[ i ] = r0[ i ] / sigma[ 0 ];
   double rho_0 = 1.0, alpha = 0.0, omega = 1.0;
   std::fill( u0, u0 + n, 0.0 );
   it.it = 0;
   it.setResidue( sigma[ 0 ] / b_norm );
   while( it.checkNext() ) {
      rho_0 = -omega * rho_0;
      for( int j = 0; j < ell; j++ ) {
         it.it++;
         const double rho_1 = dotp( r_ast.data(), col( R, j ) );
         const double beta = alpha * rho_1 / rho_0;
         rho_0 = rho_1;
         // U_[0:j] := R_[0:j] - beta U_[0:j]  (geam: 1.0 * R + (-beta) * U)
         for( int c = 0; c <= j; c++ )
            for( I i = 0; i < n; i++ )
               col( U, c )[ i ] = 1.0 * col( R, c )[ i ] + -beta * col( U, c )[ i ];
         matvec( col( U, j ), col( U, j + 1 ) );
         const double gamma = dotp( r_ast.data(), col( U, j + 1 ) );
         alpha = rho_0 / gamma;
         // R_[0:j] := R_[0:j] - alpha U_[1:j+1]
         for( int c = 0; c <= j; c++ )
            for( I i = 0; i < n; i++ )
               col( R, c )[ i ] = 1.0 * col( R, c )[ i ] + -alpha * col( U, c + 1 )[ i ];
         matvec( col( R, j ), col( R, j + 1 ) );
         for( I i = 0; i < n; i++ )
            x[ i ] = x[ i ] + alpha * u0[ i ];
      }
      for( int j = 1; j <= ell; j++ ) {
         double* rj = col( R, j );
         for( int i = 1; i < j; i++ ) {
            const double* ri = col( R, i );
            const int ij = ( i - 1 ) + ( j - 1 ) * ell;
            T[ ij ] = dotp( ri, rj ) / sigma[ i ];
            for( I k = 0; k < n; k++ )
               rj[ k ] = rj[ k ] - T[ ij ] * ri[ k ];
         }
         sigma[ j ] = dotp( rj, rj );
         g1[ j ] = dotp( r0, rj ) / sigma[ j ];
      }
      omega = g1[ ell ];
      for( int j = ell; j >= 1; j-- ) {
         g0[ j ] = g1[ j ];
         for( int i = j + 1; i <= ell; i++ )
            g0[ j ] -= T[ ( j - 1 ) + ( i - 1 ) * ell ] * g0[ i ];
      }
      for( int j = 1; j < ell; j++ ) {
         g2[ j ] = g0[ j + 1 ];
         for( int i = j + 1; i < ell; i++ )
            g2[ j ] += T[ ( j - 1 ) + ( i - 1 ) * ell ] * g0[ i + 1 ];
      }
      g2[ 0 ] = g0[ 1 ];
      gemv< double >( n, ell, 1.0, R.data(), (I) ld, g2.data(), 1.0, x.data() );
      gemv< double >( n, ell, -1.0, R.data() + ld, (I) ld, &g1[ 1 ], 1.0, r0 );
      gemv< double >( n, ell, -1.0, U.data() + ld, (I) ld, &g0[ 1 ], 1.0, u0 );
      if( exact ) {
         compute_residue( res.data() );
         sigma[ 0 ] = norm2( res );
      }
      else
         sigma[ 0 ] = std::sqrt( dotp( r0, r0 ) );
      it.setResidue( sigma[ 0 ] / b_norm );
   }
   return it.checkConvergence() ? 1 : 0;
}

bool solveTFQMR( const Csr& A, const Precond& M, const Vec& b, Vec& x, Iter& it )
{
   // TFQMR.hpp:17-125: odd iterations take alpha = rho / (v, r_ast), even ones recompute
   // A u; the residue is the quasi-residual bound tau sqrt(iter + 1) / |b|
   const I n = A.n;
   Vec r( n ), w( n ), u( n ), v( n ), d( n ), r_ast( n ), Au( n ), Mt( n );
   auto matvec = [ & ]( const Vec& src, Vec& dst )
   {
      if( M.on ) {
         A.spmv( src, Mt );
         M.solve( Mt, dst );
      }
      else
         A.spmv( src, dst );
   };
   double tau, theta, eta, rho, alpha, b_norm, w_norm;
   if( M.on ) {
      M.solve( b, Mt );
      b_norm = norm2( Mt );
      A.spmv( x, Mt );
      for( I i = 0; i < n; i++ )
         Mt[ i ] = b[ i ] - Mt[ i ];
      M.solve( Mt, r );
   }
   else {
      b_norm = norm2( b );
      A.spmv( x, r );
      for( I i = 0; i < n; i++ )
         r[ i ] = b[ i ] - r[ i ];
   }
   w = r;
   u = r;
   matvec( u, Au );
   v = Au;
   std::fill( d.begin(), d.end(), 0.0 );
   tau = norm2( r );
   theta = eta = 0.0;
   r_ast = r;
   rho = dot( r_ast, r );
   alpha = 0.0;
   if( b_norm == 0.0 )
      b_norm = 1.0;
   it.it = 0;
   it.setResidue( tau / b_norm );
   while( it.next() ) {
      const I iter = it.it;
      if( iter % 2 == 1 )
         alpha = rho / dot( v, r_ast );
      else
         matvec( u, Au );
      for( I i = 0; i < n; i++ )
         w[ i ] -= alpha * Au[ i ];
      const double dc = theta * theta * eta / alpha;
      for( I i = 0; i < n; i++ )
         d[ i ] = u[ i ] + dc * d[ i ];
      w_norm = norm2( w );
      theta = w_norm / tau;
      const double c = 1.0 / std::sqrt( 1.0 + theta * theta );
      tau = tau * theta * c;
      eta = c * c * alpha;
      for( I i = 0; i < n; i++ )
         x[ i ] += eta * d[ i ];
      it.setResidue( tau * std::sqrt( iter + 1 ) / b_norm );
      if( iter > it.minIt && it.res < it.conv )
         break;
      if( iter % 2 == 0 ) {
         const double rho_new = dot( w, r_ast );
         const double beta = rho_new / rho;
         rho = rho_new;
         for( I i = 0; i < n; i++ )
            u[ i ] = w[ i ] + beta * u[ i ];
         for( I i = 0; i < n; i++ )
            v[ i ] = beta * Au[ i ] + ( beta * beta ) * v[ i ];
         matvec( u, Au );
         for( I i = 0; i < n; i++ )
            v[ i ] += Au[ i ];
      }
      else
         for( I i = 0; i < n; i++ )
            u[ i ] -= alpha * v[ i ];
   }
   return it.checkConvergence();
}

struct Gmres
{
   // GMRES.hpp:79-801, Devices::Host instantiation (localOffset = 0, no MPI)
   enum { CGS = 0, CGSR = 1, MGS = 2, MGSR = 3, CWY = 4 };
   const Csr& A;
   const Precond& M;
   Iter& it;
   int variant, rmin, rmax, smin, smax;
   I size = 0, ldSize = 0;
   Vec r, w, Mt, z, V, Y, cs, sn, H, s, T, YL;

   Gmres( const Csr& A_, const Precond& M_, Iter& it_, const orc_solver_params& p )
   : A( A_ ), M( M_ ), it( it_ ), variant( p.gmres_variant ), rmin( p.restarting_min ), rmax( p.restarting_max ),
     smin( p.restarting_step_min ), smax( p.restarting_step_max )
   {}

   void setSize()
   {
      size = A.n;
      const I mult = 256 / sizeof( double );
      ldSize = ( ( size + mult - 1 ) / mult ) * mult + 1;  // host: +1 against false sharing
      const int m = rmax;
      r.assign( size, 0 );
      w.assign( size, 0 );
      Mt.assign( size, 0 );
      cs.assign( m + 1, 0 );
      sn.assign( m + 1, 0 );
      H.assign( ( m + 1 ) * m, 0 );
      s.assign( m + 1, 0 );
      if( variant == CWY ) {
         z.assign( size, 0 );
         Y.assign( (size_t) ldSize * ( m + 1 ), 0 );
         T.assign( ( m + 1 ) * ( m + 1 ), 0 );
         YL.assign( ( m + 1 ) * ( m + 1 ), 0 );
         V.assign( size, 0 );
      }
      else
         V.assign( (size_t) ldSize * ( m + 1 ), 0 );
   }

   void compute_residue( Vec& out, const Vec& x, const Vec& b )
   {
      if( M.on ) {
         A.spmv( x, Mt );
         for( I i = 0; i < size; i++ )
            Mt[ i ] = b[ i ] - Mt[ i ];
         M.solve( Mt, out );
      }
      else {
         A.spmv( x, out );
         for( I i = 0; i < size; i++ )
            out[ i ] = b[ i ] - out[ i ];
      }
   }

   // w = M^{-1} A v (v given as a pointer into V / a vector)
   void matvec( Vec& out, const double* v )
   {
      Vec vv( v, v + size );
      if( M.on ) {
         A.spmv( vv, Mt );
         M.solve( Mt, out );
      }
      else
         A.spmv( vv, out );
   }

   static void generatePlaneRotation( double& dx, double& dy, double& c, double& sn_ )
   {
      if( dy == 0.0 ) {
         c = 1.0;
         sn_ = 0.0;
      }
      else if( std::fabs( dy ) > std::fabs( dx ) ) {
         const double t = dx / dy;
         sn_ = 1.0 / std::sqrt( 1.0 + t * t );
         c = t * sn_;
      }
      else {
         const double t = dy / dx;
         c = 1.0 / std::sqrt( 1.0 + t * t );
         sn_ = t * c;
      }
   }
   static void applyPlaneRotation( double& dx, double& dy, double& c, double& sn_ )
   {
      const double t = c * dx + sn_ * dy;
      dy = c * dy - sn_ * dx;
      dx = t;
   }
   void apply_givens_rotations( int i, int m )
   {
      for( int k = 0; k < i; k++ )
         applyPlaneRotation( H[ k + i * ( m + 1 ) ], H[ k + 1 + i * ( m + 1 ) ], cs[ k ], sn[ k ] );
      if( H[ i + 1 + i * ( m + 1 ) ] != 0.0 ) {
         generatePlaneRotation( H[ i + i * ( m + 1 ) ], H[ i + 1 + i * ( m + 1 ) ], cs[ i ], sn[ i ] );
         applyPlaneRotation( H[ i + i * ( m + 1 ) ], H[ i + 1 + i * ( m + 1 ) ], cs[ i ], sn[ i ] );
         applyPlaneRotation( s[ i ], s[ i + 1 ], cs[ i ], sn[ i ] );
      }
   }

   int orthogonalize_CGS( int m, double normb, double beta )
   {
      double* V0 = V.data();
      for( I j = 0; j < size; j++ )
         V0[ j ] = ( 1.0 / beta ) * r[ j ];
      std::fill( H.begin(), H.end(), 0.0 );
      std::fill( s.begin(), s.end(), 0.0 );
      s[ 0 ] = beta;
      for( int i = 0; i < m && it.next(); i++ ) {
         double* vi = V.data() + (size_t) i * ldSize;
         matvec( w, vi );
         for( int k = 0; k <= i; k++ )
            H[ k + i * ( m + 1 ) ] = 0.0;
         const int reorth = variant == CGSR ? 2 : 1;
         for( int l = 0; l < reorth; l++ ) {
            std::vector< double > Hl( i + 1 );
            const double* _V = V.data();
            const double* _w = w.data();
            const I ld = ldSize;
            seqMultiReduce< double >(
               size, i + 1, [ & ]( I idx, int k ) { return _V[ idx + (size_t) k * ld ] * _w[ idx ]; }, Hl.data() );
            for( int k = 0; k <= i; k++ )
               H[ k + i * ( m + 1 ) ] += Hl[ k ];
            gemv< double >( size, i + 1, -1.0, V.data(), ldSize, Hl.data(), 1.0, w.data() );
         }
         const double normw = norm2( w );
         H[ i + 1 + i * ( m + 1 ) ] = normw;
         double* vn = V.data() + (size_t) ( i + 1 ) * ldSize;
         for( I j = 0; j < size; j++ )
            vn[ j ] = ( 1.0 / normw ) * w[ j ];
         apply_givens_rotations( i, m );
         it.setResidue( std::fabs( s[ i + 1 ] ) / normb );
         if( ! it.checkNext() )
            return i;
      }
      return m;
   }

   int orthogonalize_MGS( int m, double normb, double beta )
   {
      double* V0 = V.data();
      for( I j = 0; j < size; j++ )
         V0[ j ] = ( 1.0 / beta ) * r[ j ];
      std::fill( H.begin(), H.end(), 0.0 );
      std::fill( s.begin(), s.end(), 0.0 );
      s[ 0 ] = beta;
      for( int i = 0; i < m && it.next(); i++ ) {
         double* vi = V.data() + (size_t) i * ldSize;
         matvec( w, vi );
         for( int k = 0; k <= i; k++ )
            H[ k + i * ( m + 1 ) ] = 0.0;
         const int reorth = variant == MGSR ? 2 : 1;
         for( int l = 0; l < reorth; l++ )
            for( int k = 0; k <= i; k++ ) {
               const double* vk = V.data() + (size_t) k * ldSize;
               const double Hki =
                  seqReduce< double >( size, [ & ]( int64_t j ) { return w[ j ] * vk[ j ]; } );
               H[ k + i * ( m + 1 ) ] += Hki;
               for( I j = 0; j < size; j++ )
                  w[ j ] = w[ j ] - Hki * vk[ j ];
            }
         const double normw = norm2( w );
         H[ i + 1 + i * ( m + 1 ) ] = normw;
         double* vn = V.data() + (size_t) ( i + 1 ) * ldSize;
         for( I j = 0; j < size; j++ )
            vn[ j ] = ( 1.0 / normw ) * w[ j ];
         apply_givens_rotations( i, m );
         it.setResidue( std::fabs( s[ i + 1 ] ) / normb );
         if( ! it.checkNext() )
            return i;
      }
      return m;
   }

   void hauseholder_generate( int i, double* y_i, const Vec& zz )
   {
      for( I j = 0; j < size; j++ )
         y_i[ j ] = j < i ? 0.0 : zz[ j ];
      const double normz =
         std::sqrt( seqReduce< double >( size, [ & ]( int64_t j ) { return y_i[ j ] * y_i[ j ]; } ) );
      const double y_ii = y_i[ i ];
      if( y_ii > 0.0 )
         y_i[ i ] = y_ii + normz;
      else
         y_i[ i ] = y_ii - normz;
      const double norm_yi_squared = 2 * ( normz * normz + std::fabs( y_ii ) * normz );
      const double t_i = 2.0 / norm_yi_squared;
      T[ i + i * ( rmax + 1 ) ] = t_i;
      if( i > 0 ) {
         std::vector< double > aux( i );
         const double* _Y = Y.data();
         const I ld = ldSize;
         seqMultiReduce< double >(
            size, i, [ & ]( I idx, int k ) { return _Y[ idx + (size_t) k * ld ] * y_i[ idx ]; }, aux.data() );
         for( int k = 0; k < i; k++ ) {
            T[ k + i * ( rmax + 1 ) ] = 0.0;
            for( int j = k; j < i; j++ )
               T[ k + i * ( rmax + 1 ) ] -= T[ k + j * ( rmax + 1 ) ] * ( t_i * aux[ j ] );
         }
      }
   }

   void hauseholder_apply_trunc( double* out, int i, const double* y_i, const Vec& zz )
   {
      double* YL_i = &YL[ i * ( rmax + 1 ) ];
      for( int k = 0; k < rmax + 1; k++ )
         YL_i[ k ] = y_i[ k ];
      const double aux = 1.0;
      for( int k = 0; k <= i; k++ )
         out[ k ] = zz[ k ] - y_i[ k ] * aux;
   }

   void hauseholder_cwy( Vec& v, int i )
   {
      std::vector< double > aux( i + 1 );
      for( int k = 0; k <= i; k++ )
         aux[ k ] = YL[ i + k * ( rmax + 1 ) ];
      for( int k = 0; k <= i; k++ ) {
         double a2 = 0.0;
         for( int j = k; j <= i; j++ )
            a2 += T[ k + j * ( rmax + 1 ) ] * aux[ j ];
         aux[ k ] = a2;
      }
      gemv< double >( size, i + 1, -1.0, Y.data(), ldSize, aux.data(), 0.0, v.data() );
      v[ i ] = 1.0 + v[ i ];
   }

   void hauseholder_cwy_transposed( Vec& zz, int i, const Vec& ww )
   {
      std::vector< double > aux( i + 1 );
      const double* _Y = Y.data();
      const I ld = ldSize;
      seqMultiReduce< double >(
         size, i + 1, [ & ]( I idx, int k ) { return _Y[ idx + (size_t) k * ld ] * ww[ idx ]; }, aux.data() );
      for( int k = i; k >= 0; k-- ) {
         double a2 = 0.0;
         for( int j = 0; j <= k; j++ )
            a2 += T[ j + k * ( rmax + 1 ) ] * aux[ j ];
         aux[ k ] = a2;
      }
      zz = ww;
      gemv< double >( size, i + 1, -1.0, Y.data(), ldSize, aux.data(), 1.0, zz.data() );
   }

   int orthogonalize_CWY( int m, double normb, double /*beta*/ )
   {
      z = r;
      std::fill( H.begin(), H.end(), 0.0 );
      std::fill( s.begin(), s.end(), 0.0 );
      std::fill( T.begin(), T.end(), 0.0 );
      for( int i = 0; i <= m && it.next(); i++ ) {
         double* y_i = Y.data() + (size_t) i * ldSize;
         hauseholder_generate( i, y_i, z );
         if( i == 0 )
            hauseholder_apply_trunc( s.data(), i, y_i, z );
         else
            hauseholder_apply_trunc( &H[ ( i - 1 ) * ( m + 1 ) ], i, y_i, z );
         hauseholder_cwy( V, i );
         if( i < m ) {
            matvec( w, V.data() );
            hauseholder_cwy_transposed( z, i, w );
         }
         if( i > 0 )
            apply_givens_rotations( i - 1, m );
         it.setResidue( std::fabs( s[ i ] ) / normb );
         if( i > 0 && ! it.checkNext() )
            return i - 1;
      }
      return m;
   }

   // returns false on a zero Hessenberg diagonal (TNL throws)
   bool update( int k, int m, Vec& x )
   {
      std::vector< double > y( m + 1 );
      for( int i = 0; i <= m; i++ )
         y[ i ] = s[ i ];
      for( int i = k; i >= 0; i-- ) {
         if( H[ i + i * ( m + 1 ) ] == 0 )
            return false;
         y[ i ] /= H[ i + i * ( m + 1 ) ];
         for( int j = i - 1; j >= 0; j-- )
            y[ j ] -= H[ j + i * ( m + 1 ) ] * y[ i ];
      }
      if( variant != CWY )
         gemv< double >( size, k + 1, 1.0, V.data(), ldSize, y.data(), 1.0, x.data() );
      else {
         std::vector< double > aux( k + 1, 0.0 );
         for( int j = 0; j <= k; j++ )
            for( int i = 0; i <= j; i++ )
               aux[ i ] += y[ j ] * YL[ j + i * ( rmax + 1 ) ];
         for( int i = 0; i <= k; i++ ) {
            double a2 = 0.0;
            for( int j = i; j <= k; j++ )
               a2 += T[ i + j * ( rmax + 1 ) ] * aux[ j ];
            aux[ i ] = a2;
         }
         gemv< double >( size, k + 1, -1.0, Y.data(), ldSize, aux.data(), 1.0, x.data() );
         for( int j = 0; j <= k; j++ )
            x[ j ] = x[ j ] + y[ j ];
      }
      return true;
   }

   int solve( const Vec& b, Vec& x )
   {
      if( rmin <= 0 || rmax <= 0 || rmin > rmax )
         return 0;
      if( smin < 0 || smax < 0 || smin > smax )
         return 0;
      setSize();
      double normb;
      if( M.on ) {
         M.solve( b, Mt );
         normb = norm2( Mt );
      }
      else
         normb = norm2( b );
      if( normb == 0.0 )
         normb = 1.0;
      compute_residue( r, x, b );
      double beta = norm2( r );
      it.it = 0;
      it.setResidue( beta / normb );
      double beta_ratio = 1;
      const double max_beta_ratio = 0.99, min_beta_ratio = 0.175;
      int restart_cycles = 0;
      int m = rmax;
      while( it.checkNext() ) {
         if( rmax > rmin && restart_cycles > 0 ) {
            if( beta_ratio > max_beta_ratio )
               m = rmax;
            else if( beta_ratio >= min_beta_ratio ) {
               const int step = smin + (float) ( smax - smin ) / ( rmax - rmin ) * ( m - rmin );
               if( m - step >= rmin )
                  m -= step;
               else
                  m = rmax;
            }
         }
         int o_steps = 0;
         switch( variant ) {
            case CGS:
            case CGSR: o_steps = orthogonalize_CGS( m, normb, beta ); break;
            case MGS:
            case MGSR: o_steps = orthogonalize_MGS( m, normb, beta ); break;
            case CWY: o_steps = orthogonalize_CWY( m, normb, beta ); break;
         }
         if( o_steps < m ) {
            if( ! update( o_steps, m, x ) )
               return -1;
            return it.checkConvergence() ? 1 : 0;
         }
         if( ! update( m - 1, m, x ) )
            return -1;
         compute_residue( r, x, b );
         const double beta_old = beta;
         beta = norm2( r );
         it.setResidue( beta / normb );
         ++restart_cycles;
         beta_ratio = beta / beta_old;
      }
      return it.checkConvergence() ? 1 : 0;
   }
};

}  // namespace

extern "C" double
orc_dot_f64( int64_t n, const double* a, const double* b )
{
   return seqReduce< double >( n, [ & ]( int64_t i ) { return a[ i ] * b[ i ]; } );
}

extern "C" float
orc_dot_f32( int64_t n, const float* a, const float* b )
{
   return seqReduce< float >( n, [ & ]( int64_t i ) { return a[ i ] * b[ i ]; } );
}

extern "C" void
orc_solver_params_default( orc_solver_params* p )
{
   // IterativeSolver.h:226-245, GMRES.h:220-237
   p->max_iterations = 1000000000;
   p->min_iterations = 0;
   p->convergence_residue = 1e-6;
   p->divergence_residue = std::numeric_limits< double >::max();
   p->gmres_variant = 4;
   p->restarting_min = 10;
   p->restarting_max = 10;
   p->restarting_step_min = 3;
   p->restarting_step_max = 3;
   p->bicgstab_exact_residue = 0;
   p->preconditioner = 0;
   p->bicgstab_ell = 1;  // BICGStabL.hpp:21
}

// SparseMatrixView::vectorProduct for MatrixType::isSymmetric() (SparseMatrixView.hpp:390-491):
// outVector *= beta when beta != 1, then symmetricFetch adds the transposed entry of every
// stored a_rc with c < r to y_c (AtomicOperations, sequential on one Host thread) and
// keeperGeneral adds alpha * (row sum) to y_r.
extern "C" void
orc_symmetric_spmv( int32_t n, const int32_t* rp, const int32_t* cols, const double* vals, const double* x, double* y,
                    double alpha, double beta, int32_t first, int32_t last )
{
   if( last == 0 )
      last = n;
   if( beta != 1.0 )
      for( I i = 0; i < n; i++ )
         y[ i ] *= beta;
   for( I r = first; r < last; r++ ) {
      double s = 0;
      for( I g = rp[ r ]; g < rp[ r + 1 ]; g++ ) {
         const I c = cols[ g ];
         if( c == -1 )
            continue;
         if( c < r )
            y[ c ] += alpha * vals[ g ] * x[ r ];
         s = s + vals[ g ] * x[ c ];
      }
      y[ r ] += alpha * s;
   }
}

// isBinary() fetch (SparseMatrixView.hpp:413-437): x_c itself, summed in ComputeReal = int; the
// general keepers (SparseMatrixView.hpp:439-491) with int multiplicators.  Sums wrap as uint32.
extern "C" void
orc_binary_spmv( int32_t n, const int32_t* rp, const int32_t* cols, const int32_t* x, int32_t* y, int32_t alpha,
                 int32_t beta, int32_t first, int32_t last )
{
   if( last == 0 )
      last = n;
   for( I r = first; r < last; r++ ) {
      uint32_t s = 0;
      for( I g = rp[ r ]; g < rp[ r + 1 ]; g++ )
         if( cols[ g ] != -1 )
            s += (uint32_t) x[ cols[ g ] ];
      const uint32_t v = alpha == 1 ? s : (uint32_t) alpha * s;
      y[ r ] = (int32_t) ( beta == 0 ? v : (uint32_t) beta * (uint32_t) y[ r ] + v );
   }
}

extern "C" int32_t
orc_ilu0( int32_t n, const int32_t* row_ptr, const int32_t* cols, const double* vals, int32_t* nnzL, int32_t* nnzU,
          int32_t* Lrp, int32_t* Lc, double* Lv, int32_t* Urp, int32_t* Uc, double* Uv, const double* b, double* x )
{
   Csr A{ n, row_ptr, cols, vals };
   Ilu0 f;
   f.update( A );
   *nnzL = (int32_t) f.Lc.size();
   *nnzU = (int32_t) f.Uc.size();
   if( Lrp ) {
      std::copy( f.Lp.begin(), f.Lp.end(), Lrp );
      std::copy( f.Lc.begin(), f.Lc.end(), Lc );
      std::copy( f.Lv.begin(), f.Lv.end(), Lv );
   }
   if( Urp ) {
      std::copy( f.Up.begin(), f.Up.end(), Urp );
      std::copy( f.Uc.begin(), f.Uc.end(), Uc );
      std::copy( f.Uv.begin(), f.Uv.end(), Uv );
   }
   if( b && x ) {
      Vec xx( n );
      f.solve( Vec( b, b + n ), xx );
      std::copy( xx.begin(), xx.end(), x );
   }
   return 0;
}

extern "C" int32_t
orc_solve( int32_t method, int32_t n, const int32_t* row_ptr, const int32_t* cols, const double* vals, const double* b,
           double* x, const orc_solver_params* p, int32_t* iterations, double* residue, double* history,
           int32_t history_cap, int32_t* history_len )
{
   Csr A{ n, row_ptr, cols, vals };
   Precond M;
   if( p->preconditioner == 1 || p->preconditioner == 2 )
      M.update( A, p->preconditioner );
   if( p->preconditioner == 2 )
      for( I i = 0; i < n; i++ )
         if( M.ilu.Up[ i ] == M.ilu.Up[ i + 1 ] )
            return -1;  // a row without entries in U: no U_ii to divide by
   std::vector< double > hist;
   Iter it{ p->max_iterations, p->min_iterations, p->convergence_residue, p->divergence_residue, 0, 0.0, &hist };
   Vec bb( b, b + n ), xx( x, x + n );
   int32_t rc = 0;
   if( method == 0 )
      rc = solveCG( A, M, bb, xx, it ) ? 1 : 0;
   else if( method == 1 )
      rc = solveBiCGStab( A, M, bb, xx, it, p->bicgstab_exact_residue != 0 ) ? 1 : 0;
   else if( method == 2 ) {
      Gmres g( A, M, it, *p );
      rc = g.solve( bb, xx );
   }
   else if( method == 3 )
      rc = solveTFQMR( A, M, bb, xx, it ) ? 1 : 0;
   else if( method == 4 )
      rc = solveBiCGStabL( A, M, bb, xx, it, p->bicgstab_ell, p->bicgstab_exact_residue != 0 );
   std::copy( xx.begin(), xx.end(), x );
   *iterations = it.it;
   *residue = it.res;
   const int32_t len = (int32_t) std::min< size_t >( hist.size(), (size_t) history_cap );
   if( history )
      std::copy( hist.begin(), hist.begin() + len, history );
   if( history_len )
      *history_len = (int32_t) hist.size();
   return rc;
}
