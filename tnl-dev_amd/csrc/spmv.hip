// Sparse matrix-vector product y = alpha * A x + beta * y over every segment format,
// written for CDNA4 wavefront-64 execution.  SpMV is HBM-bound (about 0.14 flop per byte
// in FP64), so every kernel is organised around coalesced streaming of values and column
// indexes; x is gathered through L2 / MALL.
//
// Reference semantics: SparseMatrixView::vectorProduct (Matrices/SparseMatrixView.hpp:390-491):
// fetch = values[g] * x[columns[g]], padding (column -1) contributes nothing, keeper chosen
// from (alpha, beta).  Each format's Host traversal sums a row left to right in local
// order; the one-lane-per-row kernels here (CSR scalar, Ellpack, SlicedEllpack) keep that
// order and round each product and sum separately, so they reproduce the Host result bit
// for bit; the cooperative kernels reassociate the row sum.
#include "spmv.hpp"
#include "rowdist.hpp"

namespace tnlhip {

template< bool EXACT, typename A >
__device__ __forceinline__ A madd( A acc, A a, A b )
{
   if constexpr( EXACT )
      return mulAddExact( acc, a, b );
   else
      return acc + a * b;
}

// Streaming loads of the matrix arrays bypass the caches' normal retention (non-temporal),
// so the 256 MiB MALL and the L2 keep x, which is gathered many times.
__device__ __forceinline__ int32_t ldnt( const int32_t* p ) { return __builtin_nontemporal_load( p ); }
__device__ __forceinline__ double ldnt( const double* p ) { return __builtin_nontemporal_load( p ); }
__device__ __forceinline__ float ldnt( const float* p ) { return __builtin_nontemporal_load( p ); }
__device__ __forceinline__ bf16 ldnt( const bf16* p )
{
   return __ushort_as_bfloat16( __builtin_nontemporal_load( reinterpret_cast< const unsigned short* >( p ) ) );
}

__device__ __forceinline__ int64_t logicalBlock( int remap )
{
   return logicalBlockOf( blockIdx.x, gridDim.x, remap );
}

template< int P >
__device__ __forceinline__ RowSeq rowSequence( int64_t base, int64_t last, int64_t per, int dist )
{
   return rowSequenceAt< P >( base, last, per, dist, LaunchIndex{ blockIdx.x, gridDim.x, blockDim.x, threadIdx.x } );
}

// ------------------------------------------------------------------ CSR: G lanes per row

// G = 1: CSRScalar (one lane per row, exact Host order); G = 2..32: CSRLight / CSRHybrid;
// G = 64: CSRVector (a whole wavefront per row).  Groups of G lanes stride over the row
// with contiguous loads, then reduce across the group with cross-lane shuffles.  Every group
// keeps R rows in flight (rows r, r + stride, ...): their offsets, first chunk of columns /
// values and the gathers of x are issued together, so short rows do not serialize on the
// offsets -> columns -> x latency chain.  Each lane still sums its elements of a row in
// increasing order, so the result does not depend on R.
template< int G, int R, bool EXACT, typename V, typename A, typename Y >
__global__ __launch_bounds__( 256 ) void csrGroupedKernel( const int32_t* __restrict__ off,
                                                           const int32_t* __restrict__ col,
                                                           const V* __restrict__ val, const Y* __restrict__ x,
                                                           Keeper< A, Y > keep, int32_t first, int32_t last )
{
   if( keep.closed() )
      return;
   constexpr int GPB = 256 / G;
   const int gl = threadIdx.x % G;
   const int64_t stride = (int64_t) gridDim.x * GPB;
   if constexpr( R == 1 ) {
      for( int64_t r = first + (int64_t) blockIdx.x * GPB + threadIdx.x / G; r < last; r += stride ) {
         const int32_t b = off[ r ], e = off[ r + 1 ];
         A s = 0;
         for( int32_t k = b + gl; k < e; k += G ) {
            const int32_t c = col[ k ];
            if( c >= 0 )
               s = madd< EXACT >( s, toCompute( val[ k ], A() ), (A) x[ c ] );
         }
         if constexpr( G > 1 )
            s = groupSum< G >( s );
         if( gl == 0 )
            keep( r, s );
      }
      return;
   }
   for( int64_t r0 = first + (int64_t) blockIdx.x * GPB + threadIdx.x / G; r0 < last; r0 += stride * R ) {
      int32_t b[ R ], e[ R ];
#pragma unroll
      for( int j = 0; j < R; j++ ) {
         const int64_t r = r0 + j * stride;
         const bool valid = r < last;
         b[ j ] = valid ? off[ r ] : 0;
         e[ j ] = valid ? off[ r + 1 ] : 0;
      }
      int32_t c[ R ];
      V v[ R ];
#pragma unroll
      for( int j = 0; j < R; j++ ) {
         const int32_t k = b[ j ] + gl;
         const bool in = k < e[ j ];
         c[ j ] = in ? col[ k ] : -1;
         v[ j ] = in ? val[ k ] : V( 0 );
      }
      Y xv[ R ];
#pragma unroll
      for( int j = 0; j < R; j++ )
         xv[ j ] = x[ c[ j ] < 0 ? 0 : c[ j ] ];
      A s[ R ];
#pragma unroll
      for( int j = 0; j < R; j++ ) {
         s[ j ] = 0;
         if( c[ j ] >= 0 )
            s[ j ] = madd< EXACT >( s[ j ], toCompute( v[ j ], A() ), (A) xv[ j ] );
         for( int32_t k = b[ j ] + gl + G; k < e[ j ]; k += G ) {
            const int32_t ck = col[ k ];
            if( ck >= 0 )
               s[ j ] = madd< EXACT >( s[ j ], toCompute( val[ k ], A() ), (A) x[ ck ] );
         }
      }
#pragma unroll
      for( int j = 0; j < R; j++ ) {
         A t = s[ j ];
         if constexpr( G > 1 )
            t = groupSum< G >( t );
         const int64_t r = r0 + j * stride;
         if( gl == 0 && r < last )
            keep( r, t );
      }
   }
}

// ------------------------------------------------------------------ CSR scalar, staged

// CSRScalar (one lane per row, Kernels/CSRScalarKernel.hpp) with the elements staged through
// LDS.  Each wavefront owns 64 consecutive rows.  When their elements fit the wave's LDS slice
// (K x 64 of them: 512 FP64 or 1024 FP32 products), the wave loads them with coalesced loads,
// kStageLoads per lane in flight per round, stores each rounded product in LDS (a padding slot
// stores -0.0, the identity of IEEE addition, so adding it equals skipping it), and every lane
// then sums its own row left to right out of LDS: the Host traversal's result bit for bit, at
// streaming speed.  Waves whose rows hold more elements walk their rows in global memory, one
// lane per row, with the same arithmetic.  One wave per 64 rows and no grid-stride loop: at
// 160^3 FP64 neither a persistent grid that prefetches the next block's offsets (78-92 us by
// grid size) nor two 64-row blocks per wave (83 us; FP32 61.1 against 59.4 us) beat it (79 us).
constexpr int kStageLoads = 8;
constexpr int kShortRow = 8;  // mean row length up to which every CSR kernel takes the staged path

template< int K, typename V, typename A, typename Y >
__global__ __launch_bounds__( 256 ) void csrScalarStagedKernel( const int32_t* __restrict__ off,
                                                                const int32_t* __restrict__ col,
                                                                const V* __restrict__ val, const Y* __restrict__ x,
                                                                Keeper< A, Y > keep, int32_t first, int32_t last,
                                                                int remap )
{
   if( keep.closed() )
      return;
   __shared__ A lds[ 4 ][ K * kWave ];
   const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
   const int64_t r0 = first + ( logicalBlock( remap ) * 4 + wave ) * kWave;
   if( r0 >= last )  // uniform over the wave; the kernel has no workgroup barrier
      return;
   const int64_t r = r0 + lane;
   const int32_t used = (int32_t) min( (int64_t) kWave, last - r0 );  // rows of this wave, >= 1
   const bool mine = lane < used;
   const int32_t rb = mine ? off[ r ] : 0;
   const int32_t re = mine ? off[ r + 1 ] : 0;
   const int32_t b = __shfl( rb, 0, kWave );
   const int32_t size = __shfl( re, used - 1, kWave ) - b;
   if( size <= 0 ) {
      if( mine )
         keep( r, (A) 0 );
      return;
   }
   if( size <= K * kWave ) {
      // rounds of kStageLoads loads per lane (one for FP64, up to two for FP32) keep the
      // registers, hence the waves per SIMD, of the one-round case
#pragma unroll
      for( int q = 0; q < K / kStageLoads; q++ ) {
         const int32_t e0 = q * kStageLoads * kWave;
         if( e0 >= size )  // uniform over the wave
            break;
         int32_t c[ kStageLoads ];
         V v[ kStageLoads ];
#pragma unroll
         for( int k = 0; k < kStageLoads; k++ ) {
            const int32_t g = min( e0 + lane + k * kWave, size - 1 );
            c[ k ] = ldnt( col + b + g );
            v[ k ] = ldnt( val + b + g );
         }
         Y xv[ kStageLoads ];
#pragma unroll
         for( int k = 0; k < kStageLoads; k++ )
            xv[ k ] = c[ k ] >= 0 ? x[ c[ k ] ] : Y( 0 );  // x may be empty when every slot is padding
#pragma unroll
         for( int k = 0; k < kStageLoads; k++ ) {
            const int32_t g = e0 + lane + k * kWave;
            if( g < size )
               lds[ wave ][ g ] = c[ k ] >= 0 ? toCompute( v[ k ], A() ) * (A) xv[ k ] : (A) -0.0;
         }
      }
      waveSync();
      if( mine ) {
         A s = 0;
         for( int32_t k = rb - b; k < re - b; k++ )
            s += lds[ wave ][ k ];
         keep( r, s );
      }
      return;
   }
   if( mine ) {
      A s = 0;
      for( int32_t k = rb; k < re; k++ ) {
         const int32_t ck = col[ k ];
         if( ck >= 0 )
            s = mulAddExact( s, toCompute( val[ k ], A() ), (A) x[ ck ] );
      }
      keep( r, s );
   }
}

// ------------------------------------------------------------------ CSR-Adaptive

// One wavefront per block descriptor (the scheme of Kernels/CSRAdaptiveKernelView.hpp:33-170,
// with wave64 budgets).  STREAM: the wave loads the elements of several short rows with fully
// coalesced loads (rounds of kStageLoads per lane in flight), keeps the products in
// its own LDS slice, then every lane sums whole rows out of LDS.  VECTOR: one row per wave.
// LONG: one row split over the waves left in the workgroup, partial sums combined in LDS in
// a fixed order.  One descriptor per wave and no grid-stride loop, so the single workgroup
// barrier (needed only by LONG rows) is uniform.
constexpr int kPreSets = 4;  // STREAM blocks: offsets of the first 4 x 64 rows are preloaded

template< int STREAM, typename V, typename A, typename Y >
__global__ __launch_bounds__( 256 ) void csrAdaptiveKernel( int32_t nblocks, const int4* __restrict__ desc,
                                                            const int32_t* __restrict__ off,
                                                            const int32_t* __restrict__ col,
                                                            const V* __restrict__ val, const Y* __restrict__ x,
                                                            Keeper< A, Y > keep, int remap )
{
   if( keep.closed() )
      return;
   constexpr int K = STREAM / kWave;
   __shared__ A lds[ 4 ][ STREAM ];
   __shared__ A longPart[ 4 ];
   const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
   // remap: XCD x (blockIdx % 8) takes the x-th eighth of the descriptors, in order
   const int64_t lb = remap ? logicalBlock( 1 ) : blockIdx.x;
   const int64_t di = lb * 4 + wave;
   const bool active = di < nblocks;
   const int4 d = active ? desc[ di ] : make_int4( BLOCK_VECTOR, 0, 0, 0 );
   const int type = d.x & 0xff;
   const int32_t firstRow = d.y;
   const int32_t b = active ? off[ firstRow ] : 0;
   A vsum = 0;
   if( active && type == BLOCK_STREAM && d.z == 0 ) {
      for( int32_t r = lane; r < d.w; r += kWave )  // a block of empty rows
         keep( firstRow + r, (A) 0 );
   }
   else if( active && type == BLOCK_STREAM ) {
      const int32_t size = d.z;
      // offsets of the first kPreRows rows of the block, fetched alongside the element stream
      int32_t rbPre[ kPreSets ];
#pragma unroll
      for( int q = 0; q < kPreSets; q++ ) {
         const int32_t r = lane + q * kWave;
         rbPre[ q ] = r <= d.w ? off[ firstRow + r ] : 0;
      }
      // rounds of kStageLoads loads per lane (one for FP64, up to two for FP32 blocks of 1024
      // elements): the FP32 kernel keeps the FP64 kernel's registers and waves per SIMD
#pragma unroll
      for( int q = 0; q < K / kStageLoads; q++ ) {
         const int32_t e0 = q * kStageLoads * kWave;
         if( e0 >= size )  // uniform over the wave
            break;
         int32_t c[ kStageLoads ];
         V v[ kStageLoads ];
#pragma unroll
         for( int k = 0; k < kStageLoads; k++ ) {
            const int32_t e = min( e0 + lane + k * kWave, size - 1 );
            c[ k ] = ldnt( col + b + e );
            v[ k ] = ldnt( val + b + e );
         }
         Y xv[ kStageLoads ];
#pragma unroll
         for( int k = 0; k < kStageLoads; k++ )
            xv[ k ] = x[ c[ k ] < 0 ? 0 : c[ k ] ];
#pragma unroll
         for( int k = 0; k < kStageLoads; k++ ) {
            const int32_t e = e0 + lane + k * kWave;
            if( e < size )
               lds[ wave ][ e ] = c[ k ] >= 0 ? toCompute( v[ k ], A() ) * (A) xv[ k ] : (A) 0;
         }
      }
      waveSync();
#pragma unroll
      for( int q = 0; q < kPreSets; q++ ) {
         // row r = lane + 64 q ends where row r + 1 starts: the next lane's offset, or the
         // first of the next set for lane 63
         const int32_t nextLane = __shfl( rbPre[ q ], ( lane + 1 ) % kWave, kWave );
         const int32_t nextSet = q + 1 < kPreSets ? __shfl( rbPre[ q + 1 < kPreSets ? q + 1 : q ], 0, kWave ) : 0;
         const int32_t r = lane + q * kWave;
         if( r < d.w ) {
            const int32_t rb = rbPre[ q ] - b;
            const int32_t re = ( lane + 1 < kWave ? nextLane : q + 1 < kPreSets ? nextSet : off[ firstRow + r + 1 ] ) - b;
            A s = 0;
            for( int32_t k = rb; k < re; k++ )
               s += lds[ wave ][ k ];
            keep( firstRow + r, s );
         }
      }
      for( int32_t r = lane + kPreSets * kWave; r < d.w; r += kWave ) {
         const int32_t rb = off[ firstRow + r ] - b;
         const int32_t re = off[ firstRow + r + 1 ] - b;
         A s = 0;
         for( int32_t k = rb; k < re; k++ )
            s += lds[ wave ][ k ];
         keep( firstRow + r, s );
      }
   }
   else if( active ) {
      const int32_t e = off[ firstRow + 1 ];
      const int32_t widx = type == BLOCK_LONG ? ( d.x >> 8 ) & 0xff : 0;
      const int32_t wcnt = type == BLOCK_LONG ? ( d.x >> 16 ) & 0xff : 1;
      A p = 0;
      for( int32_t k = b + widx * kWave + lane; k < e; k += kWave * wcnt ) {
         const int32_t c = ldnt( col + k );
         if( c >= 0 )
            p += toCompute( ldnt( val + k ), A() ) * (A) x[ c ];
      }
      vsum = groupSum< kWave >( p );
      if( type == BLOCK_VECTOR && lane == 0 )
         keep( firstRow, vsum );
      if( type == BLOCK_LONG && lane == 0 )
         longPart[ wave ] = vsum;
   }
   __syncthreads();
   if( active && type == BLOCK_LONG && ( ( d.x >> 8 ) & 0xff ) == 0 && lane == 0 ) {
      const int32_t wcnt = ( d.x >> 16 ) & 0xff;
      A s = 0;
      for( int32_t w = 0; w < wcnt; w++ )
         s += longPart[ wave + w ];
      keep( firstRow, s );
   }
}

// ------------------------------------------------------------------ Ellpack / SlicedEllpack

// One lane per row, column-major storage: consecutive lanes read consecutive addresses
// for every local index (EllpackView.hpp:361-378, SlicedEllpackView.hpp:248-290).  The
// slots of a row are read eight at a time with unconditional loads (indexes past the row
// end are clamped to its last slot, padding gathers x[0]), so 24 loads per lane are in
// flight; the sum still runs left to right over the non-padding slots only, which is the
// Host result bit for bit.  Each workgroup owns `rowsPerBlock` consecutive rows.
constexpr int kChunk = 8;
constexpr int kRowsInFlight = 2;  // rows per lane per iteration of the padded kernels

// Two vertically adjacent slots (rows r, r+1 at one local index) are one aligned 8-byte
// index load and one 2-element value load.
typedef double double2v __attribute__( ( ext_vector_type( 2 ) ) );
typedef float float2v __attribute__( ( ext_vector_type( 2 ) ) );
template< typename V >
struct Pair;
template<>
struct Pair< double >
{
   using T = double2v;
};
template<>
struct Pair< float >
{
   using T = float2v;
};
template<>
struct Pair< bf16 >
{
   using T = unsigned int;
};
__device__ __forceinline__ void split( double2v p, double& a, double& b )
{
   a = p.x;
   b = p.y;
}
__device__ __forceinline__ void split( float2v p, float& a, float& b )
{
   a = p.x;
   b = p.y;
}
__device__ __forceinline__ void split( unsigned int p, bf16& a, bf16& b )
{
   a = __ushort_as_bfloat16( (unsigned short) ( p & 0xffff ) );
   b = __ushort_as_bfloat16( (unsigned short) ( p >> 16 ) );
}
template< typename V >
__device__ __forceinline__ typename Pair< V >::T ldnt2( const V* p )
{
   using P = typename Pair< V >::T;
   return __builtin_nontemporal_load( reinterpret_cast< const P* >( p ) );
}
__device__ __forceinline__ int2 ldnt2( const int32_t* p )
{
   const long long v = __builtin_nontemporal_load( reinterpret_cast< const long long* >( p ) );
   return make_int2( (int32_t) ( v & 0xffffffff ), (int32_t) ( v >> 32 ) );
}

// One or two (PAIR) rows per lane.  Slot index of row r at local index l: base( r ) + step * l.
template< bool PAIR, typename V, typename A, typename Y, typename Slots >
__device__ __forceinline__ void paddedRows( const Slots& slots, const int32_t* __restrict__ col,
                                            const V* __restrict__ val, const Y* __restrict__ x, Keeper< A, Y > keep,
                                            int32_t first, RowSeq seq )
{
   if constexpr( ! PAIR ) {
      // kRowsInFlight rows of the sequence per lane (r, r + step, ...), their loads issued
      // together; each row keeps its own left-to-right sum
      constexpr int Q = kRowsInFlight;
      const int64_t step = slots.step();
      int64_t r = seq.start;
      for( ; r + ( Q - 1 ) * seq.step < seq.end; r += Q * seq.step ) {
         int64_t b[ Q ];
         int32_t w[ Q ];
         int32_t wmax = 0;
#pragma unroll
         for( int q = 0; q < Q; q++ ) {
            b[ q ] = slots.base( r + q * seq.step );
            w[ q ] = slots.width( r + q * seq.step );
            wmax = max( wmax, w[ q ] );
         }
         A s[ Q ];
#pragma unroll
         for( int q = 0; q < Q; q++ )
            s[ q ] = 0;
         for( int32_t l0 = 0; l0 < wmax; l0 += kChunk ) {
            int32_t c[ Q ][ kChunk ];
            V v[ Q ][ kChunk ];
#pragma unroll
            for( int q = 0; q < Q; q++ )
#pragma unroll
               for( int j = 0; j < kChunk; j++ ) {
                  const bool in = l0 + j < w[ q ];
                  const int64_t g = b[ q ] + step * ( l0 + j );
                  c[ q ][ j ] = in ? ldnt( col + g ) : -1;
                  v[ q ][ j ] = in ? ldnt( val + g ) : V( 0 );
               }
            Y xv[ Q ][ kChunk ];
#pragma unroll
            for( int q = 0; q < Q; q++ )
#pragma unroll
               for( int j = 0; j < kChunk; j++ )
                  xv[ q ][ j ] = x[ c[ q ][ j ] < 0 ? 0 : c[ q ][ j ] ];
#pragma unroll
            for( int q = 0; q < Q; q++ )
#pragma unroll
               for( int j = 0; j < kChunk; j++ )
                  if( c[ q ][ j ] >= 0 )
                     s[ q ] = mulAddExact( s[ q ], toCompute( v[ q ][ j ], A() ), (A) xv[ q ][ j ] );
         }
#pragma unroll
         for( int q = 0; q < Q; q++ )
            if( r + q * seq.step >= first )
               keep( r + q * seq.step, s[ q ] );
      }
      seq.start = r;  // the rows left over run one per iteration below
   }
   for( int64_t r = seq.start; r < seq.end; r += seq.step ) {
      const int64_t b = slots.base( r );
      const int32_t width = slots.width( r );
      const int64_t step = slots.step();
      A s0 = 0, s1 = 0;
      for( int32_t l0 = 0; l0 < width; l0 += kChunk ) {
         if constexpr( PAIR ) {
            int2 c[ kChunk ];
            typename Pair< V >::T v[ kChunk ];
#pragma unroll
            for( int j = 0; j < kChunk; j++ ) {
               const int64_t g = b + step * min( l0 + j, width - 1 );
               c[ j ] = ldnt2( col + g );
               v[ j ] = ldnt2( val + g );
            }
            Y x0[ kChunk ], x1[ kChunk ];
#pragma unroll
            for( int j = 0; j < kChunk; j++ ) {
               x0[ j ] = x[ c[ j ].x < 0 ? 0 : c[ j ].x ];
               x1[ j ] = x[ c[ j ].y < 0 ? 0 : c[ j ].y ];
            }
#pragma unroll
            for( int j = 0; j < kChunk; j++ ) {
               V a, bb;
               split( v[ j ], a, bb );
               if( l0 + j < width && c[ j ].x >= 0 )
                  s0 = mulAddExact( s0, toCompute( a, A() ), (A) x0[ j ] );
               if( l0 + j < width && c[ j ].y >= 0 )
                  s1 = mulAddExact( s1, toCompute( bb, A() ), (A) x1[ j ] );
            }
         }
         else {
            int32_t c[ kChunk ];
            V v[ kChunk ];
#pragma unroll
            for( int j = 0; j < kChunk; j++ ) {
               const int64_t g = b + step * min( l0 + j, width - 1 );
               c[ j ] = ldnt( col + g );
               v[ j ] = ldnt( val + g );
            }
            Y xv[ kChunk ];
#pragma unroll
            for( int j = 0; j < kChunk; j++ )
               xv[ j ] = x[ c[ j ] < 0 ? 0 : c[ j ] ];
#pragma unroll
            for( int j = 0; j < kChunk; j++ )
               if( l0 + j < width && c[ j ] >= 0 )
                  s0 = mulAddExact( s0, toCompute( v[ j ], A() ), (A) xv[ j ] );
         }
      }
      if( r >= first )
         keep( r, s0 );
      if( PAIR && r + 1 < seq.end )
         keep( r + 1, s1 );
   }
}

struct EllpackSlots
{
   int32_t segmentSize, aligned;
   __device__ int64_t base( int64_t r ) const { return r; }
   __device__ int32_t width( int64_t ) const { return segmentSize; }
   __device__ int64_t step() const { return aligned; }
};

struct SlicedSlots
{
   const int32_t* sliceOffsets;
   const int32_t* sliceSizes;
   int32_t logSlice;  // log2 of the slice size (TNL's SliceSize, 32 by default)
   __device__ int64_t base( int64_t r ) const
   {
      return sliceOffsets[ r >> logSlice ] + ( r & ( ( 1 << logSlice ) - 1 ) );
   }
   __device__ int32_t width( int64_t r ) const { return sliceSizes[ r >> logSlice ]; }
   __device__ int64_t step() const { return (int64_t) 1 << logSlice; }
};

template< bool PAIR, typename V, typename A, typename Y >
__global__ __launch_bounds__( 256 ) void ellpackKernel( EllpackSlots slots, const int32_t* __restrict__ col,
                                                        const V* __restrict__ val, const Y* __restrict__ x,
                                                        Keeper< A, Y > keep, int32_t first, int32_t last,
                                                        int32_t rowsPerBlock, int dist )
{
   if( keep.closed() )
      return;
   const int64_t base = PAIR ? ( first & ~1 ) : first;
   paddedRows< PAIR >( slots, col, val, x, keep, first, rowSequence< PAIR ? 2 : 1 >( base, last, rowsPerBlock, dist ) );
}

template< bool PAIR, typename V, typename A, typename Y >
__global__ __launch_bounds__( 256 ) void slicedKernel( SlicedSlots slots, const int32_t* __restrict__ col,
                                                       const V* __restrict__ val, const Y* __restrict__ x,
                                                       Keeper< A, Y > keep, int32_t first, int32_t last,
                                                       int32_t rowsPerBlock, int dist )
{
   if( keep.closed() )
      return;
   const int64_t base = PAIR ? ( first & ~1 ) : first;
   paddedRows< PAIR >( slots, col, val, x, keep, first, rowSequence< PAIR ? 2 : 1 >( base, last, rowsPerBlock, dist ) );
}

// ------------------------------------------------------------------ ChunkedEllpack

// One 256-lane workgroup per slice, one lane per chunk (the format is built around 256
// chunks per slice): chunk c's elements sit at pointer + c + 256*k, coalesced across the
// workgroup.  Chunk sums meet in LDS and each row adds up its chunk range
// (rowToChunkMapping is an inclusive per-slice prefix of chunk counts).  A chunk's padding
// follows its elements, so with `stopAtPadding` a wavefront stops reading once every lane's
// chunk has reached padding: a slice's chunk size is set by its longest row, and short rows
// beside long ones would otherwise stream their padding in full.
template< typename V, typename A, typename Y >
__global__ __launch_bounds__( 256 ) void chunkedKernel( int32_t nslices, const int4* __restrict__ slices,
                                                        const int32_t* __restrict__ rowToChunk,
                                                        const int32_t* __restrict__ col,
                                                        const V* __restrict__ val, const Y* __restrict__ x,
                                                        Keeper< A, Y > keep, int32_t first, int32_t last,
                                                        int stopAtPadding, int remap )
{
   if( keep.closed() )
      return;
   __shared__ A part[ kChunksInSlice ];
   const int t = threadIdx.x;
   // remap: XCD x (blockIdx % 8) takes the x-th eighth of the slices, in order
   for( int64_t sl = logicalBlock( remap ); sl < nslices; sl += gridDim.x ) {
      const int4 s = slices[ sl ];  // size, chunkSize, firstSegment, pointer
      if( s.z >= last || s.z + s.x <= first )
         continue;  // uniform over the workgroup
      // every row owns exactly one chunk when the slice uses as many chunks as it has rows
      const bool onePerRow = s.x > 0 && rowToChunk[ s.z + s.x - 1 ] == s.x;
      A p = 0;
      for( int32_t k0 = 0; k0 < s.y; k0 += kChunk ) {
         int32_t c[ kChunk ];
         V v[ kChunk ];
#pragma unroll
         for( int j = 0; j < kChunk; j++ ) {
            const int64_t g = (int64_t) s.w + t + (int64_t) kChunksInSlice * min( k0 + j, s.y - 1 );
            c[ j ] = ldnt( col + g );
            v[ j ] = ldnt( val + g );
         }
         Y xv[ kChunk ];
#pragma unroll
         for( int j = 0; j < kChunk; j++ )
            xv[ j ] = x[ c[ j ] < 0 ? 0 : c[ j ] ];
#pragma unroll
         for( int j = 0; j < kChunk; j++ )
            if( k0 + j < s.y && c[ j ] >= 0 )
               p += toCompute( v[ j ], A() ) * (A) xv[ j ];
         // the last slot read is padding: so is the rest of this lane's chunk
         if( stopAtPadding && __all( c[ kChunk - 1 ] < 0 ) )
            break;  // wave-uniform; the loop holds no barrier
      }
      if( onePerRow ) {  // uniform: chunk t is row s.z + t, no exchange through LDS
         const int32_t row = s.z + t;
         if( t < s.x && row >= first && row < last )
            keep( row, (A) 0 + p );
         continue;
      }
      part[ t ] = p;
      __syncthreads();
      if( t < s.x ) {
         const int32_t row = s.z + t;
         if( row >= first && row < last ) {
            const int32_t cb = t == 0 ? 0 : rowToChunk[ row - 1 ], ce = rowToChunk[ row ];
            A acc = 0;
            for( int32_t c = cb; c < ce; c++ )
               acc += part[ c ];
            keep( row, acc );
         }
      }
      __syncthreads();
   }
}

// ------------------------------------------------------------------ BiEllpack

// A wavefront takes two strips of 32 rows, one per half-wave.  Group g of a strip is a
// dense height x width block (height 32 >> g) stored column-major, i.e. contiguous: lane l
// streams elements l, l+32, ... which all belong to rank l % height, then lanes with equal
// l % height are folded with xor-shuffles.  Lane l < height keeps rank l's running sum; at
// the end each lane fetches the sum of its own row's rank and stores it, so y is written
// coalesced.  The rank comes from a one-byte copy of rowPermArray[row] - 32 * strip (4 MB
// instead of 16 MB at 160^3).
constexpr int kBiUnroll = 8;

template< typename V, typename A, typename Y >
__global__ __launch_bounds__( 256 ) void biellpackKernel( int32_t n, int32_t strips, const uint8_t* __restrict__ rowRank,
                                                          const int32_t* __restrict__ groupPointers,
                                                          const int32_t* __restrict__ col,
                                                          const V* __restrict__ val, const Y* __restrict__ x,
                                                          Keeper< A, Y > keep, int32_t first, int32_t last )
{
   if( keep.closed() )
      return;
   const int lane = threadIdx.x % kWave, half = lane / 32, l = lane % 32;
   const int32_t wave = ( blockIdx.x * blockDim.x + threadIdx.x ) / kWave;
   const int32_t waves = gridDim.x * ( blockDim.x / kWave );
   // group pointers of the half-wave's strip are fetched one iteration ahead
   int32_t gp[ kLogStrip + 1 ];
   auto fetch = [ & ]( int32_t strip, int32_t( &g )[ kLogStrip + 1 ] ) {
#pragma unroll
      for( int k = 0; k <= kLogStrip; k++ )
         g[ k ] = strip < strips ? groupPointers[ (int64_t) strip * kGroupsPerStrip + k ] : 0;
   };
   fetch( 2 * wave + half, gp );
   for( int32_t pair = wave; 2 * pair < strips; pair += waves ) {
      const int32_t strip = 2 * pair + half;
      const bool valid = strip < strips;
      const int64_t row = (int64_t) strip * kStrip + l;
      const bool mine = valid && row < n;
      const int rank = mine ? rowRank[ row ] : 0;
      int32_t gpNext[ kLogStrip + 1 ];
      fetch( 2 * ( pair + waves ) + half, gpNext );
      A acc = 0;
      if( valid ) {
         // group 5 (height 1) is always empty: the widths are taken at ranks 15, 7, 3, 1, 0
#pragma unroll
         for( int grp = 0; grp < kLogStrip; grp++ ) {
            const int32_t base = gp[ grp ], size = gp[ grp + 1 ] - base;
            const int h = kStrip >> grp;
            A p = 0;
            // 8 elements per lane in flight: a 32-row group up to 8 wide is one round trip
            for( int32_t e0 = l; e0 < size; e0 += kBiUnroll * kStrip ) {
               int32_t c[ kBiUnroll ];
               V v[ kBiUnroll ];
#pragma unroll
               for( int j = 0; j < kBiUnroll; j++ ) {
                  const int32_t e = min( e0 + j * kStrip, size - 1 );
                  c[ j ] = ldnt( col + base + e );
                  v[ j ] = ldnt( val + base + e );
               }
               Y xv[ kBiUnroll ];
#pragma unroll
               for( int j = 0; j < kBiUnroll; j++ )
                  xv[ j ] = x[ c[ j ] < 0 ? 0 : c[ j ] ];
#pragma unroll
               for( int j = 0; j < kBiUnroll; j++ )
                  if( e0 + j * kStrip < size && c[ j ] >= 0 )
                     p += toCompute( v[ j ], A() ) * (A) xv[ j ];
            }
            for( int o = h; o < kStrip; o <<= 1 )
               p += __shfl_xor( p, o, kWave );
            if( l < h )
               acc += p;
         }
      }
      const A v = __shfl( acc, half * 32 + rank, kWave );
      if( mine && row >= first && row < last )
         keep( row, v );
#pragma unroll
      for( int k = 0; k <= kLogStrip; k++ )
         gp[ k ] = gpNext[ k ];
   }
}

// Rows without any stored element: y = keep( 0 ).
template< typename A, typename Y >
__global__ __launch_bounds__( 256 ) void zeroRowsKernel( Keeper< A, Y > keep, int32_t first, int32_t last )
{
   if( keep.closed() )
      return;
   const int64_t stride = (int64_t) gridDim.x * blockDim.x;
   for( int64_t r = first + (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < last; r += stride )
      keep( r, (A) 0 );
}

// ------------------------------------------------------------------ dispatch

namespace {

template< typename V, typename A, typename Y >
void launchTyped( const Matrix& M, const Y* x, Y* y, double alpha, double beta, int32_t first, int32_t last,
                  const Y* diag, const int32_t* gate, int gateMode, hipStream_t st, const Y* rhs = nullptr,
                  const Y* prev = nullptr )
{
   Keeper< A, Y > keep{ y, (A) alpha, (A) beta, diag, 0, gate, gate ? gateMode : 0 };
   if( rhs ) {
      keep.mode = 5;
      keep.rhs = rhs;
      keep.prev = prev;
   }
   else if( diag )
      keep.mode = 4;
   else if( beta == 0.0 )
      keep.mode = alpha == 1.0 ? 0 : 1;
   else
      keep.mode = alpha == 1.0 ? 2 : 3;
   const V* val = reinterpret_cast< const V* >( M.values.ptr );
   const int32_t* col = M.columns.ptr;
   const int64_t rows = last - first;
   if( rows <= 0 )
      return;
   auto launchStaged = [ & ] {
      constexpr int K = kAdaptiveStreamBytesPerWave / sizeof( A ) / kWave;
      unsigned grid = (unsigned) ( ( rows + 255 ) / 256 );
      if( M.xcdRemap )
         grid = ( grid + kXcds - 1 ) / kXcds * kXcds;  // logicalBlock needs a multiple of 8
      csrScalarStagedKernel< K, V, A, Y >
         <<< grid, 256, 0, st >>>( M.offsets.ptr, col, val, x, keep, first, last, M.xcdRemap ? 1 : 0 );
   };
   if( M.streamY >= 0 )
      keep.streamY = M.streamY != 0;
   if( M.cols == 0 ) {
      // every stored slot is padding and x is empty: the kernels below gather x[0] for padding
      // slots, so a matrix without columns takes the all-zero product without touching x
      zeroRowsKernel< A, Y ><<< gridFor( rows, 256 ), 256, 0, st >>>( keep, first, last );
      TNLHIP_CHECK( hipGetLastError() );
      return;
   }
   if( M.stripes && kindOf( M.format ) == KIND_CSR ) {
      stripeLaunch< V, A, Y >( M, x, keep, first, last, st );
      return;
   }
   switch( M.format ) {
      case TNLHIP_CSR_SCALAR:
         launchStaged();
         break;
      case TNLHIP_CSR_VECTOR:
      case TNLHIP_CSR_LIGHT:
      case TNLHIP_CSR_HYBRID:
      case TNLHIP_CSR_ADAPTIVE: {
         if( M.format == TNLHIP_CSR_ADAPTIVE && first == 0 && last == M.rows && M.adaptiveBlocks > 0 ) {
            constexpr int STREAM = kAdaptiveStreamBytesPerWave / sizeof( A );
            // y through L2 on request (STREAM blocks end rows anywhere in a cache line of y, so
            // non-temporal stores write some lines twice); non-temporal by default, which measured
            // 79-80 us at 160^3 FP64 against 81-83 us through L2
            keep.streamY = M.streamY != 0;
            unsigned groups = ( M.adaptiveBlocks + 3 ) / 4;
            if( M.xcdRemap )
               groups = ( groups + kXcds - 1 ) / kXcds * kXcds;  // logicalBlock needs a multiple of 8
            csrAdaptiveKernel< STREAM, V, A, Y ><<< groups, 256, 0, st >>>(
               M.adaptiveBlocks, reinterpret_cast< const int4* >( M.adaptive.ptr ), M.offsets.ptr, col, val, x, keep,
               M.xcdRemap ? 1 : 0 );
            break;
         }
         // rows of at most kShortRow stored elements on average: most lanes of a cooperative row
         // group would idle (a wavefront per 7-element row uses 7 of 64), so unless lanes_per_row
         // was set the rows take CSRScalar's staged kernel (the Host result bit for bit)
         if( ! M.lanesExplicit && M.meta.allocated <= (int64_t) kShortRow * M.rows ) {
            launchStaged();
            break;
         }
         const int G = M.format == TNLHIP_CSR_VECTOR && ! M.lanesExplicit ? 64 : M.lanes;
         const int gpb = 256 / G;
         const unsigned grid = gridFor( rows, gpb );
         switch( G ) {
            case 1:
               csrGroupedKernel< 1, 1, false, V, A, Y ><<< grid, 256, 0, st >>>( M.offsets.ptr, col, val, x, keep, first, last );
               break;
            case 2:
               csrGroupedKernel< 2, 4, false, V, A, Y ><<< grid, 256, 0, st >>>( M.offsets.ptr, col, val, x, keep, first, last );
               break;
            case 4:
               csrGroupedKernel< 4, 4, false, V, A, Y ><<< grid, 256, 0, st >>>( M.offsets.ptr, col, val, x, keep, first, last );
               break;
            case 8:
               csrGroupedKernel< 8, 4, false, V, A, Y ><<< grid, 256, 0, st >>>( M.offsets.ptr, col, val, x, keep, first, last );
               break;
            case 16:
               csrGroupedKernel< 16, 2, false, V, A, Y ><<< grid, 256, 0, st >>>( M.offsets.ptr, col, val, x, keep, first, last );
               break;
            case 32:
               csrGroupedKernel< 32, 2, false, V, A, Y ><<< grid, 256, 0, st >>>( M.offsets.ptr, col, val, x, keep, first, last );
               break;
            default:
               csrGroupedKernel< 64, 1, false, V, A, Y ><<< grid, 256, 0, st >>>( M.offsets.ptr, col, val, x, keep, first, last );
               break;
         }
         break;
      }
      case TNLHIP_ELLPACK:
      case TNLHIP_SLICED_ELLPACK: {
         // grid a multiple of the XCD count; contiguous row blocks when dist != 0
         // a row pair is one 8- / 16-byte load of two vertically adjacent slots: SlicedEllpack
         // with slice size 1 has no such neighbours (row r + 1 lies in the next slice), so it
         // takes one row per lane whatever pair_rows says
         const bool pair = M.pairRows != 0 && ( M.format == TNLHIP_ELLPACK || M.meta.sliceSize >= 2 );
         if( ! knownDistribution( M.xcdRemap ) )
            throw Error( TNLHIP_ERROR_INVALID_ARGUMENT, "unknown row distribution" );
         const PaddedLaunch pl = paddedLaunch( first, last, pair, M.blocksPerCu, deviceProps().cus );
         const unsigned grid = (unsigned) pl.grid;
         const int64_t per = pl.per;
         const EllpackSlots es{ M.meta.segmentSize, M.meta.alignedSize };
         const SlicedSlots ss{ M.sliceOffsets.ptr, M.sliceSizes.ptr, log2Exact( M.meta.sliceSize ) };
         if( M.format == TNLHIP_ELLPACK ) {
            if( pair )
               ellpackKernel< true, V, A, Y ><<< grid, 256, 0, st >>>( es, col, val, x, keep, first, last, (int32_t) per,
                                                                        M.xcdRemap );
            else
               ellpackKernel< false, V, A, Y ><<< grid, 256, 0, st >>>( es, col, val, x, keep, first, last,
                                                                         (int32_t) per, M.xcdRemap );
         }
         else {
            if( pair )
               slicedKernel< true, V, A, Y ><<< grid, 256, 0, st >>>( ss, col, val, x, keep, first, last, (int32_t) per,
                                                                       M.xcdRemap );
            else
               slicedKernel< false, V, A, Y ><<< grid, 256, 0, st >>>( ss, col, val, x, keep, first, last,
                                                                        (int32_t) per, M.xcdRemap );
         }
         break;
      }
      case TNLHIP_CHUNKED_ELLPACK:
         if( M.meta.numberOfSlices > 0 )
            // one workgroup per slice: at 160^3 70.6 us, against 78.1 / 72.7 us for grids of 8 / 32
            // workgroups per CU striding over the slices
            chunkedKernel< V, A, Y ><<< M.xcdRemap ? ( M.meta.numberOfSlices + kXcds - 1 ) / kXcds * kXcds
                                                   : M.meta.numberOfSlices,
                                        256, 0, st >>>(
               M.meta.numberOfSlices, reinterpret_cast< const int4* >( M.slices.ptr ), M.rowToChunk.ptr, col, val, x,
               keep, first, last, M.trailingPadding ? 1 : 0, M.xcdRemap ? 1 : 0 );
         break;
      case TNLHIP_BIELLPACK: {
         const int32_t strips = M.meta.virtualRows / kStrip;
         const int32_t pairs = ( strips + 1 ) / 2;
         if( strips > 0 && M.meta.allocated > 0 )
            // 32 workgroups per CU (two strip pairs per wave at 160^3): 79.8 us, against 83.0 us
            // at 8 and 84.3 us with one workgroup per 4 pairs
            biellpackKernel< V, A, Y ><<< gridFor( pairs, 4, 32 ), 256, 0, st >>>(
               M.rows, strips, M.biRank.ptr, M.groupPointers.ptr, col, val, x, keep, first, last );
         else
            // no stored element at all (TNL's Host path returns without touching y here,
            // BiEllpackView.hpp:248-249); every row of A x is zero
            zeroRowsKernel< A, Y ><<< gridFor( rows, 256 ), 256, 0, st >>>( keep, first, last );
         break;
      }
      default: throw Error( TNLHIP_ERROR_INVALID_ARGUMENT, "unknown format" );
   }
   TNLHIP_CHECK( hipGetLastError() );
}

}  // namespace

void spmvLaunchGated( const Matrix& A, const void* x, void* y, double alpha, double beta, const void* diag,
                      const int32_t* gate, int gateMode, hipStream_t stream, int32_t first, int32_t last )
{
   if( last < 0 )
      last = A.rows;
   if( A.dtype == TNLHIP_BINARY ) {
      if( A.symmetric || diag )
         throw Error( TNLHIP_ERROR_UNSUPPORTED, "symmetric binary matrices and preconditioned binary products" );
      spmvBinary( A, x, y, alpha, beta, first, last, Gate{ gate, gate ? gateMode : 0 }, stream );
      return;
   }
   if( A.symmetric ) {
      spmvSymmetric( A, x, y, alpha, beta, first, last, diag, Gate{ gate, gate ? gateMode : 0 }, stream );
      return;
   }
   switch( A.dtype ) {
      case TNLHIP_F64:
         launchTyped< double, double, double >( A, (const double*) x, (double*) y, alpha, beta, first, last,
                                                (const double*) diag, gate, gateMode, stream );
         break;
      case TNLHIP_F32:
         launchTyped< float, float, float >( A, (const float*) x, (float*) y, alpha, beta, first, last,
                                             (const float*) diag, gate, gateMode, stream );
         break;
      case TNLHIP_BF16:
         launchTyped< bf16, float, float >( A, (const float*) x, (float*) y, alpha, beta, first, last,
                                            (const float*) diag, gate, gateMode, stream );
         break;
      default: throw Error( TNLHIP_ERROR_INVALID_ARGUMENT, "unknown value type" );
   }
}

void spmvJacobiStepGated( const Matrix& P, const void* z, void* znew, const void* rhs, const void* diag,
                          const int32_t* gate, int gateMode, hipStream_t stream )
{
   if( P.dtype == TNLHIP_BINARY || P.symmetric )
      throw Error( TNLHIP_ERROR_UNSUPPORTED, "the Jacobi-polynomial preconditioner needs a general real-valued matrix" );
   TNLHIP_REQUIRE( z != znew && rhs != znew, "a Jacobi step writes a vector of its own" );
   switch( P.dtype ) {
      case TNLHIP_F64:
         launchTyped< double, double, double >( P, (const double*) z, (double*) znew, 1.0, 0.0, 0, P.rows,
                                                (const double*) diag, gate, gateMode, stream, (const double*) rhs,
                                                (const double*) z );
         break;
      case TNLHIP_F32:
         launchTyped< float, float, float >( P, (const float*) z, (float*) znew, 1.0, 0.0, 0, P.rows,
                                             (const float*) diag, gate, gateMode, stream, (const float*) rhs,
                                             (const float*) z );
         break;
      default:
         launchTyped< bf16, float, float >( P, (const float*) z, (float*) znew, 1.0, 0.0, 0, P.rows,
                                            (const float*) diag, gate, gateMode, stream, (const float*) rhs,
                                            (const float*) z );
         break;
   }
}

void spmvLaunch( const Matrix& A, const void* x, void* y, double alpha, double beta, int32_t first, int32_t last,
                 const void* diag, hipStream_t stream )
{
   spmvLaunchGated( A, x, y, alpha, beta, diag, nullptr, 0, stream, first, last );
}

}  // namespace tnlhip
