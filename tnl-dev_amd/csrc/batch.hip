// Solve server: many uneven systems, each with 1-16 right-hand sides, solved together by one
// fixed sequence of persistent kernels that a hipGraph replays.
//
// Each request is (A + shift I) X = B for a resident matrix A, k right-hand sides, FP64
// iterative refinement over an FP32 conjugate-gradient inner solve (the recurrences of
// Solvers/Linear/CG.hpp:17-107, unpreconditioned).  There is no batched counterpart in the
// reference: TNL solves one system per solve() call.
//
// Layout in HBM:
//   matrices     SlicedEllpack (slice 32, column-major: SlicedEllpackView.hpp:108-132), one
//                column array and FP64 + FP32 value arrays shared by all resident matrices
//   vectors      per request, n x k entries interleaved (entry = row * k + j):
//                FP64 x, b, r and FP32 d, rr, p, q
//   tiles        a tile is one workgroup's share of one request: 256 / kp rows times kp lanes
//                (kp = k rounded up to a power of two); tiles of a request are numbered in row
//                order and every reduction of a request sums its tiles' partials in that order,
//                so a request's result does not depend on what else is resident.
//
// One iteration of the replayed sequence runs six tile kernels over all tiles; each acts on
// the right-hand sides in one phase (per-system phase machine below), requests with no system
// in that phase skip their tiles, and each reduction is finished by a second kernel with one
// workgroup per request (sums of the tile partials in tile order, then the scalar step):
//   innerMatvec  INNER    q = (A32 + s I) p, (p, q)          -> alpha
//   innerUpdate  INNER    d += alpha p, rr -= alpha q, (rr, rr) -> beta, inner stop test
//   innerDir     INNER    p = rr + beta p
//   correct      CORRECT  x += |r| d                           (FP64)
//   outerResid   CORRECT  r = b - (A64 + s I) x, (r, r), (b, b) -> outer stop test (FP64)
//   restart      RESTART  rr = r / |r|, d = 0, p = rr, (rr, rr)
//
// GMRES requests (nonsymmetric-capable, Solvers/Linear/GMRES.hpp:79-180) take a second track:
// each right-hand side is a job in a FIFO that `gmres_workers` host threads, each with its own
// non-blocking HIP stream, take as they free up, so the k right-hand sides of a request run side
// by side.  A job is FP64 refinement over an FP32 GMRES(30) inner solve of (A + shift I) with the
// degree-3 Jacobi polynomial over BF16 values as the inner preconditioner.  Each worker keeps,
// per resident matrix, FP64 / FP32 / BF16 CSR copies of A with the layout of the resident FP64
// CSR copy; a job only rewrites their diagonal entries to a + shift (rounded from the FP64
// value), and every buffer grows and stays, so after the first job on a size nothing is
// allocated or freed (a hipFree waits for the whole device, CG launches included).  The jobs'
// kernels overlap on the device with each other and with the batched CG launches; each result
// depends only on its own request (deterministic reductions), never on the mix.
#include "common.hpp"
#include "matrix.hpp"
#include "refine.hpp"

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace tnlhip {

namespace {

constexpr int kTile = 256;
constexpr int kMaxRhs = 16;
constexpr int kServerGrid = 2048;  // 8 workgroups of 256 per CU
constexpr int kServerXcds = 8;     // workgroups are dispatched round-robin over the 8 XCDs
// The two product kernels (at most 128 VGPRs: 4 workgroups of 256 per CU) launch exactly what is
// resident at once, so the workgroups of an XCD walk its tiles together, front to back, and the x
// rows one tile gathers are still in L2 when the tiles a plane further on gather them again; a
// larger grid leaves every other block of tiles to a second round of workgroups that starts once
// the first has finished its whole walk, when those rows are long gone.
constexpr int kProductGrid = 1024;
static_assert( kServerGrid % kServerXcds == 0 && kProductGrid % kServerXcds == 0,
               "server grids must be multiples of the XCD count" );
constexpr int kRowsPerLane = 16;    // rows per lane in a tile (8 and 16 measured: 16 is 10 % faster
                                    // in the product, whose per-tile round trips it halves)

enum Phase : int32_t { P_IDLE = 0, P_INNER = 1, P_CORRECT = 2, P_RESTART = 3, P_DONE = 4, P_FAILED = 5 };
constexpr int32_t kLive = ( 1 << P_INNER ) | ( 1 << P_CORRECT ) | ( 1 << P_RESTART );

// host-written at admission
struct ReqStatic
{
   int32_t n, k, kp, rpt;        // rows, right-hand sides, padded k, rows per tile
   int32_t tiles, sliceBase, maxOuter, maxInner;
   int64_t vec;                  // entry offset in the vector arena
   int64_t mat;                  // slot offset in the matrix arena
   double shift, tol, innerTol2;
   int32_t stagnation, pad;
};

// device-written
struct ReqDyn
{
   int32_t mask;        // bit p set when some system of the request is in phase p
   uint32_t bits[ 6 ];  // bits[ p ]: bit j set when system j is in phase p
};

struct Sys
{
   int32_t phase, innerIt, outerIt, slow;
   int64_t innerTotal;
   double normb, scale, rho, rho0, alpha, beta, res, prevRes;
};

// The tile walk: walk positions [xcdStart[x], xcdStart[x + 1]) of tileOrder are XCD x's share.
struct Ctl
{
   int32_t numTiles;
   int32_t xcdStart[ kServerXcds + 1 ];
};

struct Arena
{
   const Ctl* ctl;
   const int32_t* tileReq;
   const int32_t* tileFirst;
   const int32_t* tileOrder;  // tile ids in walk order
   const ReqStatic* rs;
   ReqDyn* rd;
   Sys* sys;
   double* part;  // [tile][2][16]
   const int32_t* cols;
   const double* v64;
   const float* v32;
   const int32_t* slices;
   double* a64;   // x, b, r per request: 3 E
   float* a32;    // d, rr, p, q per request: 4 E
   int maxRequests;
};

// A tile: rpt = 256 / kp rows per row group, kRowsPerLane row groups; lane (lr, j) holds
// rows base + q * rpt + lr, right-hand side j.
struct TileCtx
{
   int32_t base, lr, j;
};

__device__ __forceinline__ TileCtx tileCtx( const Arena& A, int t, int req, const ReqStatic& R )
{
   TileCtx c;
   const int ti = t - A.tileFirst[ req ];
   c.lr = threadIdx.x / R.kp;
   c.j = threadIdx.x % R.kp;
   c.base = ti * R.rpt * kRowsPerLane;
   return c;
}

__device__ __forceinline__ bool laneOn( const TileCtx& c, const ReqStatic& R, const ReqDyn& D, int phase )
{
   return c.lr < R.rpt && c.j < R.k && ( ( D.bits[ phase ] >> c.j ) & 1u );
}

// sum[q] = sum_k A[row_q, k] * v[col * K + j] over the SlicedEllpack rows of the lane (padding
// column -1), left to right.  Rows go in pairs, eight slots of both rows per round trip (loads
// past a row's width are masked off), so 16 column, value and x loads are in flight.
constexpr int kSlotChunk = 8;

template< typename V, typename X, typename Acc >
__device__ __forceinline__ void rowProducts( const Arena& A, const ReqStatic& R, const TileCtx& c, const V* vals,
                                             const X* v, Acc ( &sum )[ kRowsPerLane ] )
{
#pragma unroll
   for( int q0 = 0; q0 < kRowsPerLane; q0 += 2 ) {
      int32_t width[ 2 ];
      int64_t idx[ 2 ];
#pragma unroll
      for( int u = 0; u < 2; u++ ) {
         const int32_t row = c.base + ( q0 + u ) * R.rpt + c.lr;
         width[ u ] = 0;
         idx[ u ] = 0;
         if( row < R.n ) {
            const int32_t s = row >> 5;
            const int32_t b0 = A.slices[ R.sliceBase + s ], b1 = A.slices[ R.sliceBase + s + 1 ];
            width[ u ] = ( b1 - b0 ) >> 5;
            if( width[ u ] > 0 )  // else slot 0 of the arena: loaded, never used
               idx[ u ] = R.mat + b0 + ( row & 31 );
         }
      }
      Acc acc[ 2 ] = { 0, 0 };
      const int32_t maxw = max( width[ 0 ], width[ 1 ] );
      const int32_t last[ 2 ] = { max( width[ 0 ] - 1, 0 ), max( width[ 1 ] - 1, 0 ) };
      for( int32_t k0 = 0; k0 < maxw; k0 += kSlotChunk ) {
         int32_t col[ 2 ][ kSlotChunk ];
#pragma unroll
         for( int u = 0; u < 2; u++ )
#pragma unroll
            for( int k = 0; k < kSlotChunk; k++ )
               col[ u ][ k ] = A.cols[ idx[ u ] + 32 * (int64_t) min( k0 + k, last[ u ] ) ];
         Acc term[ 2 ][ kSlotChunk ];
#pragma unroll
         for( int u = 0; u < 2; u++ )
#pragma unroll
            for( int k = 0; k < kSlotChunk; k++ ) {
               // every load is issued (clamped slot; row 0 of x for padding and for slots past
               // the row) so that all of them are in flight together, and the product is formed
               // unconditionally: a select between it and 0 is turned into a branch around the
               // loads, one round trip per slot; the sum below skips the dropped slots
               const bool on = k0 + k < width[ u ] && col[ u ][ k ] >= 0;
               const Acc a = (Acc) vals[ idx[ u ] + 32 * (int64_t) min( k0 + k, last[ u ] ) ];
               const Acc b = (Acc) v[ (int64_t) ( on ? col[ u ][ k ] : 0 ) * R.k + c.j ];
               term[ u ][ k ] = a * b;
               col[ u ][ k ] = on ? 0 : -1;
            }
#pragma unroll
         for( int u = 0; u < 2; u++ )
#pragma unroll
            for( int k = 0; k < kSlotChunk; k++ )
               if( col[ u ][ k ] >= 0 )
                  acc[ u ] += term[ u ][ k ];
      }
      sum[ q0 ] = acc[ 0 ];
      sum[ q0 + 1 ] = acc[ 1 ];
   }
}

// Vector form of a tile for requests whose k is a multiple of V (2 or 4): with L = kp / V lanes
// per row, lane (lr, jq) holds rows base + u * (256 / L) + lr for u < kRowsPerLane / V and the V
// right-hand sides V jq .. V jq + V - 1, which are adjacent in the interleaved vector arena.  Each
// slot then costs one V-wide load of x, and one column and one value load shared by V systems,
// instead of V scalar gathers and V copies of the column and value loads.  The tile covers the
// same rows as the scalar form.
template< typename T, int V >
struct VecOf
{
   typedef T type __attribute__( ( ext_vector_type( V ) ) );
};

struct VecCtx
{
   int32_t base, lr, jq, rpp;
};

template< int V >
__device__ __forceinline__ VecCtx vecCtx( int t, int32_t tileFirst, const ReqStatic& R )
{
   const int L = R.kp / V;
   VecCtx c;
   c.lr = threadIdx.x / L;
   c.jq = threadIdx.x % L;
   c.rpp = kTile / L;
   c.base = ( t - tileFirst ) * R.rpt * kRowsPerLane;
   return c;
}

// sum[u] = sum_k A[row_u, k] * v[col * K + V jq + i] for i < V, slots left to right; rows in
// pairs, kSlotChunk slots of both per round trip, as rowProducts.
template< typename T, int V >
__device__ __forceinline__ void rowProductsVec( const Arena& A, const ReqStatic& R, const VecCtx& c, const T* vals,
                                                const T* v, typename VecOf< T, V >::type ( &sum )[ kRowsPerLane / V ] )
{
   using VT = typename VecOf< T, V >::type;
   // the slice offsets of all the lane's rows in one round trip (rows past n read the last
   // row's slice and get width 0; an empty row reads slot 0 of the arena, never used)
   constexpr int U = kRowsPerLane / V;
   int32_t wAll[ U ];
   int64_t iAll[ U ];
#pragma unroll
   for( int u = 0; u < U; u++ ) {
      const int32_t row = c.base + u * c.rpp + c.lr;
      const int32_t s = min( row, R.n - 1 ) >> 5;
      const int32_t b0 = A.slices[ R.sliceBase + s ], b1 = A.slices[ R.sliceBase + s + 1 ];
      wAll[ u ] = row < R.n ? ( b1 - b0 ) >> 5 : 0;
      iAll[ u ] = wAll[ u ] > 0 ? R.mat + b0 + ( row & 31 ) : 0;
   }
#pragma unroll
   for( int u0 = 0; u0 < U; u0 += 2 ) {
      int32_t width[ 2 ] = { wAll[ u0 ], wAll[ u0 + 1 ] };
      int64_t idx[ 2 ] = { iAll[ u0 ], iAll[ u0 + 1 ] };
      VT acc[ 2 ] = { (VT) 0, (VT) 0 };
      const int32_t maxw = max( width[ 0 ], width[ 1 ] );
      const int32_t last[ 2 ] = { max( width[ 0 ] - 1, 0 ), max( width[ 1 ] - 1, 0 ) };
      const T* vj = v + V * c.jq;
      for( int32_t k0 = 0; k0 < maxw; k0 += kSlotChunk ) {
         int32_t col[ 2 ][ kSlotChunk ];
#pragma unroll
         for( int u = 0; u < 2; u++ )
#pragma unroll
            for( int k = 0; k < kSlotChunk; k++ )
               col[ u ][ k ] = A.cols[ idx[ u ] + 32 * (int64_t) min( k0 + k, last[ u ] ) ];
         VT term[ 2 ][ kSlotChunk ];
#pragma unroll
         for( int u = 0; u < 2; u++ )
#pragma unroll
            for( int k = 0; k < kSlotChunk; k++ ) {
               const bool on = k0 + k < width[ u ] && col[ u ][ k ] >= 0;
               const T a = vals[ idx[ u ] + 32 * (int64_t) min( k0 + k, last[ u ] ) ];
               const VT b = *reinterpret_cast< const VT* >( vj + (int64_t) ( on ? col[ u ][ k ] : 0 ) * R.k );
               term[ u ][ k ] = a * b;  // unconditional, as in rowProducts
               col[ u ][ k ] = on ? 0 : -1;
            }
#pragma unroll
         for( int u = 0; u < 2; u++ )
#pragma unroll
            for( int k = 0; k < kSlotChunk; k++ )
               if( col[ u ][ k ] >= 0 )
                  acc[ u ] += term[ u ][ k ];
      }
      sum[ u0 ] = acc[ 0 ];
      sum[ u0 + 1 ] = acc[ 1 ];
   }
}

// Tile partials of the vector form: lane sums over the lanes of one jq (xor offsets down to L),
// then the four waves in order; part[t][0][j] for j < kp.
template< int V >
__device__ __forceinline__ void tilePartialsVec( double* part, int t, int kp, const double ( &v )[ V ] )
{
   __shared__ double lds[ kTile / 64 ][ kMaxRhs ];
   const int L = kp / V;
   const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
#pragma unroll
   for( int i = 0; i < V; i++ ) {
      double x = v[ i ];
      for( int o = 32; o >= L; o >>= 1 )
         x += __shfl_xor( x, o );
      if( lane < L )
         lds[ wave ][ V * lane + i ] = x;
   }
   __syncthreads();
   if( (int) threadIdx.x < kp ) {
      const int j = threadIdx.x;
      double s = lds[ 0 ][ j ];
      for( int q = 1; q < kTile / 64; q++ )
         s += lds[ q ][ j ];
      part[ (int64_t) t * 2 * kMaxRhs + j ] = s;
   }
}

// Walks the tile table, fetching the next tile and its request one step ahead; f( t, req ).
// Workgroup b runs on XCD b % 8 (round-robin dispatch) and walks that XCD's share of the walk
// order, which holds a contiguous eighth of every request's tiles (rebuildTiles): the rows a
// tile's stencil reaches (the neighbouring tiles, a plane away) are then taken on the same XCD,
// whose L2 keeps that x, and every request still spreads over all eight XCDs.
template< typename F >
__device__ __forceinline__ void forTiles( const Arena& A, F&& f )
{
   const int xcd = blockIdx.x % kServerXcds;
   const int per = gridDim.x / kServerXcds;
   const int k1 = A.ctl->xcdStart[ xcd + 1 ];
   int k = A.ctl->xcdStart[ xcd ] + (int) ( blockIdx.x / kServerXcds );
   int t = k < k1 ? A.tileOrder[ k ] : 0;
   int req = k < k1 ? A.tileReq[ t ] : 0;
   while( k < k1 ) {
      const int kn = k + per;
      const int tn = kn < k1 ? A.tileOrder[ kn ] : 0;
      const int rn = kn < k1 ? A.tileReq[ tn ] : 0;
      f( t, req );
      k = kn;
      t = tn;
      req = rn;
   }
}

// Per-tile sums of W values per right-hand side into part[t][w][j], in a fixed order.
template< int W >
__device__ __forceinline__ void tilePartials( double* part, int t, int kp, const double ( &v )[ W ] )
{
   __shared__ double lds[ kTile / 64 ][ W ][ kMaxRhs ];
   const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
#pragma unroll
   for( int w = 0; w < W; w++ ) {
      double x = v[ w ];
      for( int o = 32; o >= kp; o >>= 1 )
         x += __shfl_xor( x, o );
      if( lane < kp )
         lds[ wave ][ w ][ lane ] = x;
   }
   __syncthreads();
   if( threadIdx.x < W * kp ) {
      const int w = threadIdx.x / kp, j = threadIdx.x % kp;
      double s = lds[ 0 ][ w ][ j ];
      for( int q = 1; q < kTile / 64; q++ )
         s += lds[ q ][ w ][ j ];
      part[ (int64_t) t * 2 * kMaxRhs + w * kMaxRhs + j ] = s;
   }
}

// Totals over the request's tiles in a fixed order (each of the 256 lanes sums a strided
// share of one (w, j) column, then a fixed tree); returns LDS holding tot[w * kMaxRhs + j],
// valid in every thread on return.
template< int W >
__device__ __forceinline__ const double* requestTotals( const double* part, int first, int tiles, int kp )
{
   __shared__ double red[ kTile ];
   __shared__ double out[ 2 * kMaxRhs ];
   const int pairs = W * kp;  // a power of two <= 32
   const int G = kTile / pairs;
   const int tid = threadIdx.x;
   const int pair = tid % pairs, g = tid / pairs;
   const int w = pair / kp, j = pair % kp;
   const double* p = part + (int64_t) first * 2 * kMaxRhs + w * kMaxRhs + j;
   // kStreams independent loads in flight per lane (the partials come from memory: a request of
   // 8k tiles is 30 dependent trips per lane instead of 120), folded in a fixed tree
   constexpr int kStreams = 16;
   double a[ kStreams ];
#pragma unroll
   for( int u = 0; u < kStreams; u++ )
      a[ u ] = 0;
   int ti = g;
   for( ; ti + ( kStreams - 1 ) * G < tiles; ti += kStreams * G ) {
#pragma unroll
      for( int u = 0; u < kStreams; u++ )
         a[ u ] += p[ (int64_t) ( ti + u * G ) * 2 * kMaxRhs ];
   }
   for( ; ti < tiles; ti += G )
      a[ 0 ] += p[ (int64_t) ti * 2 * kMaxRhs ];
#pragma unroll
   for( int h = kStreams / 2; h >= 1; h /= 2 ) {
#pragma unroll
      for( int u = 0; u < h; u++ )
         a[ u ] += a[ u + h ];
   }
   red[ tid ] = a[ 0 ];
   __syncthreads();
   for( int step = G / 2; step >= 1; step >>= 1 ) {
      if( g < step )
         red[ tid ] += red[ tid + step * pairs ];
      __syncthreads();
   }
   if( tid < pairs )
      out[ w * kMaxRhs + j ] = red[ tid ];
   __syncthreads();
   return out;
}

// Per-request finish of a reduction kernel: one workgroup per resident request whose mask
// holds `phase` sums its tiles' partials, lanes j < k update system j, and the request's phase
// mask and bits are recomputed.  Step( Sys&, const double* tot, int j, const ReqStatic& ).
template< int W, typename Step >
__global__ __launch_bounds__( kTile ) void requestFinalKernel( Arena A, int phase, Step step )
{
   __shared__ int m;
   __shared__ uint32_t bits[ 6 ];
   for( int req = blockIdx.x; req < A.maxRequests; req += gridDim.x ) {
      if( ! ( A.rd[ req ].mask & ( 1 << phase ) ) )
         continue;
      const ReqStatic R = A.rs[ req ];
      const double* tot = requestTotals< W >( A.part, A.tileFirst[ req ], R.tiles, R.kp );
      if( threadIdx.x == 0 )
         m = 0;
      if( threadIdx.x < 6 )
         bits[ threadIdx.x ] = 0;
      if( threadIdx.x < R.k ) {
         Sys& s = A.sys[ req * kMaxRhs + threadIdx.x ];
         step( s, tot, threadIdx.x, R );
      }
      __syncthreads();
      if( threadIdx.x < R.k ) {
         const int ph = A.sys[ req * kMaxRhs + threadIdx.x ].phase;
         atomicOr( &m, 1 << ph );
         atomicOr( &bits[ ph ], 1u << threadIdx.x );
      }
      __syncthreads();
      if( threadIdx.x == 0 )
         A.rd[ req ].mask = m;
      if( threadIdx.x < 6 )
         A.rd[ req ].bits[ threadIdx.x ] = bits[ threadIdx.x ];
      __syncthreads();
   }
}

__device__ __forceinline__ float* a32( const Arena& A, const ReqStatic& R, int which )
{
   return A.a32 + 4 * R.vec + (int64_t) which * R.n * R.k;
}
__device__ __forceinline__ double* a64( const Arena& A, const ReqStatic& R, int which )
{
   return A.a64 + 3 * R.vec + (int64_t) which * R.n * R.k;
}
enum { X64 = 0, B64 = 1, R64 = 2 };
enum { D32 = 0, RR32 = 1, P32 = 2, Q32 = 3 };

// entry index of row group q of the lane, or -1 past the last row
__device__ __forceinline__ int64_t entryOf( const TileCtx& c, const ReqStatic& R, int q )
{
   const int32_t row = c.base + q * R.rpt + c.lr;
   return row < R.n ? (int64_t) row * R.k + c.j : -1;
}

// ---------------------------------------------------------------- tile kernels

// q = (A32 + s I) p and the (p, q) partials over the vector form of the tile (k % V == 0)
template< int V >
__device__ __forceinline__ void innerMatvecVec( const Arena& A, const ReqStatic& R, const ReqDyn& D, int t,
                                                int32_t tileFirst )
{
   using VT = typename VecOf< float, V >::type;
   constexpr uint32_t all = ( 1u << V ) - 1u;
   const VecCtx c = vecCtx< V >( t, tileFirst, R );
   const uint32_t act = V * c.jq < R.k ? ( D.bits[ P_INNER ] >> ( V * c.jq ) ) & all : 0u;
   double v[ V ] = {};
   if( act ) {
      const float* p = a32( A, R, P32 );
      float* qv = a32( A, R, Q32 );
      VT sum[ kRowsPerLane / V ];
      rowProductsVec< float, V >( A, R, c, A.v32, p, sum );
#pragma unroll
      for( int u = 0; u < kRowsPerLane / V; u++ ) {
         const int32_t row = c.base + u * c.rpp + c.lr;
         if( row < R.n ) {
            const int64_t e = (int64_t) row * R.k + V * c.jq;
            const VT pe = *reinterpret_cast< const VT* >( p + e );
            const VT r = sum[ u ] + (float) R.shift * pe;
            if( act == all )
               *reinterpret_cast< VT* >( qv + e ) = r;
            else {
#pragma unroll
               for( int i = 0; i < V; i++ )
                  if( ( act >> i ) & 1u )
                     qv[ e + i ] = r[ i ];
            }
#pragma unroll
            for( int i = 0; i < V; i++ )
               if( ( act >> i ) & 1u )
                  v[ i ] += (double) pe[ i ] * (double) r[ i ];
         }
      }
   }
   tilePartialsVec< V >( A.part, t, R.kp, v );
   __syncthreads();
}

// at most 128 VGPRs (4 waves per SIMD, no spills): the unbounded build took 143 (3 waves)
__global__ __launch_bounds__( kTile, 4 ) void innerMatvecKernel( Arena A )
{
   forTiles( A, [ & ]( int t, int req ) {
      // the request's state, constants and first tile in one round trip, before the phase test
      const ReqDyn D = A.rd[ req ];
      const ReqStatic R = A.rs[ req ];
      const int32_t tileFirst = A.tileFirst[ req ];
      if( ! ( D.mask & ( 1 << P_INNER ) ) )
         return;
      if( R.k % 4 == 0 ) {
         innerMatvecVec< 4 >( A, R, D, t, tileFirst );
         return;
      }
      if( R.k % 2 == 0 ) {
         innerMatvecVec< 2 >( A, R, D, t, tileFirst );
         return;
      }
      const TileCtx c = tileCtx( A, t, req, R );
      double v[ 1 ] = { 0 };
      if( laneOn( c, R, D, P_INNER ) ) {
         const float* p = a32( A, R, P32 );
         float* qv = a32( A, R, Q32 );
         float sum[ kRowsPerLane ];
         rowProducts< float, float, float >( A, R, c, A.v32, p, sum );
#pragma unroll
         for( int q = 0; q < kRowsPerLane; q++ ) {
            const int64_t e = entryOf( c, R, q );
            if( e >= 0 ) {
               const float pe = p[ e ];
               const float r = sum[ q ] + (float) R.shift * pe;
               qv[ e ] = r;
               v[ 0 ] += (double) pe * (double) r;
            }
         }
      }
      tilePartials< 1 >( A.part, t, R.kp, v );
      __syncthreads();
   } );
}

struct StepAlpha
{
   __device__ void operator()( Sys& s, const double* tot, int j, const ReqStatic& ) const
   {
      if( s.phase != P_INNER )
         return;
      const double pq = tot[ j ];
      if( ! ( pq > 0 ) || ! isfinite( pq ) )
         s.phase = P_CORRECT;  // breakdown: keep the current correction
      else
         s.alpha = s.rho / pq;
   }
};

__global__ __launch_bounds__( kTile ) void innerUpdateKernel( Arena A )
{
   forTiles( A, [ & ]( int t, int req ) {
      const ReqDyn D = A.rd[ req ];
      if( ! ( D.mask & ( 1 << P_INNER ) ) )
         return;
      const ReqStatic R = A.rs[ req ];
      const TileCtx c = tileCtx( A, t, req, R );
      double v[ 1 ] = { 0 };
      if( laneOn( c, R, D, P_INNER ) ) {
         const float alpha = (float) A.sys[ req * kMaxRhs + c.j ].alpha;
         float* d = a32( A, R, D32 );
         float* rr = a32( A, R, RR32 );
         const float* p = a32( A, R, P32 );
         const float* qv = a32( A, R, Q32 );
#pragma unroll
         for( int q = 0; q < kRowsPerLane; q++ ) {
            const int64_t e = entryOf( c, R, q );
            if( e >= 0 ) {
               d[ e ] += alpha * p[ e ];
               const float r = rr[ e ] - alpha * qv[ e ];
               rr[ e ] = r;
               v[ 0 ] += (double) r * (double) r;
            }
         }
      }
      tilePartials< 1 >( A.part, t, R.kp, v );
      __syncthreads();
   } );
}

struct StepBeta
{
   __device__ void operator()( Sys& s, const double* tot, int j, const ReqStatic& R ) const
   {
      if( s.phase != P_INNER )
         return;
      const double rho = tot[ j ];
      s.innerIt++;
      s.innerTotal++;
      if( ! isfinite( rho ) || rho <= R.innerTol2 * s.rho0 || s.innerIt >= R.maxInner )
         s.phase = P_CORRECT;
      else {
         s.beta = rho / s.rho;
         s.rho = rho;
      }
   }
};

__global__ __launch_bounds__( kTile ) void innerDirKernel( Arena A )
{
   forTiles( A, [ & ]( int t, int req ) {
      const ReqDyn D = A.rd[ req ];
      if( ! ( D.mask & ( 1 << P_INNER ) ) )
         return;
      const ReqStatic R = A.rs[ req ];
      const TileCtx c = tileCtx( A, t, req, R );
      if( laneOn( c, R, D, P_INNER ) ) {
         const float beta = (float) A.sys[ req * kMaxRhs + c.j ].beta;
         float* p = a32( A, R, P32 );
         const float* rr = a32( A, R, RR32 );
#pragma unroll
         for( int q = 0; q < kRowsPerLane; q++ ) {
            const int64_t e = entryOf( c, R, q );
            if( e >= 0 )
               p[ e ] = rr[ e ] + beta * p[ e ];
         }
      }
   } );
}

__global__ __launch_bounds__( kTile ) void correctKernel( Arena A )
{
   forTiles( A, [ & ]( int t, int req ) {
      const ReqDyn D = A.rd[ req ];
      if( ! ( D.mask & ( 1 << P_CORRECT ) ) )
         return;
      const ReqStatic R = A.rs[ req ];
      const TileCtx c = tileCtx( A, t, req, R );
      if( laneOn( c, R, D, P_CORRECT ) ) {
         const double scale = A.sys[ req * kMaxRhs + c.j ].scale;
         double* x = a64( A, R, X64 );
         const float* d = a32( A, R, D32 );
#pragma unroll
         for( int q = 0; q < kRowsPerLane; q++ ) {
            const int64_t e = entryOf( c, R, q );
            if( e >= 0 )
               x[ e ] += scale * (double) d[ e ];
         }
      }
   } );
}

// at most 128 VGPRs, as innerMatvecKernel
__global__ __launch_bounds__( kTile, 4 ) void outerResidKernel( Arena A )
{
   forTiles( A, [ & ]( int t, int req ) {
      const ReqDyn D = A.rd[ req ];
      if( ! ( D.mask & ( 1 << P_CORRECT ) ) )
         return;
      const ReqStatic R = A.rs[ req ];
      const TileCtx c = tileCtx( A, t, req, R );
      double v[ 2 ] = { 0, 0 };
      if( laneOn( c, R, D, P_CORRECT ) ) {
         const double* x = a64( A, R, X64 );
         const double* b = a64( A, R, B64 );
         double* r = a64( A, R, R64 );
         double sum[ kRowsPerLane ];
         rowProducts< double, double, double >( A, R, c, A.v64, x, sum );
#pragma unroll
         for( int q = 0; q < kRowsPerLane; q++ ) {
            const int64_t e = entryOf( c, R, q );
            if( e >= 0 ) {
               const double bb = b[ e ];
               const double re = bb - ( sum[ q ] + R.shift * x[ e ] );
               r[ e ] = re;
               v[ 0 ] += re * re;
               v[ 1 ] += bb * bb;
            }
         }
      }
      tilePartials< 2 >( A.part, t, R.kp, v );
      __syncthreads();
   } );
}

struct StepOuter
{
   __device__ void operator()( Sys& s, const double* tot, int j, const ReqStatic& R ) const
   {
      if( s.phase != P_CORRECT )
         return;
      if( s.normb < 0 ) {
         s.normb = sqrt( tot[ kMaxRhs + j ] );
         if( s.normb == 0 )
            s.normb = 1;
      }
      const double nr = sqrt( tot[ j ] );
      const double res = nr / s.normb;
      s.slow = ( s.outerIt > 0 && res > 0.5 * s.prevRes ) ? s.slow + 1 : 0;
      s.res = res;
      if( res < R.tol || nr == 0 )
         s.phase = P_DONE;
      else if( ! isfinite( res ) || s.outerIt >= R.maxOuter || ( R.stagnation > 0 && s.slow >= R.stagnation ) )
         s.phase = P_FAILED;
      else {
         s.phase = P_RESTART;
         s.scale = nr;
         s.prevRes = res;
      }
   }
};

__global__ __launch_bounds__( kTile ) void restartKernel( Arena A )
{
   forTiles( A, [ & ]( int t, int req ) {
      const ReqDyn D = A.rd[ req ];
      if( ! ( D.mask & ( 1 << P_RESTART ) ) )
         return;
      const ReqStatic R = A.rs[ req ];
      const TileCtx c = tileCtx( A, t, req, R );
      double v[ 1 ] = { 0 };
      if( laneOn( c, R, D, P_RESTART ) ) {
         const double scale = A.sys[ req * kMaxRhs + c.j ].scale;
         const double* r = a64( A, R, R64 );
         float* rr = a32( A, R, RR32 );
         float* p = a32( A, R, P32 );
         float* d = a32( A, R, D32 );
#pragma unroll
         for( int q = 0; q < kRowsPerLane; q++ ) {
            const int64_t e = entryOf( c, R, q );
            if( e >= 0 ) {
               const float x = (float) ( r[ e ] / scale );
               rr[ e ] = x;
               p[ e ] = x;
               d[ e ] = 0.0f;
               v[ 0 ] += (double) x * (double) x;
            }
         }
      }
      tilePartials< 1 >( A.part, t, R.kp, v );
      __syncthreads();
   } );
}

struct StepRestart
{
   __device__ void operator()( Sys& s, const double* tot, int j, const ReqStatic& ) const
   {
      if( s.phase != P_RESTART )
         return;
      s.rho = s.rho0 = tot[ j ];
      s.innerIt = 0;
      s.outerIt++;
      s.phase = P_INNER;
   }
};

// admission: x = 0, d = 0 over the request's entries
__global__ __launch_bounds__( 256 ) void admitKernel( int64_t E, double* x, float* d )
{
   const int64_t stride = (int64_t) gridDim.x * blockDim.x;
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < E; i += stride ) {
      x[ i ] = 0.0;
      d[ i ] = 0.0f;
   }
}

// CSR -> SlicedEllpack scatter of one matrix into the arena, one lane per row
__global__ __launch_bounds__( 256 ) void scatterKernel( int32_t n, const int32_t* rp, const int32_t* ci, const double* vv,
                                                        const int32_t* slices, int64_t mat, int32_t* cols, double* v64,
                                                        float* v32 )
{
   const int32_t row = blockIdx.x * blockDim.x + threadIdx.x;
   if( row >= n )
      return;
   const int32_t s = row >> 5;
   const int32_t width = ( slices[ s + 1 ] - slices[ s ] ) >> 5;
   const int64_t base = mat + slices[ s ] + ( row & 31 );
   const int32_t b = rp[ row ], len = rp[ row + 1 ] - b;
   for( int32_t kk = 0; kk < width; kk++ ) {
      const int64_t idx = base + 32 * (int64_t) kk;
      if( kk < len ) {
         cols[ idx ] = ci[ b + kk ];
         v64[ idx ] = vv[ b + kk ];
         v32[ idx ] = (float) vv[ b + kk ];
      }
      else {
         cols[ idx ] = -1;
         v64[ idx ] = 0.0;
         v32[ idx ] = 0.0f;
      }
   }
}

// ---------------------------------------------------------------- host side

// first-fit allocator over [0, capacity)
struct RangeAlloc
{
   int64_t capacity = 0;
   std::map< int64_t, int64_t > freeList;  // offset -> length
   void init( int64_t cap )
   {
      capacity = cap;
      freeList.clear();
      if( cap > 0 )
         freeList[ 0 ] = cap;
   }
   int64_t alloc( int64_t len )
   {
      if( len <= 0 )
         len = 1;
      for( auto it = freeList.begin(); it != freeList.end(); ++it )
         if( it->second >= len ) {
            const int64_t off = it->first, rest = it->second - len;
            freeList.erase( it );
            if( rest > 0 )
               freeList[ off + len ] = rest;
            return off;
         }
      return -1;
   }
   void release( int64_t off, int64_t len )
   {
      if( len <= 0 )
         len = 1;
      auto it = freeList.emplace( off, len ).first;
      auto next = std::next( it );
      if( next != freeList.end() && it->first + it->second == next->first ) {
         it->second += next->second;
         freeList.erase( next );
      }
      if( it != freeList.begin() ) {
         auto prev = std::prev( it );
         if( prev->first + prev->second == it->first ) {
            prev->second += it->second;
            freeList.erase( it );
         }
      }
   }
   int64_t largest() const
   {
      int64_t m = 0;
      for( auto& f : freeList )
         m = std::max( m, f.second );
      return m;
   }
   int64_t freeTotal() const
   {
      int64_t m = 0;
      for( auto& f : freeList )
         m += f.second;
      return m;
   }
};

struct ResidentMatrix
{
   int32_t n = 0;
   int64_t mat = -1, slots = 0;
   int32_t sliceBase = -1, sliceCount = 0;
   int users = 0;
   bool removing = false;
   // GMRES track: the host CSR (kept from add_matrix) and, built on the first GMRES request, the
   // FP64 SlicedEllpack matrix each request's shifted copy is made from
   std::shared_ptr< const std::vector< int32_t > > rowPointers, columns;
   std::shared_ptr< const std::vector< double > > values;
   std::shared_ptr< tnlhip_matrix_s > base;
   uint64_t generation = 0;  // distinct for every add_matrix: keys the workers' cached operators
};

struct Request
{
   int64_t ticket;
   int32_t matrix, k;
   double shift;
   const double* b;
   double* x;
   tnlhip_request_params params;
   int slot = -1;
   double submitted = 0;
};

// val[k] = (V)( base[k] + shift ) on every stored diagonal entry of a CSR matrix laid out as the
// FP64 CSR `base` (the shifted operator of a request; off-diagonal entries are left as they are)
template< typename V >
__device__ __forceinline__ V roundTo( double v );
template<>
__device__ __forceinline__ double roundTo< double >( double v ) { return v; }
template<>
__device__ __forceinline__ float roundTo< float >( double v ) { return (float) v; }
template<>
__device__ __forceinline__ bf16 roundTo< bf16 >( double v ) { return __float2bfloat16( (float) v ); }

template< typename V >
__global__ __launch_bounds__( 256 ) void setShiftedDiagonalKernel( int32_t n, const int32_t* __restrict__ off,
                                                                   const int32_t* __restrict__ col,
                                                                   const double* __restrict__ base,
                                                                   V* __restrict__ val, double shift )
{
   const int64_t stride = (int64_t) gridDim.x * blockDim.x;
   for( int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride )
      for( int32_t k = off[ r ]; k < off[ r + 1 ]; k++ )
         if( col[ k ] == r )
            val[ k ] = roundTo< V >( base[ k ] + shift );
}

}  // namespace

}  // namespace tnlhip

using namespace tnlhip;

// Owning handles of the GMRES track (the C ABI's destroy as the deleter).
struct MatrixHandle
{
   tnlhip_matrix m = nullptr;
   MatrixHandle() = default;
   MatrixHandle( const MatrixHandle& ) = delete;
   MatrixHandle& operator=( const MatrixHandle& ) = delete;
   ~MatrixHandle()
   {
      if( m )
         tnlhip_matrix_destroy( m );
   }
};

// A GMRES request on the worker track: its right-hand sides are separate jobs.
struct GmresTask
{
   Request q;
   int remaining = 0;     // jobs not finished
   bool started = false;  // a job was taken
   tnlhip_request_result out{};
};

struct GmresJob
{
   std::shared_ptr< GmresTask > task;
   int j = 0;  // right-hand side
};

// One worker's copies of a resident matrix: FP64, FP32 and BF16 CSR with the resident FP64 CSR
// copy's layout, their diagonal set for `shift`.
struct GmresOperators
{
   MatrixHandle A64, A32, A16;
   double shift = std::nan( "" );
};

// Everything a GMRES worker keeps between jobs.
struct GmresWorkerState
{
   RefineContext ctx;
   DeviceArray< double > db, dx;
   std::vector< double > hb, hx;
   std::map< uint64_t, std::unique_ptr< GmresOperators > > ops;  // by resident-matrix generation
};

struct tnlhip_server_s
{
   tnlhip_server_config cfg{};
   hipStream_t stream = nullptr;
   DeviceArray< int32_t > cols, slices, tileReq, tileFirst, tileOrder;
   DeviceArray< double > v64, a64, part;
   DeviceArray< float > v32, a32;
   DeviceArray< ReqStatic > rs;
   DeviceArray< ReqDyn > rd;
   DeviceArray< Sys > sys;
   DeviceArray< Ctl > ctl;
   DeviceArray< int32_t > staging;  // CSR upload staging (row pointers + columns), int32
   DeviceArray< double > stagingV;
   RangeAlloc matAlloc, sliceAlloc, vecAlloc;
   std::vector< ResidentMatrix > matrices;
   std::deque< Request > queue;
   std::vector< Request > slots;  // per resident slot (slot < 0: free)
   std::vector< bool > slotUsed;
   std::map< int64_t, tnlhip_request_result > finished;
   int64_t nextTicket = 1;
   int64_t maxTiles = 0;
   bool tilesDirty = false;
   // pinned mirrors
   int32_t* hTileReq = nullptr;
   int32_t* hTileFirst = nullptr;
   int32_t* hTileOrder = nullptr;
   ReqDyn* hDyn = nullptr;
   ReqStatic* hStatic = nullptr;
   Sys* hSys = nullptr;
   Ctl* hCtl = nullptr;
   hipGraph_t graph = nullptr;
   hipGraphExec_t exec = nullptr;
   int64_t launches = 0;
   // GMRES track: one job per right-hand side (under gmresMutex)
   int device = 0;
   std::vector< std::thread > workers;
   std::mutex gmresMutex;
   std::condition_variable gmresWake;
   std::deque< GmresJob > gmresQueue;
   std::vector< std::pair< Request, tnlhip_request_result > > gmresDone;
   int gmresBusy = 0;          // requests with a job taken and not all finished
   int gmresWaiting = 0;       // requests with no job taken yet
   uint64_t nextGeneration = 1;
   bool stopping = false;

   void gmresWorker();
   tnlhip_request_result solveGmresRhs( const Request& q, int j, hipStream_t st, GmresWorkerState& w );

   Arena arena() const
   {
      return Arena{ ctl.ptr, tileReq.ptr, tileFirst.ptr, tileOrder.ptr, rs.ptr, rd.ptr, sys.ptr, part.ptr, cols.ptr,
                    v64.ptr, v32.ptr, slices.ptr, a64.ptr, a32.ptr, cfg.max_requests };
   }

   void enqueueIterations( int iterations )
   {
      const Arena A = arena();
      const int fg = std::min( cfg.max_requests, kServerGrid );
      for( int i = 0; i < iterations; i++ ) {
         innerMatvecKernel<<< kProductGrid, kTile, 0, stream >>>( A );
         requestFinalKernel< 1 ><<< fg, kTile, 0, stream >>>( A, P_INNER, StepAlpha{} );
         innerUpdateKernel<<< kServerGrid, kTile, 0, stream >>>( A );
         requestFinalKernel< 1 ><<< fg, kTile, 0, stream >>>( A, P_INNER, StepBeta{} );
         innerDirKernel<<< kServerGrid, kTile, 0, stream >>>( A );
         correctKernel<<< kServerGrid, kTile, 0, stream >>>( A );
         outerResidKernel<<< kProductGrid, kTile, 0, stream >>>( A );
         requestFinalKernel< 2 ><<< fg, kTile, 0, stream >>>( A, P_CORRECT, StepOuter{} );
         restartKernel<<< kServerGrid, kTile, 0, stream >>>( A );
         requestFinalKernel< 1 ><<< fg, kTile, 0, stream >>>( A, P_RESTART, StepRestart{} );
      }
   }

   ~tnlhip_server_s()
   {
      {
         std::lock_guard< std::mutex > lock( gmresMutex );
         stopping = true;
      }
      gmresWake.notify_all();
      for( std::thread& t : workers )
         t.join();
      if( exec )
         (void) hipGraphExecDestroy( exec );
      if( graph )
         (void) hipGraphDestroy( graph );
      if( stream )
         (void) hipStreamSynchronize( stream );
      for( void* p : { (void*) hTileReq, (void*) hTileFirst, (void*) hTileOrder, (void*) hDyn, (void*) hStatic,
                       (void*) hSys, (void*) hCtl } )
         if( p )
            (void) hipHostFree( p );
      if( stream )
         (void) hipStreamDestroy( stream );
   }
};

// One GMRES worker: its own non-blocking stream, jobs (right-hand sides) in FIFO order until the
// server stops.
void tnlhip_server_s::gmresWorker()
{
   (void) hipSetDevice( device );
   hipStream_t st = nullptr;
   if( hipStreamCreateWithFlags( &st, hipStreamNonBlocking ) != hipSuccess )
      st = nullptr;
   {
      GmresWorkerState w;
      for( ;; ) {
         GmresJob job;
         {
            std::unique_lock< std::mutex > lock( gmresMutex );
            gmresWake.wait( lock, [ & ] { return stopping || ! gmresQueue.empty(); } );
            if( gmresQueue.empty() )
               break;
            job = gmresQueue.front();
            gmresQueue.pop_front();
            if( ! job.task->started ) {
               job.task->started = true;
               gmresWaiting--;
               gmresBusy++;
            }
         }
         const tnlhip_request_result r = solveGmresRhs( job.task->q, job.j, st, w );
         {
            std::lock_guard< std::mutex > lock( gmresMutex );
            GmresTask& t = *job.task;
            t.out.converged = t.out.converged && r.converged;
            t.out.outer_iterations = std::max( t.out.outer_iterations, r.outer_iterations );
            t.out.inner_iterations += r.inner_iterations;
            t.out.residue = std::isnan( r.residue ) || std::isnan( t.out.residue ) ? std::nan( "" )
                                                                                   : std::max( t.out.residue, r.residue );
            if( --t.remaining == 0 ) {
               gmresDone.emplace_back( t.q, t.out );
               gmresBusy--;
            }
         }
      }
   }
   if( st )
      (void) hipStreamDestroy( st );
}

// (A + shift I) x_j = b_j for right-hand side j: FP64 refinement over an FP32 GMRES(30) inner solve
// (Jacobi polynomial over BF16 values); b and x are the request's interleaved host arrays (entry
// row * k + j).
tnlhip_request_result tnlhip_server_s::solveGmresRhs( const Request& q, int j, hipStream_t st, GmresWorkerState& w )
{
   tnlhip_request_result out{};
   out.converged = 1;
   const int status = guarded( [ & ] {
      std::shared_ptr< tnlhip_matrix_s > base;
      int32_t n = 0;
      uint64_t generation = 0;
      std::shared_ptr< const std::vector< int32_t > > rp, ci;
      std::shared_ptr< const std::vector< double > > va;
      {
         std::lock_guard< std::mutex > lock( gmresMutex );
         ResidentMatrix& M = matrices[ q.matrix ];
         n = M.n;
         rp = M.rowPointers;
         ci = M.columns;
         va = M.values;
         base = M.base;
         generation = M.generation;
      }
      if( ! base ) {
         // built outside the lock (an upload); a second worker may build it too, one copy is kept
         tnlhip_matrix h = nullptr;
         TNLHIP_CHECK_STATUS( tnlhip_matrix_create_csr( TNLHIP_CSR_SCALAR, TNLHIP_F64, n, n, rp->data(), ci->data(),
                                                        va->data(), 1, &h ) );
         std::shared_ptr< tnlhip_matrix_s > built( h, []( tnlhip_matrix m ) { tnlhip_matrix_destroy( m ); } );
         std::lock_guard< std::mutex > lock( gmresMutex );
         ResidentMatrix& M = matrices[ q.matrix ];
         if( ! M.base )
            M.base = built;
         base = M.base;
      }
      const Matrix& B = *reinterpret_cast< const Matrix* >( base.get() );
      auto it = w.ops.find( generation );
      if( it == w.ops.end() ) {
         // this worker's first job on the matrix: drop copies of matrices no longer resident, then
         // copy A in the resident CSR layout (a conversion keeps the source's entry order and its
         // stored entries per row)
         {
            std::lock_guard< std::mutex > lock( gmresMutex );
            for( auto o = w.ops.begin(); o != w.ops.end(); ) {
               bool live = false;
               for( const ResidentMatrix& M : matrices )
                  live = live || ( M.generation == o->first && M.mat >= 0 );
               o = live ? std::next( o ) : w.ops.erase( o );
            }
         }
         auto ops = std::make_unique< GmresOperators >();
         TNLHIP_CHECK_STATUS( tnlhip_matrix_convert( base.get(), TNLHIP_CSR_SCALAR, TNLHIP_F64, &ops->A64.m ) );
         TNLHIP_CHECK_STATUS( tnlhip_matrix_convert( base.get(), TNLHIP_CSR_SCALAR, TNLHIP_F32, &ops->A32.m ) );
         TNLHIP_CHECK_STATUS( tnlhip_matrix_convert( base.get(), TNLHIP_CSR_SCALAR, TNLHIP_BF16, &ops->A16.m ) );
         it = w.ops.emplace( generation, std::move( ops ) ).first;
      }
      GmresOperators& o = *it->second;
      if( std::memcmp( &o.shift, &q.shift, sizeof( double ) ) != 0 ) {
         // the request's operator: a + shift on the diagonal, rounded to each copy's value type
         auto setDiag = [ & ]( tnlhip_matrix h, auto* tag ) {
            using V = std::remove_pointer_t< decltype( tag ) >;
            Matrix& C = *reinterpret_cast< Matrix* >( h );
            C.diagonalValid = false;
            if( n > 0 )
               setShiftedDiagonalKernel< V ><<< gridFor( n, 256 ), 256, 0, st >>>(
                  n, B.offsets.ptr, B.columns.ptr, reinterpret_cast< const double* >( B.values.ptr ),
                  reinterpret_cast< V* >( C.values.ptr ), q.shift );
         };
         setDiag( o.A64.m, (double*) nullptr );
         setDiag( o.A32.m, (float*) nullptr );
         setDiag( o.A16.m, (bf16*) nullptr );
         TNLHIP_CHECK( hipGetLastError() );
         o.shift = q.shift;
      }
      tnlhip_refinement_params rp_{};
      tnlhip_refinement_params_default( &rp_ );
      rp_.inner.preconditioner = 3;
      rp_.inner.poly_degree = 3;
      rp_.poly_matrix = o.A16.m;
      rp_.tolerance = q.params.tolerance;
      rp_.max_outer_iterations = q.params.max_outer_iterations;
      rp_.stagnation_limit = q.params.stagnation_limit;
      rp_.inner.convergence_residue = q.params.inner_tolerance;
      rp_.inner.max_iterations = q.params.max_inner_iterations;
      growTo( w.db, std::max( n, 1 ) );
      growTo( w.dx, std::max( n, 1 ) );
      w.hb.resize( n );
      w.hx.resize( n );
      for( int32_t i = 0; i < n; i++ )
         w.hb[ i ] = q.b[ (int64_t) i * q.k + j ];
      TNLHIP_CHECK( hipMemcpyAsync( w.db.ptr, w.hb.data(), sizeof( double ) * n, hipMemcpyHostToDevice, st ) );
      TNLHIP_CHECK( hipMemsetAsync( w.dx.ptr, 0, sizeof( double ) * n, st ) );
      tnlhip_refinement_result res{};
      refineWith( w.ctx, o.A64.m, o.A32.m, &rp_, w.db.ptr, w.dx.ptr, &res, st );
      TNLHIP_CHECK( hipMemcpyAsync( w.hx.data(), w.dx.ptr, sizeof( double ) * n, hipMemcpyDeviceToHost, st ) );
      TNLHIP_CHECK( hipStreamSynchronize( st ) );
      for( int32_t i = 0; i < n; i++ )
         q.x[ (int64_t) i * q.k + j ] = w.hx[ i ];
      out.converged = res.converged;
      out.outer_iterations = res.outer_iterations;
      out.inner_iterations = res.inner_iterations;
      out.residue = res.residue;
   } );
   if( status != TNLHIP_SUCCESS ) {
      out.converged = 0;
      out.residue = std::nan( "" );
   }
   return out;
}

namespace {

int64_t requestEntries( const tnlhip_server_s& S, const Request& q )
{
   return (int64_t) S.matrices[ q.matrix ].n * q.k;
}

int kpOf( int k )
{
   int kp = 1;
   while( kp < k )
      kp <<= 1;
   return kp;
}

// admit the head of the queue while it fits; returns the number admitted
int admit( tnlhip_server_s& S )
{
   int count = 0;
   while( ! S.queue.empty() ) {
      Request& q = S.queue.front();
      const int64_t E = requestEntries( S, q );
      int slot = -1;
      for( int i = 0; i < S.cfg.max_requests; i++ )
         if( ! S.slotUsed[ i ] ) {
            slot = i;
            break;
         }
      if( slot < 0 )
         break;
      const int64_t off = S.vecAlloc.alloc( E );
      if( off < 0 )
         break;
      const ResidentMatrix& M = S.matrices[ q.matrix ];
      ReqStatic R{};
      R.n = M.n;
      R.k = q.k;
      R.kp = kpOf( q.k );
      R.rpt = kTile / R.kp;
      R.tiles = ( M.n + R.rpt * kRowsPerLane - 1 ) / ( R.rpt * kRowsPerLane );
      R.sliceBase = M.sliceBase;
      R.maxOuter = q.params.max_outer_iterations;
      R.maxInner = q.params.max_inner_iterations;
      R.vec = off;
      R.mat = M.mat;
      R.shift = q.shift;
      R.tol = q.params.tolerance;
      R.innerTol2 = q.params.inner_tolerance * q.params.inner_tolerance;
      R.stagnation = q.params.stagnation_limit;
      S.hStatic[ slot ] = R;
      ReqDyn dyn{};
      dyn.mask = 1 << P_CORRECT;
      dyn.bits[ P_CORRECT ] = ( 1u << q.k ) - 1u;
      S.hDyn[ slot ] = dyn;
      for( int j = 0; j < kMaxRhs; j++ ) {
         Sys s{};
         s.phase = j < q.k ? P_CORRECT : P_IDLE;
         s.normb = -1;
         s.scale = 0;
         S.hSys[ slot * kMaxRhs + j ] = s;
      }
      hipStream_t st = S.stream;
      TNLHIP_CHECK( hipMemcpyAsync( S.rs.ptr + slot, S.hStatic + slot, sizeof( ReqStatic ), hipMemcpyHostToDevice, st ) );
      TNLHIP_CHECK( hipMemcpyAsync( S.rd.ptr + slot, S.hDyn + slot, sizeof( ReqDyn ), hipMemcpyHostToDevice, st ) );
      TNLHIP_CHECK( hipMemcpyAsync( S.sys.ptr + slot * kMaxRhs, S.hSys + slot * kMaxRhs, sizeof( Sys ) * kMaxRhs,
                                    hipMemcpyHostToDevice, st ) );
      double* b = S.a64.ptr + 3 * off + E;
      TNLHIP_CHECK( hipMemcpyAsync( b, q.b, sizeof( double ) * E, hipMemcpyHostToDevice, st ) );
      admitKernel<<< gridFor( E, 256 ), 256, 0, st >>>( E, S.a64.ptr + 3 * off, S.a32.ptr + 4 * off );
      TNLHIP_CHECK( hipGetLastError() );
      q.slot = slot;
      S.slotUsed[ slot ] = true;
      S.slots[ slot ] = q;
      S.matrices[ q.matrix ].users++;
      S.queue.pop_front();
      S.tilesDirty = true;
      count++;
   }
   return count;
}

void rebuildTiles( tnlhip_server_s& S )
{
   int32_t t = 0;
   for( int i = 0; i < S.cfg.max_requests; i++ ) {
      S.hTileFirst[ i ] = t;
      if( ! S.slotUsed[ i ] )
         continue;
      const ReqStatic& R = S.hStatic[ i ];
      for( int q = 0; q < R.tiles; q++ )
         S.hTileReq[ t + q ] = i;
      t += R.tiles;
   }
   TNLHIP_REQUIRE( t <= S.maxTiles, "tile table overflow" );
   S.hCtl->numTiles = t;
   // walk order: XCD x takes tiles [x T / 8, (x + 1) T / 8) of every request of T tiles, in
   // request order; tile ids (and so the partials' order) stay request-contiguous
   int32_t k = 0;
   for( int x = 0; x < kServerXcds; x++ ) {
      S.hCtl->xcdStart[ x ] = k;
      for( int i = 0; i < S.cfg.max_requests; i++ ) {
         if( ! S.slotUsed[ i ] )
            continue;
         const int64_t T = S.hStatic[ i ].tiles;
         const int32_t a = (int32_t) ( x * T / kServerXcds ), b = (int32_t) ( ( x + 1 ) * T / kServerXcds );
         for( int32_t q = a; q < b; q++ )
            S.hTileOrder[ k++ ] = S.hTileFirst[ i ] + q;
      }
   }
   S.hCtl->xcdStart[ kServerXcds ] = k;
   TNLHIP_REQUIRE( k == t, "tile walk does not cover the tile table" );
   hipStream_t st = S.stream;
   if( t > 0 ) {
      TNLHIP_CHECK( hipMemcpyAsync( S.tileReq.ptr, S.hTileReq, sizeof( int32_t ) * t, hipMemcpyHostToDevice, st ) );
      TNLHIP_CHECK( hipMemcpyAsync( S.tileOrder.ptr, S.hTileOrder, sizeof( int32_t ) * t, hipMemcpyHostToDevice, st ) );
   }
   TNLHIP_CHECK( hipMemcpyAsync( S.tileFirst.ptr, S.hTileFirst, sizeof( int32_t ) * S.cfg.max_requests,
                                 hipMemcpyHostToDevice, st ) );
   TNLHIP_CHECK( hipMemcpyAsync( S.ctl.ptr, S.hCtl, sizeof( Ctl ), hipMemcpyHostToDevice, st ) );
   S.tilesDirty = false;
}

void releaseMatrixIfUnused( tnlhip_server_s& S, int id )
{
   ResidentMatrix& M = S.matrices[ id ];
   if( M.removing && M.users == 0 && M.mat >= 0 ) {
      S.matAlloc.release( M.mat, M.slots );
      S.sliceAlloc.release( M.sliceBase, M.sliceCount );
      M.mat = -1;
      M.sliceBase = -1;
      M.base.reset();
      M.rowPointers.reset();
      M.columns.reset();
      M.values.reset();
   }
}

}  // namespace

extern "C" {

void tnlhip_server_config_default( tnlhip_server_config* c )
{
   if( ! c )
      return;
   std::memset( c, 0, sizeof( *c ) );
   c->matrix_capacity = 64ll << 20;   // 64 Mi slots: 1 GiB (int32 column + FP64 + FP32 value)
   c->slice_capacity = 4ll << 20;
   c->vector_capacity = 64ll << 20;   // 64 Mi entries: 2.5 GiB (3 FP64 + 4 FP32 vectors)
   c->max_requests = 256;
   c->iterations_per_launch = 8;
   c->use_graph = 1;
   c->staging_capacity = 32ll << 20;
   c->gmres_workers = 4;
}

void tnlhip_request_params_default( tnlhip_request_params* p )
{
   if( ! p )
      return;
   std::memset( p, 0, sizeof( *p ) );
   p->tolerance = 1e-10;
   p->inner_tolerance = 3e-6;
   p->max_outer_iterations = 20;
   p->max_inner_iterations = 5000;
   p->method = TNLHIP_CG;
   p->stagnation_limit = 2;
}

int tnlhip_server_footprint( int32_t rows, int64_t nnz, int32_t max_row, int32_t nrhs, int64_t* matrix_bytes,
                             int64_t* request_bytes )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( rows >= 0 && nnz >= 0 && max_row >= 0 && nrhs >= 1 && nrhs <= kMaxRhs, "bad sizes" );
      const int64_t slots = std::max< int64_t >( nnz, (int64_t) ( ( rows + 31 ) / 32 ) * 32 * max_row );
      if( matrix_bytes )
         *matrix_bytes = slots * ( 4 + 8 + 4 ) + 4 * ( rows / 32 + 2 );
      if( request_bytes )
         *request_bytes = (int64_t) rows * nrhs * ( 3 * 8 + 4 * 4 );
   } );
}

int tnlhip_server_create( const tnlhip_server_config* config, tnlhip_server* out )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( out, "NULL argument" );
      *out = nullptr;
      auto S = std::make_unique< tnlhip_server_s >();
      tnlhip_server_config_default( &S->cfg );
      if( config )
         S->cfg = *config;
      tnlhip_server_config& c = S->cfg;
      TNLHIP_REQUIRE( c.matrix_capacity > 0 && c.vector_capacity > 0 && c.slice_capacity > 1 && c.max_requests > 0
                         && c.iterations_per_launch > 0 && c.staging_capacity > 0,
                      "server capacities must be positive" );
      TNLHIP_REQUIRE( c.vector_capacity < ( 1ll << 40 ) && c.matrix_capacity < ( 1ll << 40 ), "capacity too large" );
      // non-blocking: the legacy default stream, which the conversions of the GMRES workers use,
      // does not wait for the batched launches (nor they for it)
      TNLHIP_CHECK( hipStreamCreateWithFlags( &S->stream, hipStreamNonBlocking ) );
      S->cols.resize( c.matrix_capacity );
      S->v64.resize( c.matrix_capacity );
      S->v32.resize( c.matrix_capacity );
      S->slices.resize( c.slice_capacity );
      S->a64.resize( 3 * c.vector_capacity );
      S->a32.resize( 4 * c.vector_capacity );
      S->maxTiles = c.vector_capacity / 128 + 2 * (int64_t) c.max_requests;
      TNLHIP_REQUIRE( S->maxTiles < ( 1ll << 31 ), "too many tiles" );
      S->tileReq.resize( S->maxTiles );
      S->tileOrder.resize( S->maxTiles );
      S->tileFirst.resize( c.max_requests );
      S->part.resize( S->maxTiles * 2 * kMaxRhs );
      S->rs.resize( c.max_requests );
      S->rd.resize( c.max_requests );
      S->sys.resize( (int64_t) c.max_requests * kMaxRhs );
      S->ctl.resize( 1 );
      S->staging.resize( c.staging_capacity );
      S->stagingV.resize( c.staging_capacity );
      S->matAlloc.init( c.matrix_capacity );
      S->sliceAlloc.init( c.slice_capacity );
      S->vecAlloc.init( c.vector_capacity );
      S->slots.resize( c.max_requests );
      S->slotUsed.assign( c.max_requests, false );
      TNLHIP_CHECK( hipHostMalloc( (void**) &S->hTileReq, sizeof( int32_t ) * S->maxTiles ) );
      TNLHIP_CHECK( hipHostMalloc( (void**) &S->hTileFirst, sizeof( int32_t ) * c.max_requests ) );
      TNLHIP_CHECK( hipHostMalloc( (void**) &S->hTileOrder, sizeof( int32_t ) * S->maxTiles ) );
      TNLHIP_CHECK( hipHostMalloc( (void**) &S->hDyn, sizeof( ReqDyn ) * c.max_requests ) );
      TNLHIP_CHECK( hipHostMalloc( (void**) &S->hStatic, sizeof( ReqStatic ) * c.max_requests ) );
      TNLHIP_CHECK( hipHostMalloc( (void**) &S->hSys, sizeof( Sys ) * c.max_requests * kMaxRhs ) );
      TNLHIP_CHECK( hipHostMalloc( (void**) &S->hCtl, sizeof( Ctl ) ) );
      std::memset( S->hStatic, 0, sizeof( ReqStatic ) * c.max_requests );
      std::memset( S->hDyn, 0, sizeof( ReqDyn ) * c.max_requests );
      TNLHIP_CHECK( hipMemcpyAsync( S->rs.ptr, S->hStatic, sizeof( ReqStatic ) * c.max_requests, hipMemcpyHostToDevice,
                                    S->stream ) );
      TNLHIP_CHECK(
         hipMemcpyAsync( S->rd.ptr, S->hDyn, sizeof( ReqDyn ) * c.max_requests, hipMemcpyHostToDevice, S->stream ) );
      rebuildTiles( *S );
      TNLHIP_CHECK( hipStreamSynchronize( S->stream ) );
      if( c.use_graph ) {
         TNLHIP_CHECK( hipStreamBeginCapture( S->stream, hipStreamCaptureModeThreadLocal ) );
         S->enqueueIterations( c.iterations_per_launch );
         TNLHIP_CHECK( hipStreamEndCapture( S->stream, &S->graph ) );
         TNLHIP_CHECK( hipGraphInstantiate( &S->exec, S->graph, nullptr, nullptr, 0 ) );
      }
      TNLHIP_CHECK( hipStreamSynchronize( S->stream ) );
      TNLHIP_REQUIRE( c.gmres_workers >= 0 && c.gmres_workers <= 16, "gmres_workers must be in [0, 16]" );
      TNLHIP_CHECK( hipGetDevice( &S->device ) );
      for( int i = 0; i < c.gmres_workers; i++ )
         S->workers.emplace_back( [ srv = S.get() ] { srv->gmresWorker(); } );
      *out = S.release();
   } );
}

int tnlhip_server_destroy( tnlhip_server S )
{
   return guarded( [ & ] { delete S; } );
}

int tnlhip_server_add_matrix( tnlhip_server S, int32_t rows, const int32_t* row_pointers, const int32_t* columns,
                              const double* values, int32_t* matrix_id )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( S && row_pointers && matrix_id && rows > 0, "bad argument" );
      const int64_t nnz = row_pointers[ rows ];
      TNLHIP_REQUIRE( nnz >= 0 && ( nnz == 0 || ( columns && values ) ), "bad CSR arrays" );
      std::vector< int32_t > caps( rows );
      for( int32_t i = 0; i < rows; i++ ) {
         caps[ i ] = row_pointers[ i + 1 ] - row_pointers[ i ];
         TNLHIP_REQUIRE( caps[ i ] >= 0, "row pointers must not decrease" );
      }
      for( int64_t i = 0; i < nnz; i++ )
         TNLHIP_REQUIRE( columns[ i ] >= 0 && columns[ i ] < rows, "column index out of range (square matrices only)" );
      HostSegments seg;
      buildSegments( seg, KIND_SLICED, rows, caps.data() );
      const int64_t slots = seg.sliceOffsets.back();
      const int32_t nslices = (int32_t) seg.sliceOffsets.size();
      TNLHIP_REQUIRE( rows + 1 + nnz <= S->cfg.staging_capacity && nnz <= S->cfg.staging_capacity,
                      "matrix larger than the server's staging buffer" );
      const int64_t mat = S->matAlloc.alloc( slots );
      if( mat < 0 )
         throw Error( TNLHIP_ERROR_CAPACITY, "no room for the matrix in the server's matrix arena" );
      const int64_t sb = S->sliceAlloc.alloc( nslices );
      if( sb < 0 ) {
         S->matAlloc.release( mat, slots );
         throw Error( TNLHIP_ERROR_CAPACITY, "no room for the matrix in the server's slice arena" );
      }
      hipStream_t st = S->stream;
      TNLHIP_CHECK( hipStreamSynchronize( st ) );
      TNLHIP_CHECK( hipMemcpyAsync( S->slices.ptr + sb, seg.sliceOffsets.data(), sizeof( int32_t ) * nslices,
                                    hipMemcpyHostToDevice, st ) );
      int32_t* drp = S->staging.ptr;
      int32_t* dci = S->staging.ptr + rows + 1;
      TNLHIP_CHECK( hipMemcpyAsync( drp, row_pointers, sizeof( int32_t ) * ( rows + 1 ), hipMemcpyHostToDevice, st ) );
      if( nnz > 0 ) {
         TNLHIP_CHECK( hipMemcpyAsync( dci, columns, sizeof( int32_t ) * nnz, hipMemcpyHostToDevice, st ) );
         TNLHIP_CHECK( hipMemcpyAsync( S->stagingV.ptr, values, sizeof( double ) * nnz, hipMemcpyHostToDevice, st ) );
      }
      scatterKernel<<< ( rows + 255 ) / 256, 256, 0, st >>>( rows, drp, dci, S->stagingV.ptr, S->slices.ptr + sb, mat,
                                                             S->cols.ptr, S->v64.ptr, S->v32.ptr );
      TNLHIP_CHECK( hipGetLastError() );
      TNLHIP_CHECK( hipStreamSynchronize( st ) );
      ResidentMatrix M;
      M.n = rows;
      M.mat = mat;
      M.slots = slots;
      M.sliceBase = (int32_t) sb;
      M.sliceCount = nslices;
      if( ! S->workers.empty() ) {
         M.rowPointers = std::make_shared< const std::vector< int32_t > >( row_pointers, row_pointers + rows + 1 );
         M.columns = std::make_shared< const std::vector< int32_t > >( columns, columns + nnz );
         M.values = std::make_shared< const std::vector< double > >( values, values + nnz );
      }
      std::lock_guard< std::mutex > lock( S->gmresMutex );
      M.generation = S->nextGeneration++;
      int id = -1;
      for( size_t i = 0; i < S->matrices.size(); i++ )
         if( S->matrices[ i ].mat < 0 && S->matrices[ i ].removing ) {
            id = (int) i;
            break;
         }
      if( id < 0 ) {
         id = (int) S->matrices.size();
         S->matrices.push_back( M );
      }
      else
         S->matrices[ id ] = M;
      *matrix_id = id;
   } );
}

int tnlhip_server_remove_matrix( tnlhip_server S, int32_t id )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( S && id >= 0 && id < (int) S->matrices.size() && ! S->matrices[ id ].removing, "unknown matrix" );
      for( const Request& q : S->queue )
         TNLHIP_REQUIRE( q.matrix != id, "the matrix has queued requests" );
      std::lock_guard< std::mutex > lock( S->gmresMutex );
      S->matrices[ id ].removing = true;
      releaseMatrixIfUnused( *S, id );
   } );
}

int tnlhip_server_submit( tnlhip_server S, int32_t matrix_id, double shift, int32_t nrhs, const double* b, double* x,
                          const tnlhip_request_params* params, int64_t* ticket )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( S && b && x && ticket, "NULL argument" );
      TNLHIP_REQUIRE( matrix_id >= 0 && matrix_id < (int) S->matrices.size() && ! S->matrices[ matrix_id ].removing,
                      "unknown matrix" );
      TNLHIP_REQUIRE( nrhs >= 1 && nrhs <= kMaxRhs, "1 to 16 right-hand sides per request" );
      Request q;
      tnlhip_request_params_default( &q.params );
      if( params )
         q.params = *params;
      if( q.params.method != TNLHIP_CG && q.params.method != TNLHIP_GMRES )
         throw Error( TNLHIP_ERROR_UNSUPPORTED, "the solve server runs CG and GMRES requests" );
      if( q.params.method == TNLHIP_GMRES && S->workers.empty() )
         throw Error( TNLHIP_ERROR_UNSUPPORTED, "the server was created without GMRES workers (gmres_workers = 0)" );
      TNLHIP_REQUIRE( q.params.tolerance > 0 && q.params.inner_tolerance > 0 && q.params.max_outer_iterations >= 0
                         && q.params.max_inner_iterations > 0,
                      "bad request parameters" );
      q.matrix = matrix_id;
      q.k = nrhs;
      q.shift = shift;
      q.b = b;
      q.x = x;
      if( q.params.method == TNLHIP_GMRES ) {
         q.ticket = S->nextTicket++;
         auto task = std::make_shared< GmresTask >();
         task->q = q;
         task->remaining = q.k;
         task->out.converged = 1;
         {
            std::lock_guard< std::mutex > lock( S->gmresMutex );
            S->matrices[ matrix_id ].users++;
            for( int j = 0; j < q.k; j++ )
               S->gmresQueue.push_back( GmresJob{ task, j } );
            S->gmresWaiting++;
         }
         S->gmresWake.notify_all();
         *ticket = q.ticket;
         return;
      }
      const int64_t E = requestEntries( *S, q );
      if( E > S->cfg.vector_capacity )
         throw Error( TNLHIP_ERROR_CAPACITY, "request larger than the server's vector arena" );
      q.ticket = S->nextTicket++;
      S->queue.push_back( q );
      *ticket = q.ticket;
   } );
}

int tnlhip_server_progress( tnlhip_server S, int64_t* finished, int32_t capacity, int32_t* count )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( S && count && ( capacity == 0 || finished ), "NULL argument" );
      *count = 0;
      hipStream_t st = S->stream;
      {
         // finished GMRES requests
         std::lock_guard< std::mutex > lock( S->gmresMutex );
         for( auto& d : S->gmresDone ) {
            d.second.launches = S->launches;
            S->finished[ d.first.ticket ] = d.second;
            S->matrices[ d.first.matrix ].users--;
            releaseMatrixIfUnused( *S, d.first.matrix );
            if( *count < capacity )
               finished[ *count ] = d.first.ticket;
            ( *count )++;
         }
         S->gmresDone.clear();
      }
      admit( *S );
      bool any = false;
      for( int i = 0; i < S->cfg.max_requests; i++ )
         any = any || S->slotUsed[ i ];
      if( ! any )
         return;
      if( S->tilesDirty )
         rebuildTiles( *S );
      if( S->exec )
         TNLHIP_CHECK( hipGraphLaunch( S->exec, st ) );
      else
         S->enqueueIterations( S->cfg.iterations_per_launch );
      TNLHIP_CHECK( hipGetLastError() );
      S->launches++;
      TNLHIP_CHECK( hipMemcpyAsync( S->hDyn, S->rd.ptr, sizeof( ReqDyn ) * S->cfg.max_requests, hipMemcpyDeviceToHost,
                                    st ) );
      TNLHIP_CHECK( hipStreamSynchronize( st ) );
      std::vector< int > done;
      for( int i = 0; i < S->cfg.max_requests; i++ )
         if( S->slotUsed[ i ] && ! ( S->hDyn[ i ].mask & kLive ) )
            done.push_back( i );
      if( done.empty() )
         return;
      for( int i : done ) {
         const Request& q = S->slots[ i ];
         const ReqStatic& R = S->hStatic[ i ];
         const int64_t E = (int64_t) R.n * R.k;
         TNLHIP_CHECK( hipMemcpyAsync( q.x, S->a64.ptr + 3 * R.vec, sizeof( double ) * E, hipMemcpyDeviceToHost, st ) );
         TNLHIP_CHECK( hipMemcpyAsync( S->hSys + i * kMaxRhs, S->sys.ptr + i * kMaxRhs, sizeof( Sys ) * kMaxRhs,
                                       hipMemcpyDeviceToHost, st ) );
      }
      TNLHIP_CHECK( hipStreamSynchronize( st ) );
      for( int i : done ) {
         const Request& q = S->slots[ i ];
         tnlhip_request_result r{};
         r.converged = 1;
         for( int j = 0; j < q.k; j++ ) {
            const Sys& s = S->hSys[ i * kMaxRhs + j ];
            r.converged = r.converged && s.phase == P_DONE;
            r.outer_iterations = std::max( r.outer_iterations, s.outerIt );
            r.inner_iterations += s.innerTotal;
            r.residue = std::max( r.residue, s.res );
         }
         r.launches = S->launches;
         S->finished[ q.ticket ] = r;
         S->vecAlloc.release( S->hStatic[ i ].vec, (int64_t) S->hStatic[ i ].n * S->hStatic[ i ].k );
         S->matrices[ q.matrix ].users--;
         releaseMatrixIfUnused( *S, q.matrix );
         S->slotUsed[ i ] = false;
         S->tilesDirty = true;
         if( *count < capacity )
            finished[ *count ] = q.ticket;
         ( *count )++;
      }
   } );
}

int tnlhip_server_result( tnlhip_server S, int64_t ticket, tnlhip_request_result* result )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( S && result, "NULL argument" );
      auto it = S->finished.find( ticket );
      TNLHIP_REQUIRE( it != S->finished.end(), "ticket not finished (or already taken)" );
      *result = it->second;
      S->finished.erase( it );
   } );
}

int tnlhip_server_stats( tnlhip_server S, tnlhip_server_stats_t* stats )
{
   return guarded( [ & ] {
      TNLHIP_REQUIRE( S && stats, "NULL argument" );
      std::memset( stats, 0, sizeof( *stats ) );
      for( int i = 0; i < S->cfg.max_requests; i++ )
         if( S->slotUsed[ i ] ) {
            stats->resident_requests++;
            stats->resident_entries += (int64_t) S->hStatic[ i ].n * S->hStatic[ i ].k;
         }
      stats->queued_requests = (int32_t) S->queue.size();
      {
         std::lock_guard< std::mutex > lock( S->gmresMutex );
         stats->queued_requests += S->gmresWaiting;
         // on a worker, or finished there and not yet reported by tnlhip_server_progress: a
         // caller that stops once nothing is resident must not lose the latter
         stats->resident_requests += S->gmresBusy + (int32_t) S->gmresDone.size();
      }
      stats->vector_free = S->vecAlloc.freeTotal();
      stats->matrix_free = S->matAlloc.freeTotal();
      stats->launches = S->launches;
      stats->tiles = S->hCtl->numTiles;
   } );
}

}  // extern "C"
