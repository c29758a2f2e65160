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
#include "views.hpp"

namespace tnlhip {

// ------------------------------------------------------------------ result keeper

// SparseMatrixView.hpp:447-490: direct / alpha / beta / general keepers; mode 4 divides by
// a diagonal, which fuses Preconditioners::Diagonal::solve (x = b / diag) into the product.
template< typename A, typename Y >
struct Keeper
{
   Y* y;
   A alpha, beta;
   const Y* diag;
   int mode;
   const int32_t* gate;  // solver flags: the launch is skipped when gate[0] != 0 (and, with
   int gateMode;         // gateMode 2, unless gate[1] != 0)
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
         default: out = v / (A) diag[ row ]; break;
      }
      y[ row ] = (Y) out;
   }
};

template< bool EXACT, typename A >
__device__ __forceinline__ A madd( A acc, A a, A b )
{
   if constexpr( EXACT )
      return mulAddExact( acc, a, b );
   else
      return acc + a * b;
}

// ------------------------------------------------------------------ CSR: G lanes per row

// G = 1: CSRScalar (one lane per row, exact Host order); G = 2..32: CSRLight / CSRHybrid;
// G = 64: CSRVector (a whole wavefront per row).  Groups of G lanes stride over the row
// with contiguous loads, then reduce across the group with cross-lane shuffles.
template< int G, bool EXACT, typename V, typename A, typename Y >
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
}

// ------------------------------------------------------------------ CSR-Adaptive

// One wavefront per block descriptor (Kernels/CSRAdaptiveKernelView.hpp:33-170 for the
// scheme).  STREAM: the wave streams the elements of several short rows into its LDS
// slice with fully coalesced loads, then every lane sums whole rows out of LDS.
// VECTOR: one row per wave.  LONG: one row split over the waves left in the workgroup,
// partial sums combined in LDS in a fixed order.
template< int STREAM, typename V, typename A, typename Y >
__global__ __launch_bounds__( 256 ) void csrAdaptiveKernel( int32_t nblocks, const int4* __restrict__ desc,
                                                            const int32_t* __restrict__ off,
                                                            const int32_t* __restrict__ col,
                                                            const V* __restrict__ val, const Y* __restrict__ x,
                                                            Keeper< A, Y > keep )
{
   if( keep.closed() )
      return;
   __shared__ A lds[ 4 ][ STREAM ];
   __shared__ A longPart[ 4 ];
   const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
   const int32_t groups = ( nblocks + 3 ) / 4;
   for( int32_t wg = blockIdx.x; wg < groups; wg += gridDim.x ) {
      const int32_t di = wg * 4 + wave;
      const bool active = di < nblocks;
      const int4 d = active ? desc[ di ] : make_int4( 0, 0, 0, 0 );
      const int type = d.x & 0xff;
      const int32_t firstRow = d.y;
      A vsum = 0;
      if( active ) {
         const int32_t b = off[ firstRow ];
         if( type == BLOCK_STREAM ) {
            for( int32_t e = lane; e < d.z; e += kWave ) {
               const int32_t c = col[ b + e ];
               lds[ wave ][ e ] = c >= 0 ? toCompute( val[ b + e ], A() ) * (A) x[ c ] : (A) 0;
            }
         }
         else {
            const int32_t e = off[ firstRow + 1 ];
            const int32_t widx = type == BLOCK_LONG ? ( d.x >> 8 ) & 0xff : 0;
            const int32_t wcnt = type == BLOCK_LONG ? ( d.x >> 16 ) & 0xff : 1;
            A p = 0;
            for( int32_t k = b + widx * kWave + lane; k < e; k += kWave * wcnt ) {
               const int32_t c = col[ k ];
               if( c >= 0 )
                  p += toCompute( val[ k ], A() ) * (A) x[ c ];
            }
            vsum = groupSum< kWave >( p );
            if( type == BLOCK_LONG && lane == 0 )
               longPart[ wave ] = vsum;
         }
      }
      __syncthreads();
      if( active ) {
         if( type == BLOCK_STREAM ) {
            const int32_t b = off[ firstRow ];
            for( int32_t r = lane; r < d.w; r += kWave ) {
               const int32_t rb = off[ firstRow + r ] - b, re = off[ firstRow + r + 1 ] - b;
               A s = 0;
               for( int32_t k = rb; k < re; k++ )
                  s += lds[ wave ][ k ];
               keep( firstRow + r, s );
            }
         }
         else if( type == BLOCK_VECTOR ) {
            if( lane == 0 )
               keep( firstRow, vsum );
         }
         else if( ( ( d.x >> 8 ) & 0xff ) == 0 && lane == 0 ) {
            const int32_t wcnt = ( d.x >> 16 ) & 0xff;
            A s = 0;
            for( int32_t w = 0; w < wcnt; w++ )
               s += longPart[ wave + w ];
            keep( firstRow, s );
         }
      }
      __syncthreads();
   }
}

// ------------------------------------------------------------------ Ellpack / SlicedEllpack

// One lane per row, column-major storage: consecutive lanes read consecutive addresses
// for every local index (EllpackView.hpp:361-378, SlicedEllpackView.hpp:248-290).
// Padding is trailing in each row, so predication on column >= 0 gives the Host sum.
template< typename V, typename A, typename Y >
__global__ __launch_bounds__( 256 ) void ellpackKernel( int32_t segmentSize, int32_t aligned,
                                                        const int32_t* __restrict__ col,
                                                        const V* __restrict__ val, const Y* __restrict__ x,
                                                        Keeper< A, Y > keep, int32_t first, int32_t last )
{
   if( keep.closed() )
      return;
   const int64_t stride = (int64_t) gridDim.x * blockDim.x;
   for( int64_t r = first + (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < last; r += stride ) {
      A s = 0;
#pragma unroll 4
      for( int32_t l = 0; l < segmentSize; l++ ) {
         const int64_t g = r + (int64_t) aligned * l;
         const int32_t c = col[ g ];
         if( c >= 0 )
            s = mulAddExact( s, toCompute( val[ g ], A() ), (A) x[ c ] );
      }
      keep( r, s );
   }
}

template< typename V, typename A, typename Y >
__global__ __launch_bounds__( 256 ) void slicedKernel( const int32_t* __restrict__ sliceOffsets,
                                                       const int32_t* __restrict__ col,
                                                       const V* __restrict__ val, const Y* __restrict__ x,
                                                       Keeper< A, Y > keep, int32_t first, int32_t last )
{
   if( keep.closed() )
      return;
   const int64_t stride = (int64_t) gridDim.x * blockDim.x;
   for( int64_t r = first + (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < last; r += stride ) {
      const int32_t sl = (int32_t) ( r / kSliceSize );
      const int32_t b = sliceOffsets[ sl ] + (int32_t) ( r % kSliceSize ), e = sliceOffsets[ sl + 1 ];
      A s = 0;
#pragma unroll 4
      for( int32_t g = b; g < e; g += kSliceSize ) {
         const int32_t c = col[ g ];
         if( c >= 0 )
            s = mulAddExact( s, toCompute( val[ g ], A() ), (A) x[ c ] );
      }
      keep( r, s );
   }
}

// ------------------------------------------------------------------ ChunkedEllpack

// One 256-lane workgroup per slice, one lane per chunk (the format is built around 256
// chunks per slice): chunk c's elements sit at pointer + c + 256*k, coalesced across the
// workgroup.  Chunk sums meet in LDS and each row adds up its chunk range
// (rowToChunkMapping is an inclusive per-slice prefix of chunk counts).
template< typename V, typename A, typename Y >
__global__ __launch_bounds__( 256 ) void chunkedKernel( int32_t nslices, const int4* __restrict__ slices,
                                                        const int32_t* __restrict__ rowToChunk,
                                                        const int32_t* __restrict__ col,
                                                        const V* __restrict__ val, const Y* __restrict__ x,
                                                        Keeper< A, Y > keep, int32_t first, int32_t last )
{
   if( keep.closed() )
      return;
   __shared__ A part[ kChunksInSlice ];
   const int t = threadIdx.x;
   for( int32_t sl = blockIdx.x; sl < nslices; sl += gridDim.x ) {
      const int4 s = slices[ sl ];  // size, chunkSize, firstSegment, pointer
      if( s.z >= last || s.z + s.x <= first )
         continue;  // uniform over the workgroup
      A p = 0;
      for( int32_t k = 0; k < s.y; k++ ) {
         const int64_t g = (int64_t) s.w + t + (int64_t) kChunksInSlice * k;
         const int32_t c = col[ g ];
         if( c >= 0 )
            p += toCompute( val[ g ], A() ) * (A) x[ c ];
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
// the end each lane fetches the sum of its own row's rank (rowPermArray) and stores it, so
// y is written coalesced.
template< typename V, typename A, typename Y >
__global__ __launch_bounds__( 256 ) void biellpackKernel( int32_t n, int32_t strips, const int32_t* __restrict__ rowPerm,
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
   for( int32_t pair = wave; 2 * pair < strips; pair += waves ) {
      const int32_t strip = 2 * pair + half;
      const bool valid = strip < strips;
      A acc = 0;
      if( valid ) {
         const int32_t* gp = groupPointers + (int64_t) strip * kGroupsPerStrip;
         // group 5 (height 1) is always empty: the widths are taken at ranks 15, 7, 3, 1, 0
#pragma unroll
         for( int grp = 0; grp < kLogStrip; grp++ ) {
            const int32_t base = gp[ grp ], size = gp[ grp + 1 ] - base;
            const int h = kStrip >> grp;
            A p = 0;
            for( int32_t e = l; e < size; e += kStrip ) {
               const int32_t c = col[ base + e ];
               if( c >= 0 )
                  p += toCompute( val[ base + e ], A() ) * (A) x[ c ];
            }
            for( int o = h; o < kStrip; o <<= 1 )
               p += __shfl_xor( p, o, kWave );
            if( l < h )
               acc += p;
         }
      }
      const int64_t row = (int64_t) strip * kStrip + l;
      const bool mine = valid && row < n;
      const int rank = mine ? rowPerm[ row ] - strip * kStrip : 0;
      const A v = __shfl( acc, half * 32 + rank, kWave );
      if( mine && row >= first && row < last )
         keep( row, v );
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
                  const Y* diag, const int32_t* gate, int gateMode, hipStream_t st )
{
   Keeper< A, Y > keep{ y, (A) alpha, (A) beta, diag, 0, gate, gate ? gateMode : 0 };
   if( diag )
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
   switch( M.format ) {
      case TNLHIP_CSR_SCALAR:
         csrGroupedKernel< 1, true, V, A, Y >
            <<< gridFor( rows, 256 ), 256, 0, st >>>( M.offsets.ptr, col, val, x, keep, first, last );
         break;
      case TNLHIP_CSR_VECTOR:
      case TNLHIP_CSR_LIGHT:
      case TNLHIP_CSR_HYBRID:
      case TNLHIP_CSR_ADAPTIVE: {
         if( M.format == TNLHIP_CSR_ADAPTIVE && first == 0 && last == M.rows && M.adaptiveBlocks > 0 ) {
            constexpr int STREAM = 8192 / sizeof( A );
            const int32_t groups = ( M.adaptiveBlocks + 3 ) / 4;
            csrAdaptiveKernel< STREAM, V, A, Y ><<< gridFor( groups, 1, 8 ), 256, 0, st >>>(
               M.adaptiveBlocks, reinterpret_cast< const int4* >( M.adaptive.ptr ), M.offsets.ptr, col, val, x, keep );
            break;
         }
         const int G = M.format == TNLHIP_CSR_VECTOR ? 64 : M.lanes;
         const int gpb = 256 / G;
         const unsigned grid = gridFor( rows, gpb );
         switch( G ) {
            case 1:
               csrGroupedKernel< 1, false, V, A, Y ><<< grid, 256, 0, st >>>( M.offsets.ptr, col, val, x, keep, first, last );
               break;
            case 2:
               csrGroupedKernel< 2, false, V, A, Y ><<< grid, 256, 0, st >>>( M.offsets.ptr, col, val, x, keep, first, last );
               break;
            case 4:
               csrGroupedKernel< 4, false, V, A, Y ><<< grid, 256, 0, st >>>( M.offsets.ptr, col, val, x, keep, first, last );
               break;
            case 8:
               csrGroupedKernel< 8, false, V, A, Y ><<< grid, 256, 0, st >>>( M.offsets.ptr, col, val, x, keep, first, last );
               break;
            case 16:
               csrGroupedKernel< 16, false, V, A, Y ><<< grid, 256, 0, st >>>( M.offsets.ptr, col, val, x, keep, first, last );
               break;
            case 32:
               csrGroupedKernel< 32, false, V, A, Y ><<< grid, 256, 0, st >>>( M.offsets.ptr, col, val, x, keep, first, last );
               break;
            default:
               csrGroupedKernel< 64, false, V, A, Y ><<< grid, 256, 0, st >>>( M.offsets.ptr, col, val, x, keep, first, last );
               break;
         }
         break;
      }
      case TNLHIP_ELLPACK:
         ellpackKernel< V, A, Y ><<< gridFor( rows, 256 ), 256, 0, st >>>( M.meta.segmentSize, M.meta.alignedSize, col,
                                                                            val, x, keep, first, last );
         break;
      case TNLHIP_SLICED_ELLPACK:
         slicedKernel< V, A, Y >
            <<< gridFor( rows, 256 ), 256, 0, st >>>( M.sliceOffsets.ptr, col, val, x, keep, first, last );
         break;
      case TNLHIP_CHUNKED_ELLPACK:
         if( M.meta.numberOfSlices > 0 )
            chunkedKernel< V, A, Y ><<< gridFor( M.meta.numberOfSlices, 1 ), 256, 0, st >>>(
               M.meta.numberOfSlices, reinterpret_cast< const int4* >( M.slices.ptr ), M.rowToChunk.ptr, col, val, x,
               keep, first, last );
         break;
      case TNLHIP_BIELLPACK: {
         const int32_t strips = M.meta.virtualRows / kStrip;
         const int32_t pairs = ( strips + 1 ) / 2;
         if( strips > 0 && M.meta.allocated > 0 )
            biellpackKernel< V, A, Y ><<< gridFor( pairs, 4 ), 256, 0, st >>>(
               M.rows, strips, M.rowPerm.ptr, M.groupPointers.ptr, col, val, x, keep, first, last );
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

void spmvLaunch( const Matrix& A, const void* x, void* y, double alpha, double beta, int32_t first, int32_t last,
                 const void* diag, hipStream_t stream )
{
   spmvLaunchGated( A, x, y, alpha, beta, diag, nullptr, 0, stream, first, last );
}

}  // namespace tnlhip
