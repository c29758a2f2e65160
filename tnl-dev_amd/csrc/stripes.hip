// Column-striped execution plan of a CSR matrix: y = alpha A x + beta y for matrices whose x
// gathers do not stay in cache.
//
// The eight XCDs of the MI355X each have a private 4 MiB L2.  A matrix with columns spread
// uniformly over millions of rows (the lognormal irregular matrix: 2 M rows, 16 MB of FP64 x)
// gathers every x entry from the MALL or HBM as a whole cache line, since no XCD's L2 holds a
// useful share of x.  The plan splits the columns into S stripes of equal width and copies the
// stored entries stripe-major (stripe, row, then the row's own order), so that
//   - the workgroups of a stripe run on the XCDs that own it (workgroups are dispatched to the
//     XCDs round-robin: workgroup b runs on XCD b mod 8), and that stripe's x (at most 2 MiB)
//     stays in those XCDs' L2 while the entries stream past it with non-temporal loads;
//   - each (stripe, row) pair yields one partial sum, stored stripe-major in `partial`, and a
//     second kernel adds a row's S partials in stripe order and applies the keeper.
// The row's elements are summed left to right within each stripe and the stripe sums in stripe
// order: a fixed order (the same bits on every call), a reassociation of the Host row sum (a row's
// entries within a stripe are added in runs of at most 8 (FP64) / 16 (FP32) consecutive products,
// the runs left to right).
// Reference semantics: SparseMatrixView::vectorProduct (Matrices/SparseMatrixView.hpp:390-491);
// the storage of the matrix is untouched (the plan is a second, reordered copy).
#include "blas.hpp"
#include "spmv.hpp"

namespace tnlhip {

namespace {

constexpr int kStripeWaveRows = 256;  // rows per wavefront (4 per lane)
constexpr int kStripeRowsPerLane = kStripeWaveRows / kWave;
constexpr int kStripeBlockRows = 4 * kStripeWaveRows;  // rows per 256-lane workgroup

__device__ __forceinline__ int32_t ldntS( const int32_t* p ) { return __builtin_nontemporal_load( p ); }
__device__ __forceinline__ double ldntS( const double* p ) { return __builtin_nontemporal_load( p ); }
__device__ __forceinline__ float ldntS( const float* p ) { return __builtin_nontemporal_load( p ); }
__device__ __forceinline__ bf16 ldntS( const bf16* p )
{
   return __ushort_as_bfloat16( __builtin_nontemporal_load( reinterpret_cast< const unsigned short* >( p ) ) );
}

// Stripe of a column: columns [s * width, (s + 1) * width) belong to stripe s.
__device__ __forceinline__ int stripeOf( int32_t c, int32_t width, int S )
{
   const int s = c / width;
   return s < S ? s : S - 1;
}

// counts[s * n + r] = stored entries of row r in stripe s (padding -1 is not stored).
template< int S >
__global__ __launch_bounds__( 256 ) void stripeCountKernel( int32_t n, const int32_t* __restrict__ off,
                                                            const int32_t* __restrict__ col, int32_t width,
                                                            int32_t* __restrict__ counts )
{
   const int64_t stride = (int64_t) gridDim.x * blockDim.x;
   for( int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride ) {
      int32_t cnt[ S ];
#pragma unroll
      for( int s = 0; s < S; s++ )
         cnt[ s ] = 0;
      for( int32_t k = off[ r ]; k < off[ r + 1 ]; k++ ) {
         const int32_t c = col[ k ];
         if( c < 0 )
            continue;
         const int st = stripeOf( c, width, S );
#pragma unroll
         for( int s = 0; s < S; s++ )
            cnt[ s ] += st == s;
      }
#pragma unroll
      for( int s = 0; s < S; s++ )
         counts[ (int64_t) s * n + r ] = cnt[ s ];
   }
}

// Copies each stored entry of row r to its stripe's range, keeping the row's order.
template< int S, typename V >
__global__ __launch_bounds__( 256 ) void stripeScatterKernel( int32_t n, const int32_t* __restrict__ off,
                                                              const int32_t* __restrict__ col,
                                                              const V* __restrict__ val, int32_t width,
                                                              const int32_t* __restrict__ soff,
                                                              int32_t* __restrict__ scol, V* __restrict__ sval )
{
   const int64_t stride = (int64_t) gridDim.x * blockDim.x;
   for( int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride ) {
      int32_t pos[ S ];
#pragma unroll
      for( int s = 0; s < S; s++ )
         pos[ s ] = soff[ (int64_t) s * n + r ];
      for( int32_t k = off[ r ]; k < off[ r + 1 ]; k++ ) {
         const int32_t c = col[ k ];
         if( c < 0 )
            continue;
         const int st = stripeOf( c, width, S );
         int32_t g = 0;
#pragma unroll
         for( int s = 0; s < S; s++ )
            if( st == s )
               g = pos[ s ]++;
         scol[ g ] = c;
         sval[ g ] = val[ k ];
      }
   }
}

// Workgroup -> (stripe, block of kStripeBlockRows rows).  S < 8: stripe s belongs to the 8/S
// XCDs x with x mod S == s, which share its row blocks; S >= 8 (a multiple of 8): XCD x owns
// stripes x, x + 8, ... and takes them one after the other.
struct StripeTile
{
   int s;
   int64_t block;
};
__device__ __forceinline__ StripeTile stripeTile( int S, int64_t blocksPerStripe )
{
   const int xcd = blockIdx.x % 8;
   const int64_t k = blockIdx.x / 8;
   if( S < 8 ) {
      const int share = 8 / S;
      return StripeTile{ xcd % S, k * share + xcd / S };
   }
   return StripeTile{ xcd + 8 * (int) ( k / blocksPerStripe ), k % blocksPerStripe };
}

// One wavefront per kStripeWaveRows rows of one stripe.  The wave's entries are taken in
// chunks of CH: coalesced non-temporal loads (kLoads per lane per round), the x gathers
// (resident in the XCD's L2), the rounded products in the wave's LDS slice.  A row's part of a
// chunk is then summed in two balanced steps instead of by its own lane alone (rows of a
// lognormal matrix are hundreds of entries long beside rows of one, and one lane walking a long
// row would hold up its whole wavefront):
//   1. lane l adds the P = CH / 64 products at positions [P l, P l + P) left to right, cutting
//      a run wherever a row starts (a flag its owner set); each run's sum replaces the product
//      at the run's first position;
//   2. the owner of each row adds, in position order, the run starting at the row's first
//      position in the chunk and the runs starting at the lane boundaries P m inside the row.
// Fixed positions and orders: the same bits on every call.  Positions are padded by one slot
// per P in LDS (index p + p / P), so the P-strided reads of step 1 spread over the banks.
template< int CH, typename V, typename A, typename Y >
__global__ __launch_bounds__( 256 ) void stripeProductKernel( int S, int64_t blocksPerStripe, int32_t n,
                                                              int32_t first, int32_t last,
                                                              const int32_t* __restrict__ soff,
                                                              const int32_t* __restrict__ scol,
                                                              const V* __restrict__ sval, const Y* __restrict__ x,
                                                              A* __restrict__ partial, Gate gate )
{
   constexpr int kLoads = 8;
   constexpr int P = CH / kWave;  // products per lane in step 1
   constexpr int CHP = CH + CH / P;
   static_assert( CH % ( kLoads * kWave ) == 0, "chunk must be whole rounds of loads" );
   if( ! gate.open() )
      return;
   __shared__ A lds[ 4 ][ CHP ];
   __shared__ uint8_t startsRow[ 4 ][ CH ];
   const StripeTile t = stripeTile( S, blocksPerStripe );
   if( t.block >= blocksPerStripe )
      return;  // uniform over the workgroup; the kernel has no workgroup barrier
   const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
   const int64_t r0 = first + t.block * kStripeBlockRows + (int64_t) wave * kStripeWaveRows;
   if( r0 >= last )
      return;  // uniform over the wave
   const int64_t rEnd = min( r0 + kStripeWaveRows, (int64_t) last );
   const int32_t* o = soff + (int64_t) t.s * n;
   int32_t rb[ kStripeRowsPerLane ], re[ kStripeRowsPerLane ];
#pragma unroll
   for( int q = 0; q < kStripeRowsPerLane; q++ ) {
      const int64_t r = r0 + lane + q * kWave;
      rb[ q ] = r < rEnd ? o[ r ] : 0;
      re[ q ] = r < rEnd ? o[ r + 1 ] : 0;
   }
   // o[n] of stripe s is o[0] of stripe s + 1 (the offsets are one stripe-major scan)
   const int32_t b = o[ r0 ], e = o[ rEnd ];
   A sum[ kStripeRowsPerLane ];
#pragma unroll
   for( int q = 0; q < kStripeRowsPerLane; q++ )
      sum[ q ] = 0;
   A* L = lds[ wave ];
   uint8_t* F = startsRow[ wave ];
#pragma unroll
   for( int i = 0; i < P; i++ )
      F[ lane * P + i ] = 0;
   for( int32_t c0 = b; c0 < e; c0 += CH ) {  // uniform over the wave
      const int32_t size = min( CH, e - c0 );
      // row starts inside this chunk (rows with entries only: their starts are distinct)
#pragma unroll
      for( int q = 0; q < kStripeRowsPerLane; q++ )
         if( re[ q ] > rb[ q ] && rb[ q ] >= c0 && rb[ q ] < c0 + size )
            F[ rb[ q ] - c0 ] = 1;
#pragma unroll
      for( int u = 0; u < CH / ( kLoads * kWave ); u++ ) {
         const int32_t g0 = u * kLoads * kWave;
         if( g0 >= size )
            break;  // uniform over the wave
         int32_t c[ kLoads ];
         V v[ kLoads ];
#pragma unroll
         for( int k = 0; k < kLoads; k++ ) {
            const int32_t g = min( g0 + lane + k * kWave, size - 1 );
            c[ k ] = ldntS( scol + c0 + g );
            v[ k ] = ldntS( sval + c0 + g );
         }
         Y xv[ kLoads ];
#pragma unroll
         for( int k = 0; k < kLoads; k++ )
            xv[ k ] = x[ c[ k ] ];  // the plan stores no padding: every column is in range
#pragma unroll
         for( int k = 0; k < kLoads; k++ ) {
            const int32_t g = g0 + lane + k * kWave;
            if( g < size )
               L[ g + g / P ] = toCompute( v[ k ], A() ) * (A) xv[ k ];
         }
      }
      waveSync();
      // step 1: runs of the lane's P positions, cut at row starts
      const int32_t p0 = lane * P;
      if( p0 < size ) {
         A pr[ P ];
         uint8_t fl[ P ];
#pragma unroll
         for( int i = 0; i < P; i++ ) {
            pr[ i ] = L[ p0 + i + lane ];
            fl[ i ] = F[ p0 + i ];
         }
         A acc = pr[ 0 ];
         int at = 0;
#pragma unroll
         for( int i = 1; i < P; i++ ) {
            const bool inChunk = p0 + i < size;
            if( inChunk && fl[ i ] ) {
               L[ p0 + at + lane ] = acc;
               acc = pr[ i ];
               at = i;
            }
            else if( inChunk )
               acc += pr[ i ];
         }
         L[ p0 + at + lane ] = acc;
      }
      waveSync();
      // step 2: each row adds its runs in position order
#pragma unroll
      for( int q = 0; q < kStripeRowsPerLane; q++ ) {
         const int32_t lo = max( rb[ q ], c0 ) - c0, hi = min( re[ q ], c0 + size ) - c0;
         if( lo < hi ) {
            A s2 = L[ lo + lo / P ];
            for( int32_t m = lo / P + 1; m * P < hi; m++ )
               s2 += L[ m * P + m ];
            sum[ q ] += s2;
         }
      }
      // the flags this lane set are cleared before the next chunk sets its own
#pragma unroll
      for( int q = 0; q < kStripeRowsPerLane; q++ )
         if( re[ q ] > rb[ q ] && rb[ q ] >= c0 && rb[ q ] < c0 + size )
            F[ rb[ q ] - c0 ] = 0;
      waveSync();
   }
#pragma unroll
   for( int q = 0; q < kStripeRowsPerLane; q++ ) {
      const int64_t r = r0 + lane + q * kWave;
      if( r < rEnd )
         __builtin_nontemporal_store( sum[ q ], partial + (int64_t) t.s * n + r );
   }
}

// y[r] = keep( partial[0][r] + partial[1][r] + ... ), stripes in order.
template< typename A, typename Y >
__global__ __launch_bounds__( 256 ) void stripeCombineKernel( int S, int32_t n, int32_t first, int32_t last,
                                                              const A* __restrict__ partial, Keeper< A, Y > keep )
{
   if( keep.closed() )
      return;
   const int64_t stride = (int64_t) gridDim.x * blockDim.x;
   for( int64_t r = first + (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < last; r += stride ) {
      A s = __builtin_nontemporal_load( partial + r );
      for( int k = 1; k < S; k++ )
         s += __builtin_nontemporal_load( partial + (int64_t) k * n + r );
      keep( r, s );
   }
}

inline int64_t stripeBlocks( int32_t rows ) { return ( (int64_t) rows + kStripeBlockRows - 1 ) / kStripeBlockRows; }

}  // namespace

void buildStripePlan( Matrix& M, int S, hipStream_t st )
{
   TNLHIP_REQUIRE( S == 0 || S == 2 || S == 4 || S == 8 || S == 16 || S == 32,
                   "column_stripes is 0 (off), 2, 4, 8, 16 or 32" );
   if( S == 0 ) {
      M.stripes.reset();
      return;
   }
   TNLHIP_REQUIRE( kindOf( M.format ) == KIND_CSR, "column stripes apply to CSR matrices" );
   if( M.symmetric || M.dtype == TNLHIP_BINARY )
      throw Error( TNLHIP_ERROR_UNSUPPORTED, "column stripes of symmetric or binary matrices" );
   auto P = std::make_unique< StripePlan >();
   P->stripes = S;
   P->width = std::max< int32_t >( 1, (int32_t) ( ( (int64_t) M.cols + S - 1 ) / S ) );
   const int32_t n = M.rows;
   const int64_t cells = (int64_t) S * n;
   P->offsets.resize( cells + 1 );
   const size_t vb = dtypeBytes( M.dtype );
   const size_t ab = M.dtype == TNLHIP_F64 ? 8 : 4;
   P->partial.resize( cells * (int64_t) ab );
   if( n > 0 ) {
      DeviceArray< int32_t > counts( cells );
      auto count = [ & ]( auto tagS ) {
         constexpr int SS = decltype( tagS )::value;
         stripeCountKernel< SS ><<< gridFor( n, 256 ), 256, 0, st >>>( n, M.offsets.ptr, M.columns.ptr, P->width,
                                                                       counts.ptr );
      };
      auto scatter = [ & ]( auto tagS, auto* tagV ) {
         constexpr int SS = decltype( tagS )::value;
         using V = std::remove_pointer_t< decltype( tagV ) >;
         stripeScatterKernel< SS, V ><<< gridFor( n, 256 ), 256, 0, st >>>(
            n, M.offsets.ptr, M.columns.ptr, reinterpret_cast< const V* >( M.values.ptr ), P->width, P->offsets.ptr,
            P->columns.ptr, reinterpret_cast< V* >( P->values.ptr ) );
      };
      auto withS = [ & ]( auto&& f ) {
         switch( S ) {
            case 2: f( std::integral_constant< int, 2 >() ); break;
            case 4: f( std::integral_constant< int, 4 >() ); break;
            case 8: f( std::integral_constant< int, 8 >() ); break;
            case 16: f( std::integral_constant< int, 16 >() ); break;
            default: f( std::integral_constant< int, 32 >() ); break;
         }
      };
      withS( count );
      TNLHIP_CHECK( hipGetLastError() );
      Workspace& ws = threadWorkspace();
      deviceExclusiveScan( cells, counts.ptr, P->offsets.ptr, P->offsets.ptr + cells, ws, st );
      int32_t total = 0;
      TNLHIP_CHECK( hipMemcpyAsync( &total, P->offsets.ptr + cells, sizeof( total ), hipMemcpyDeviceToHost, st ) );
      TNLHIP_CHECK( hipStreamSynchronize( st ) );
      P->entries = total;
      P->columns.resize( total );
      P->values.resize( (int64_t) total * (int64_t) vb );
      withValueType( M.dtype, [ & ]( auto* tagV ) { withS( [ & ]( auto tagS ) { scatter( tagS, tagV ); } ); } );
      TNLHIP_CHECK( hipGetLastError() );
      TNLHIP_CHECK( hipStreamSynchronize( st ) );
   }
   else {
      TNLHIP_CHECK( hipMemsetAsync( P->offsets.ptr, 0, sizeof( int32_t ), st ) );
      TNLHIP_CHECK( hipStreamSynchronize( st ) );
   }
   M.stripes = std::move( P );
}

template< typename V, typename A, typename Y >
void stripeLaunch( const Matrix& M, const Y* x, Keeper< A, Y > keep, int32_t first, int32_t last,
                   hipStream_t st )
{
   const StripePlan& P = *M.stripes;
   const int S = P.stripes;
   const int64_t rows = last - first;
   if( rows <= 0 )
      return;
   const int64_t blocks = stripeBlocks( (int32_t) rows );
   int64_t grid;
   if( S < 8 )
      grid = 8 * ( ( blocks + 8 / S - 1 ) / ( 8 / S ) );
   else
      grid = 8 * ( S / 8 ) * blocks;
   A* partial = reinterpret_cast< A* >( P.partial.ptr );
   const Gate gate{ keep.gate, keep.gateMode };
   constexpr int CH = kAdaptiveStreamBytesPerWave / sizeof( A );
   stripeProductKernel< CH, V, A, Y ><<< (unsigned) grid, 256, 0, st >>>(
      S, blocks, M.rows, first, last, P.offsets.ptr, P.columns.ptr, reinterpret_cast< const V* >( P.values.ptr ), x,
      partial, gate );
   stripeCombineKernel< A, Y ><<< gridFor( rows, 256 ), 256, 0, st >>>( S, M.rows, first, last, partial, keep );
   TNLHIP_CHECK( hipGetLastError() );
}

template void stripeLaunch< double, double, double >( const Matrix&, const double*, Keeper< double, double >,
                                                      int32_t, int32_t, hipStream_t );
template void stripeLaunch< float, float, float >( const Matrix&, const float*, Keeper< float, float >, int32_t,
                                                   int32_t, hipStream_t );
template void stripeLaunch< bf16, float, float >( const Matrix&, const float*, Keeper< float, float >, int32_t,
                                                  int32_t, hipStream_t );

}  // namespace tnlhip
