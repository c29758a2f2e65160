// ILU(0) preconditioner: Preconditioners/ILU0.hpp:17-137 (Host) and
// Utils/TriangularSolve.h:25-111, on the device.
//
// Symbolic part (host, once per matrix): row i of A is read in storage order (stored zeros
// included); TNL's counting rule gives nl entries with column < i and nu with column in [i, N),
// padding skipped, and the first nl stored entries become row i of L, the next nu row i of U
// (ILU0.hpp:36-90).  Rows whose split does not put every L column in [0, i) and every U column
// in [i, N), or whose U row is empty (no U_ii to divide by), are refused.
//
// Numeric factorization and the two triangular solves run on the device in level order: row i
// depends on the rows named by its L columns (factorization and forward solve) or by its U
// columns above the diagonal (backward solve), so all rows of one level are independent.  Every
// row is processed by one lane with TNL's operation order -- L_ik = L_ik / U_kk, then
// L_ij -= L_ik U_kj and U_ij -= L_ik U_kj (U_kj = 0 when row k of U has no column j); forward
// x_i = b_i - sum L_ij x_j; backward x_i = (x_i - sum_{c >= 1} U_ij x_j) / U_i[0] -- each
// product rounded before the subtraction, so the factors and the solves equal TNL's Host results
// bit for bit.  A level with more than kNarrowLevel rows is one launch over the device; runs of
// consecutive smaller levels are one launch of a single 1024-lane workgroup with a barrier
// between levels.
#include "ilu.hpp"

#include <algorithm>

namespace tnlhip {

namespace {

constexpr int kIluBlock = 256;
constexpr int kRunBlock = 1024;          // one workgroup of 16 waves for runs of small levels
constexpr int32_t kNarrowLevel = 1024;   // a level of at most this many rows: one row per lane

template< typename T >
struct IluView
{
   const int32_t* Lrp;
   const int32_t* Lc;
   T* Lv;
   const int32_t* Urp;
   const int32_t* Uc;
   T* Uv;
};

// U.getElement( k, j ): the first stored entry of row k with column j, or 0
template< typename T >
__device__ __forceinline__ T uElement( const IluView< T >& f, int32_t k, int32_t j )
{
   for( int32_t c = f.Urp[ k ]; c < f.Urp[ k + 1 ]; c++ )
      if( f.Uc[ c ] == j )
         return f.Uv[ c ];
   return (T) 0;
}

enum Stage
{
   FACTOR = 0,
   LOWER = 1,
   UPPER = 2
};

template< int STAGE, typename T >
__device__ __forceinline__ void processRow( const IluView< T >& f, int32_t i, const T* b, T* x )
{
   if constexpr( STAGE == FACTOR ) {
      const int32_t le = f.Lrp[ i + 1 ];
      for( int32_t ck = f.Lrp[ i ]; ck < le; ck++ ) {
         const int32_t k = f.Lc[ ck ];
         const T lik = f.Lv[ ck ] / uElement( f, k, k );
         f.Lv[ ck ] = lik;
         for( int32_t cj = ck + 1; cj < le; cj++ )
            f.Lv[ cj ] = mulSubExact( f.Lv[ cj ], lik, uElement( f, k, f.Lc[ cj ] ) );
         for( int32_t cj = f.Urp[ i ]; cj < f.Urp[ i + 1 ]; cj++ )
            f.Uv[ cj ] = mulSubExact( f.Uv[ cj ], lik, uElement( f, k, f.Uc[ cj ] ) );
      }
   }
   else if constexpr( STAGE == LOWER ) {
      T xi = b[ i ];
      for( int32_t c = f.Lrp[ i ]; c < f.Lrp[ i + 1 ]; c++ )
         xi = mulSubExact( xi, f.Lv[ c ], x[ f.Lc[ c ] ] );
      x[ i ] = xi;
   }
   else {
      const int32_t first = f.Urp[ i ];
      T xi = x[ i ];
      for( int32_t c = first + 1; c < f.Urp[ i + 1 ]; c++ )
         xi = mulSubExact( xi, f.Uv[ c ], x[ f.Uc[ c ] ] );
      x[ i ] = xi / f.Uv[ first ];
   }
}

// one level over the whole device
template< int STAGE, typename T >
__global__ __launch_bounds__( kIluBlock ) void levelKernel( IluView< T > f, const int32_t* __restrict__ order,
                                                            int32_t begin, int32_t end, const T* b, T* x, Gate gate )
{
   if( ! gate.open() )
      return;
   for( int32_t k = begin + blockIdx.x * kIluBlock + threadIdx.x; k < end; k += gridDim.x * kIluBlock )
      processRow< STAGE >( f, order[ k ], b, x );
}

// consecutive small levels in one workgroup; the barrier makes each level's results visible
// to the next (same CU)
template< int STAGE, typename T >
__global__ __launch_bounds__( kRunBlock ) void levelRunKernel( IluView< T > f, const int32_t* __restrict__ order,
                                                               const int32_t* __restrict__ levels, int32_t first,
                                                               int32_t last, const T* b, T* x, Gate gate )
{
   if( ! gate.open() )
      return;
   for( int32_t lv = first; lv < last; lv++ ) {
      const int32_t end = levels[ lv + 1 ];
      for( int32_t k = levels[ lv ] + threadIdx.x; k < end; k += kRunBlock )
         processRow< STAGE >( f, order[ k ], b, x );
      __syncthreads();
   }
}

// rows grouped by level (stable in the row index) and the launches that cover the levels
void buildSchedule( Ilu0::Schedule& s, const std::vector< int32_t >& level, int32_t n, hipStream_t st )
{
   int32_t nl = 0;
   for( int32_t l : level )
      nl = std::max( nl, l + 1 );
   s.levels.assign( (size_t) nl + 1, 0 );
   for( int32_t l : level )
      s.levels[ l + 1 ]++;
   for( int32_t l = 0; l < nl; l++ )
      s.levels[ l + 1 ] += s.levels[ l ];
   std::vector< int32_t > order( n ), fill( s.levels.begin(), s.levels.end() - 1 );
   for( int32_t i = 0; i < n; i++ )
      order[ fill[ level[ i ] ]++ ] = i;
   s.order.upload( order, st );
   s.levelsDevice.upload( s.levels, st );
   TNLHIP_CHECK( hipStreamSynchronize( st ) );  // order is local
   s.launches.clear();
   for( int32_t l = 0; l < nl; ) {
      if( s.levels[ l + 1 ] - s.levels[ l ] > kNarrowLevel ) {
         s.launches.push_back( { l, l + 1, true } );
         l++;
         continue;
      }
      int32_t e = l;
      while( e < nl && s.levels[ e + 1 ] - s.levels[ e ] <= kNarrowLevel )
         e++;
      s.launches.push_back( { l, e, false } );
      l = e;
   }
}

template< int STAGE, typename T >
void runSchedule( const Ilu0::Schedule& s, const IluView< T >& f, const T* b, T* x, Gate gate, hipStream_t st )
{
   for( const auto& L : s.launches ) {
      if( L.wide ) {
         const int32_t begin = s.levels[ L.firstLevel ], end = s.levels[ L.lastLevel ];
         levelKernel< STAGE, T ><<< gridFor( end - begin, kIluBlock ), kIluBlock, 0, st >>>( f, s.order.ptr, begin, end,
                                                                                             b, x, gate );
      }
      else
         levelRunKernel< STAGE, T ><<< 1, kRunBlock, 0, st >>>( f, s.order.ptr, s.levelsDevice.ptr, L.firstLevel,
                                                                L.lastLevel, b, x, gate );
   }
   TNLHIP_CHECK( hipGetLastError() );
}

template< typename T >
IluView< T > viewOf( const Ilu0& M )
{
   return IluView< T >{ M.Lrp.ptr, M.Lc.ptr, reinterpret_cast< T* >( M.Lv.ptr ),
                        M.Urp.ptr, M.Uc.ptr, reinterpret_cast< T* >( M.Uv.ptr ) };
}

template< typename T >
void uploadValues( DeviceArray< uint8_t >& dst, const std::vector< double >& v, hipStream_t st )
{
   std::vector< T > t( v.begin(), v.end() );
   dst.resize( (int64_t) t.size() * (int64_t) sizeof( T ) );
   if( ! t.empty() ) {
      TNLHIP_CHECK( hipMemcpyAsync( dst.ptr, t.data(), t.size() * sizeof( T ), hipMemcpyHostToDevice, st ) );
      TNLHIP_CHECK( hipStreamSynchronize( st ) );  // t lives on this frame
   }
}

}  // namespace

void Ilu0::build( const Matrix& A, int vectorDtype, hipStream_t st )
{
   TNLHIP_REQUIRE( A.rows == A.cols, "ILU(0) needs a square matrix" );
   if( A.symmetric )
      throw Error( TNLHIP_ERROR_UNSUPPORTED, "ILU(0) of a symmetric matrix (lower triangle stored) is not supported" );
   TNLHIP_REQUIRE( vectorDtype == TNLHIP_F64 || vectorDtype == TNLHIP_F32, "ILU(0) computes in double or float" );
   n = A.rows;
   dtype = vectorDtype;
   // the rows of A in storage order, stored zeros included
   std::vector< int32_t > rp( (size_t) n + 1, 0 ), cols;
   std::vector< double > vals;
   if( n > 0 ) {
      DeviceArray< int32_t > dcaps( n );
      capacitiesOf( A, dcaps.ptr, st );
      const std::vector< int32_t > caps = dcaps.download( st );
      for( int32_t i = 0; i < n; i++ )
         rp[ i + 1 ] = rp[ i ] + caps[ i ];
      DeviceArray< int32_t > drp, dcol( rp[ n ] );
      DeviceArray< double > dval( rp[ n ] );
      drp.upload( rp, st );
      gatherToCsr( A, drp.ptr, dcol.ptr, dval.ptr, st, true );
      cols = dcol.download( st );
      vals = dval.download( st );
   }
   // L / U split with TNL's counting rule (ILU0.hpp:36-50) and copy of the first entries
   std::vector< int32_t > lrp( (size_t) n + 1, 0 ), urp( (size_t) n + 1, 0 ), lc, uc;
   std::vector< double > lv, uv;
   for( int32_t i = 0; i < n; i++ ) {
      int32_t nl = 0, nu = 0;
      for( int32_t g = rp[ i ]; g < rp[ i + 1 ]; g++ ) {
         const int32_t c = cols[ g ];
         if( c < 0 )
            continue;
         if( c < i )
            nl++;
         else if( c < n )
            nu++;
         else
            break;
      }
      int32_t g = rp[ i ];
      while( g < rp[ i + 1 ] && cols[ g ] < 0 )
         g++;
      for( int32_t k = 0; k < nl; k++, g++ ) {
         if( cols[ g ] < 0 || cols[ g ] >= i )
            throw Error( TNLHIP_ERROR_UNSUPPORTED, "ILU(0): row " + std::to_string( i ) +
                                                      " does not store its columns below the diagonal first" );
         lc.push_back( cols[ g ] );
         lv.push_back( vals[ g ] );
      }
      for( int32_t k = 0; k < nu; k++, g++ ) {
         if( cols[ g ] < i || cols[ g ] >= n )
            throw Error( TNLHIP_ERROR_UNSUPPORTED, "ILU(0): row " + std::to_string( i ) +
                                                      " does not store its columns below the diagonal first" );
         uc.push_back( cols[ g ] );
         uv.push_back( vals[ g ] );
      }
      if( nu == 0 )
         throw Error( TNLHIP_ERROR_INVALID_ARGUMENT,
                      "ILU(0): row " + std::to_string( i ) + " has no entry on or above the diagonal" );
      lrp[ i + 1 ] = (int32_t) lc.size();
      urp[ i + 1 ] = (int32_t) uc.size();
   }
   // levels: forward / factorization over the L columns, backward over the U columns above i
   std::vector< int32_t > lo( n, 0 ), up( n, 0 );
   for( int32_t i = 0; i < n; i++ )
      for( int32_t c = lrp[ i ]; c < lrp[ i + 1 ]; c++ )
         lo[ i ] = std::max( lo[ i ], lo[ lc[ c ] ] + 1 );
   for( int32_t i = n - 1; i >= 0; i-- )
      for( int32_t c = urp[ i ]; c < urp[ i + 1 ]; c++ )
         if( uc[ c ] > i )
            up[ i ] = std::max( up[ i ], up[ uc[ c ] ] + 1 );
   buildSchedule( lower, lo, n, st );
   buildSchedule( upper, up, n, st );
   Lrp.upload( lrp, st );
   Urp.upload( urp, st );
   Lc.upload( lc, st );
   Uc.upload( uc, st );
   if( dtype == TNLHIP_F64 ) {
      uploadValues< double >( Lv, lv, st );
      uploadValues< double >( Uv, uv, st );
      runSchedule< FACTOR, double >( lower, viewOf< double >( *this ), nullptr, nullptr, Gate{ nullptr, 0 }, st );
   }
   else {
      uploadValues< float >( Lv, lv, st );
      uploadValues< float >( Uv, uv, st );
      runSchedule< FACTOR, float >( lower, viewOf< float >( *this ), nullptr, nullptr, Gate{ nullptr, 0 }, st );
   }
   TNLHIP_CHECK( hipStreamSynchronize( st ) );  // rp / cols / vals and the schedules' host vectors are local
}

void Ilu0::solve( const void* b, void* x, Gate gate, hipStream_t st ) const
{
   if( n == 0 )
      return;
   if( dtype == TNLHIP_F64 ) {
      const auto f = viewOf< double >( *this );
      runSchedule< LOWER, double >( lower, f, (const double*) b, (double*) x, gate, st );
      runSchedule< UPPER, double >( upper, f, nullptr, (double*) x, gate, st );
   }
   else {
      const auto f = viewOf< float >( *this );
      runSchedule< LOWER, float >( lower, f, (const float*) b, (float*) x, gate, st );
      runSchedule< UPPER, float >( upper, f, nullptr, (float*) x, gate, st );
   }
}

}  // namespace tnlhip
