// The row distribution of the padded-format kernels (tnl-dev_amd/csrc/rowdist.hpp) run on the host:
// for every launch the library can make -- rows [first, last) of n, one or two rows per lane,
// 1..64 workgroups per CU, distributions 0, 1 and 3 -- every row of the range is taken by exactly
// one lane, no lane takes a row outside [first & ~1, last), and every slot the kernels load lies
// inside the Ellpack / SlicedEllpack storage.  The lane loop below follows paddedRows in spmv.hip:
// two rows of the sequence in flight (r, r + step), then the rows left one at a time; with pairs,
// rows r and r + 1 from one 8-byte load at r's slot.  No GPU is used.
#include "../../tnl-dev_amd/csrc/rowdist.hpp"

#include <cstdio>
#include <vector>

using namespace tnlhip;

static int failures = 0;

static void check( int64_t n, int64_t first, int64_t last, bool pair, int bpc, int dist, int cus )
{
   const PaddedLaunch pl = paddedLaunch( first, last, pair, bpc, cus );
   // Ellpack: slot( r, l ) = r + aligned * l; SlicedEllpack: sliceOffset + r % 32 + 32 l.  The
   // widths are irrelevant to the bounds of the row part: a slot at l is inside the storage when
   // the row's base slot and its pair partner are inside the first column of the layout.
   const int64_t aligned = ( n + 31 ) / 32 * 32;
   std::vector< uint8_t > seen( n, 0 );
   const int P = pair ? 2 : 1;
   bool bad = false;
   for( int64_t b = 0; b < pl.grid && ! bad; b++ )
      for( int64_t t = 0; t < 256 && ! bad; t++ ) {
         const LaunchIndex li{ b, pl.grid, 256, t };
         const RowSeq s = P == 2 ? rowSequenceAt< 2 >( pl.base, last, pl.per, dist, li )
                                 : rowSequenceAt< 1 >( pl.base, last, pl.per, dist, li );
         auto visit = [ & ]( int64_t r, bool keep ) {
            // the loads: slot of row r (and r + 1 with pairs) in the first layout column
            if( r < 0 || r >= aligned || ( pair && ( r % 2 != 0 || r + 1 >= aligned ) ) ) {
               std::printf( "FAIL n=%lld [%lld,%lld) pair=%d bpc=%d dist=%d: load of row %lld\n", (long long) n,
                            (long long) first, (long long) last, (int) pair, bpc, dist, (long long) r );
               bad = true;
               return;
            }
            if( keep && r >= first && r < last )
               seen[ r ]++;
         };
         if( ! pair ) {
            int64_t r = s.start;
            for( ; r + s.step < s.end; r += 2 * s.step ) {
               visit( r, true );
               visit( r + s.step, true );
            }
            for( ; r < s.end; r += s.step )
               visit( r, true );
         }
         else
            for( int64_t r = s.start; r < s.end; r += s.step ) {
               visit( r, true );
               if( r + 1 < s.end )
                  seen[ r + 1 ] += ( r + 1 >= first && r + 1 < last ) ? 1 : 0;
            }
      }
   for( int64_t r = first; r < last && ! bad; r++ )
      if( seen[ r ] != 1 ) {
         std::printf( "FAIL n=%lld [%lld,%lld) pair=%d bpc=%d dist=%d: row %lld taken %d times\n", (long long) n,
                      (long long) first, (long long) last, (int) pair, bpc, dist, (long long) r, (int) seen[ r ] );
         bad = true;
      }
   for( int64_t r = 0; r < n && ! bad; r++ )
      if( ( r < first || r >= last ) && seen[ r ] ) {
         std::printf( "FAIL: row %lld outside the range written\n", (long long) r );
         bad = true;
      }
   failures += bad;
}

int main()
{
   const int64_t sizes[] = { 1, 2, 31, 32, 33, 255, 257, 1000, 4097, 65537, 262145, 4096000 };
   long long configs = 0;
   for( int64_t n : sizes )
      for( int range = 0; range < 3; range++ ) {
         const int64_t first = range == 0 ? 0 : range == 1 ? ( n > 1 ? 1 : 0 ) : n / 3;
         const int64_t last = range == 2 ? n - n / 5 : n;
         if( last <= first )
            continue;
         for( int pair = 0; pair < 2; pair++ )
            for( int bpc = 1; bpc <= 64; bpc += ( n > 100000 ? 7 : 1 ) )
               for( int dist : { 0, 1, 3 } )
                  for( int cus : { 256, 80 } ) {
                     check( n, first, last, pair != 0, bpc, dist, cus );
                     configs++;
                  }
      }
   if( failures ) {
      std::printf( "%d FAILURES in %lld launches\n", failures, configs );
      return 1;
   }
   std::printf( "ALL PASSED (%lld launches)\n", configs );
   return 0;
}
