// Autotuner: picks the segment format, the slice size of SlicedEllpack and kernel launch
// parameters (lanes per row for the CSR light / hybrid kernels, rows per lane and workgroups
// per CU for the padded kernels) for one
// matrix by timing y = A x on the device.  No counterpart in the reference: TNL fixes the
// format at compile time (SparseMatrix's Segments template parameter) and the benchmark
// (Benchmarks/SpMV/spmv.h:269-345) times each format separately.
#include "common.hpp"

#include <cstring>
#include <vector>

namespace {

struct Candidate
{
   int format;
   int lanes;      // 0: kernel default
   int blocksPerCu;  // 0: default
   int pairRows;   // -1: default
   int sliceSize;  // SlicedEllpack: 16, 32 (TNL's default), 64; 0 for other formats
   int stripes;    // CSRAdaptive: column stripes of the striped plan (0: none)
};

std::vector< Candidate > candidatesFor( int format, bool xExceedsL2 )
{
   std::vector< Candidate > c;
   switch( format ) {
      case TNLHIP_CSR_LIGHT:
      case TNLHIP_CSR_HYBRID:
         c.push_back( { format, 0, 0, -1, 0, 0 } );
         for( int l : { 2, 4, 8, 16, 32 } )
            c.push_back( { format, l, 0, -1, 0, 0 } );
         break;
      case TNLHIP_ELLPACK:
         for( int p : { 0, 1 } )
            for( int b : { 8, 16 } )
               c.push_back( { format, 0, b, p, 0, 0 } );
         break;
      case TNLHIP_SLICED_ELLPACK:
         // slice sizes outermost: a change of slice size re-lays the storage out
         for( int s : { 32, 16, 64 } )
            for( int p : { 0, 1 } )
               for( int b : { 8, 16 } )
                  c.push_back( { format, 0, b, p, s, 0 } );
         break;
      case TNLHIP_CSR_ADAPTIVE:
         c.push_back( { format, 0, 0, -1, 0, 0 } );
         // the striped plan keeps each stripe of x in the L2 of the XCDs that own it: worth
         // timing only when x does not fit one XCD's L2 already
         if( xExceedsL2 )
            for( int s : { 4, 8 } )
               c.push_back( { format, 0, 0, -1, 0, s } );
         break;
      default:
         c.push_back( { format, 0, 0, -1, 0, 0 } );
   }
   return c;
}

// Every knob of the candidate, defaults included, so a matrix configured for one candidate
// carries nothing over from the one timed before it.
void configure( tnlhip_matrix A, const Candidate& c )
{
   const bool csrGroups = c.format == TNLHIP_CSR_LIGHT || c.format == TNLHIP_CSR_HYBRID ||
                          c.format == TNLHIP_CSR_VECTOR || c.format == TNLHIP_CSR_ADAPTIVE;
   if( csrGroups )
      TNLHIP_CHECK_STATUS( tnlhip_matrix_set_param( A, "lanes_per_row", c.lanes ) );
   TNLHIP_CHECK_STATUS( tnlhip_matrix_set_param( A, "blocks_per_cu", c.blocksPerCu > 0 ? c.blocksPerCu : 16 ) );
   TNLHIP_CHECK_STATUS( tnlhip_matrix_set_param( A, "pair_rows", c.pairRows > 0 ? 1 : 0 ) );
   if( c.format == TNLHIP_SLICED_ELLPACK )
      TNLHIP_CHECK_STATUS( tnlhip_matrix_set_param( A, "slice_size", c.sliceSize > 0 ? c.sliceSize : 32 ) );
   if( c.format == TNLHIP_CSR_ADAPTIVE )
      TNLHIP_CHECK_STATUS( tnlhip_matrix_set_param( A, "column_stripes", c.stripes ) );
}

}  // namespace

using namespace tnlhip;

extern "C" {

int tnlhip_autotune_csr( int value_dtype, int32_t rows, int32_t columns, const int32_t* row_pointers,
                         const int32_t* column_indexes, const double* values, int32_t host_pointers,
                         uint32_t format_mask, int32_t repetitions, tnlhip_tuning* best, tnlhip_tuning* all,
                         int32_t all_capacity, int32_t* all_count, tnlhip_matrix* best_matrix )
{
   tnlhip_matrix keep = nullptr;
   hipStream_t st = nullptr;
   hipEvent_t e0 = nullptr, e1 = nullptr;
   void* x = nullptr;
   void* y = nullptr;
   const int status = guarded( [ & ] {
      TNLHIP_REQUIRE( best && rows >= 0 && columns >= 0 && repetitions > 0, "bad argument" );
      TNLHIP_REQUIRE( value_dtype == TNLHIP_F64 || value_dtype == TNLHIP_F32 || value_dtype == TNLHIP_BF16,
                      "unknown value type" );
      if( all_count )
         *all_count = 0;
      if( best_matrix )
         *best_matrix = nullptr;
      std::memset( best, 0, sizeof( *best ) );
      best->seconds = -1;
      const size_t vb = value_dtype == TNLHIP_F64 ? 8 : 4;
      TNLHIP_CHECK( hipStreamCreate( &st ) );
      TNLHIP_CHECK( hipEventCreate( &e0 ) );
      TNLHIP_CHECK( hipEventCreate( &e1 ) );
      TNLHIP_CHECK( hipMalloc( &x, vb * (size_t) std::max( columns, 1 ) ) );
      TNLHIP_CHECK( hipMalloc( &y, vb * (size_t) std::max( rows, 1 ) ) );
      TNLHIP_CHECK( hipMemsetAsync( x, 0, vb * (size_t) std::max( columns, 1 ), st ) );
      int count = 0;
      for( int format = 0; format <= TNLHIP_BIELLPACK; format++ ) {
         if( ! ( format_mask & ( 1u << format ) ) )
            continue;
         tnlhip_matrix A = nullptr;
         int s = tnlhip_matrix_create_csr( format, value_dtype, rows, columns, row_pointers, column_indexes, values,
                                           host_pointers, &A );
         if( s != TNLHIP_SUCCESS )
            throw Error( s, std::string( "autotune: building a candidate failed: " ) + tnlhip_last_error() );
         bool kept = false;
         const bool xExceedsL2 = (int64_t) columns * (int64_t) vb > ( 4ll << 20 );
         for( const Candidate& c : candidatesFor( format, xExceedsL2 ) ) {
            configure( A, c );
            for( int w = 0; w < 2; w++ )
               TNLHIP_CHECK_STATUS( tnlhip_spmv( A, x, y, 1.0, 0.0, 0, 0, st ) );
            TNLHIP_CHECK( hipEventRecord( e0, st ) );
            for( int r = 0; r < repetitions; r++ )
               TNLHIP_CHECK_STATUS( tnlhip_spmv( A, x, y, 1.0, 0.0, 0, 0, st ) );
            TNLHIP_CHECK( hipEventRecord( e1, st ) );
            TNLHIP_CHECK( hipEventSynchronize( e1 ) );
            float ms = 0;
            TNLHIP_CHECK( hipEventElapsedTime( &ms, e0, e1 ) );
            tnlhip_tuning t{ format, c.lanes, c.blocksPerCu, c.pairRows, c.sliceSize, c.stripes, 1e-3 * ms / repetitions };
            if( all && count < all_capacity )
               all[ count ] = t;
            count++;
            if( best->seconds < 0 || t.seconds < best->seconds ) {
               *best = t;
               if( best_matrix ) {
                  if( keep && keep != A )
                     tnlhip_matrix_destroy( keep );
                  keep = A;
                  kept = true;
               }
            }
         }
         if( kept )
            // leave the kept matrix configured exactly as its best candidate (format defaults
            // restored for the knobs the candidate leaves at their default)
            configure( A, Candidate{ best->format, best->lanes, best->blocks_per_cu, best->pair_rows, best->slice_size,
                                     best->column_stripes } );
         else
            tnlhip_matrix_destroy( A );
      }
      TNLHIP_REQUIRE( count > 0, "no format selected in format_mask" );
      if( all_count )
         *all_count = count;
      if( best_matrix ) {
         *best_matrix = keep;
         keep = nullptr;
      }
   } );
   if( keep )
      tnlhip_matrix_destroy( keep );
   if( x )
      (void) hipFree( x );
   if( y )
      (void) hipFree( y );
   if( e0 )
      (void) hipEventDestroy( e0 );
   if( e1 )
      (void) hipEventDestroy( e1 );
   if( st )
      (void) hipStreamDestroy( st );
   return status;
}

}  // extern "C"
