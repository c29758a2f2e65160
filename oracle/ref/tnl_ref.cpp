/*
 * C entry points over the real TNL headers (Devices::Host), compiled from the sources
 * where they lie by oracle/Makefile into oracle/_ref/libtnlref.so.
 *
 * TEST INFRASTRUCTURE ONLY: it pins the oracle (oracle/oracle.cpp) to TNL's own results
 * and serves as the timed CPU baseline.  Nothing in the product links it.
 * The API mirrors oracle.h so tests can compare the two value by value.
 */
#include <TNL/Algorithms/Segments/BiEllpack.h>
#include <TNL/Algorithms/Segments/CSR.h>
#include <TNL/Algorithms/Segments/ChunkedEllpack.h>
#include <TNL/Algorithms/Segments/Ellpack.h>
#include <TNL/Algorithms/Segments/SlicedEllpack.h>
#include <TNL/Algorithms/Segments/Kernels/CSRAdaptiveKernel.h>
#include <TNL/Containers/Vector.h>
#include <TNL/Matrices/SparseMatrix.h>
#include <TNL/Matrices/MatrixReader.h>
#include <TNL/Matrices/MatrixWriter.h>
#include <TNL/Solvers/Linear/BICGStab.h>
#include <TNL/Solvers/Linear/BICGStabL.h>
#include <TNL/Solvers/Linear/CG.h>
#include <TNL/Solvers/Linear/GMRES.h>
#include <TNL/Solvers/Linear/TFQMR.h>
#include <TNL/Solvers/Linear/Preconditioners/Diagonal.h>
#include <TNL/Solvers/Linear/Preconditioners/ILU0.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <unistd.h>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include <omp.h>

using namespace TNL;
namespace Seg = TNL::Algorithms::Segments;
using Host = TNL::Devices::Host;
using IVec = Containers::Vector< int, Host, int >;
using DVec = Containers::Vector< double, Host, int >;
using Alloc = typename Allocators::Default< Host >::template Allocator< int >;

namespace {

// Expose the protected metadata of each segments class.
template< Seg::ElementsOrganization O >
struct PeekEllpack : Seg::Ellpack< Host, int, Alloc, O, 32 >
{
   using Base = Seg::Ellpack< Host, int, Alloc, O, 32 >;
   std::vector< int > meta() const { return { this->segmentSize, this->size, this->alignedSize }; }
};
template< Seg::ElementsOrganization O, int SliceSize = 32 >
struct PeekSliced : Seg::SlicedEllpack< Host, int, Alloc, O, SliceSize >
{
   using Base = Seg::SlicedEllpack< Host, int, Alloc, O, SliceSize >;
   using Base::sliceOffsets;
   using Base::sliceSegmentSizes;
   std::vector< int > meta() const { return { this->size, this->alignedSize }; }
};
template< Seg::ElementsOrganization O >
struct PeekChunked : Seg::ChunkedEllpack< Host, int, Alloc, O >
{
   using Base = Seg::ChunkedEllpack< Host, int, Alloc, O >;
   using Base::chunksToSegmentsMapping;
   using Base::numberOfSlices;
   using Base::rowPointers;
   using Base::rowToChunkMapping;
   using Base::rowToSliceMapping;
   using Base::slices;
   using Base::storageSize;
};
template< Seg::ElementsOrganization O >
struct PeekBi : Seg::BiEllpack< Host, int, Alloc, O, 32 >
{
   using Base = Seg::BiEllpack< Host, int, Alloc, O, 32 >;
   using Base::groupPointers;
   using Base::rowPermArray;
   using Base::storageSize;
   using Base::virtualRows;
};

template< typename V >
std::vector< int > toStd( const V& v )
{
   std::vector< int > r( v.getSize() );
   for( int i = 0; i < v.getSize(); i++ )
      r[ i ] = v[ i ];
   return r;
}

struct Handle
{
   std::map< std::string, std::vector< int > > arrays;
   long long storage = 0;
};

template< typename Peek >
void fillChunked( Handle& h, const Peek& p )
{
   h.arrays[ "rowToChunkMapping" ] = toStd( p.rowToChunkMapping );
   h.arrays[ "rowToSliceMapping" ] = toStd( p.rowToSliceMapping );
   h.arrays[ "chunksToSegmentsMapping" ] = toStd( p.chunksToSegmentsMapping );
   h.arrays[ "rowPointers" ] = toStd( p.rowPointers );
   std::vector< int > sl;
   for( int i = 0; i < p.numberOfSlices; i++ ) {
      const auto& s = p.slices[ i ];
      sl.insert( sl.end(), { s.size, s.chunkSize, s.firstSegment, s.pointer } );
   }
   h.arrays[ "slices" ] = sl;
   h.arrays[ "chunked" ] = { p.numberOfSlices, p.storageSize };
   h.storage = p.storageSize;
}

template< typename Peek >
void fillBi( Handle& h, const Peek& p )
{
   h.arrays[ "rowPermArray" ] = toStd( p.rowPermArray );
   h.arrays[ "groupPointers" ] = toStd( p.groupPointers );
   h.storage = p.storageSize;
   h.arrays[ "biellpack" ] = { p.virtualRows, (int) ( h.storage & 0xffffffff ), (int) ( h.storage >> 32 ) };
}

template< Seg::ElementsOrganization O >
Handle* build( int kind, const IVec& caps )
{
   auto* h = new Handle;
   switch( kind ) {
      case 0: {
         Seg::CSR< Host, int > s;
         s.setSegmentsSizes( caps );
         h->arrays[ "offsets" ] = toStd( s.getOffsets() );
         h->storage = s.getStorageSize();
         break;
      }
      case 1: {
         PeekEllpack< O > s;
         s.setSegmentsSizes( caps );
         h->arrays[ "ellpack" ] = s.meta();
         h->storage = s.getStorageSize();
         break;
      }
      case 2: {
         PeekSliced< O > s;
         s.setSegmentsSizes( caps );
         h->arrays[ "sliceOffsets" ] = toStd( s.sliceOffsets );
         h->arrays[ "sliceSegmentSizes" ] = toStd( s.sliceSegmentSizes );
         h->arrays[ "sliced" ] = s.meta();
         h->storage = s.getStorageSize();
         break;
      }
      case 3: {
         PeekChunked< O > s;
         s.setSegmentsSizes( caps );
         fillChunked( *h, s );
         break;
      }
      case 4: {
         PeekBi< O > s;
         s.setSegmentsSizes( caps );
         fillBi( *h, s );
         break;
      }
      default: delete h; return nullptr;
   }
   return h;
}

template< typename Matrix >
void fillCSR( Matrix& A, int n, int ncols, const int* rp, const int* cols, const double* vals )
{
   A.setDimensions( n, ncols );
   IVec caps( n );
   for( int i = 0; i < n; i++ )
      caps[ i ] = rp[ i + 1 ] - rp[ i ];
   A.setRowCapacities( caps );
   auto& c = A.getColumnIndexes();
   auto& v = A.getValues();
   const int nnz = rp[ n ];
   for( int k = 0; k < nnz; k++ ) {
      c[ k ] = cols[ k ];
      v[ k ] = vals[ k ];
   }
}

using CsrMatrix = Matrices::SparseMatrix< double, Host, int, Matrices::GeneralMatrix, Seg::CSRDefault >;

template< template< typename, typename, typename > class S >
int convertSpmv( int n, int ncols, const int* rp, const int* cols, const double* vals, const double* x, double* y,
                 double* out_vals, int* out_cols, long long storage_cap )
{
   CsrMatrix A;
   fillCSR( A, n, ncols, rp, cols, vals );
   Matrices::SparseMatrix< double, Host, int, Matrices::GeneralMatrix, S > B;
   B = A;
   const long long st = B.getValues().getSize();
   if( st > storage_cap )
      return -1;
   for( long long g = 0; g < st; g++ ) {
      out_vals[ g ] = B.getValues()[ g ];
      out_cols[ g ] = B.getColumnIndexes()[ g ];
   }
   DVec xv( ncols ), yv( n );
   for( int i = 0; i < ncols; i++ )
      xv[ i ] = x[ i ];
   for( int i = 0; i < n; i++ )
      yv[ i ] = y[ i ];
   B.vectorProduct( xv, yv );
   for( int i = 0; i < n; i++ )
      y[ i ] = yv[ i ];
   return 0;
}

template< typename D, typename I, typename A >
using EllC = Seg::Ellpack< D, I, A, Seg::ColumnMajorOrder, 32 >;
template< typename D, typename I, typename A >
using EllRow = Seg::Ellpack< D, I, A, Seg::RowMajorOrder, 32 >;
template< typename D, typename I, typename A >
using SlC = Seg::SlicedEllpack< D, I, A, Seg::ColumnMajorOrder, 32 >;
template< typename D, typename I, typename A >
using SlR = Seg::SlicedEllpack< D, I, A, Seg::RowMajorOrder, 32 >;
template< typename D, typename I, typename A >
using ChC = Seg::ChunkedEllpack< D, I, A, Seg::ColumnMajorOrder >;
template< typename D, typename I, typename A >
using ChR = Seg::ChunkedEllpack< D, I, A, Seg::RowMajorOrder >;
template< typename D, typename I, typename A >
using BiC = Seg::BiEllpack< D, I, A, Seg::ColumnMajorOrder, 32 >;
template< typename D, typename I, typename A >
using BiR = Seg::BiEllpack< D, I, A, Seg::RowMajorOrder, 32 >;

// residual history through IterativeSolver::setResidue (IterativeSolver.hpp:189-200)
template< typename Solver >
struct WithHistory : Solver
{
   void openHistory( const std::string& path )
   {
      this->residualHistoryFile.open( path );
      this->residualHistoryFile.precision( 17 );
   }
   void closeHistory() { this->residualHistoryFile.close(); }
};


// SparseMatrix< Real, Host, int, General, S >::operator=( CSR ) then save( path ): TNL's
// File/Object format (Object.hpp:17-75, SparseMatrixView.hpp:911-920).
template< typename Real, template< typename, typename, typename > class S >
int saveAs( int n, int ncols, const int* rp, const int* cols, const double* vals, const char* path )
{
   CsrMatrix A;
   fillCSR( A, n, ncols, rp, cols, vals );
   Matrices::SparseMatrix< Real, Host, int, Matrices::GeneralMatrix, S > B;
   B = A;
   B.save( std::string( path ) );
   return 0;
}

// SparseMatrix< Real, Host, int, General, S >::load( path ) then vectorProduct( x, y ).
template< typename Real, template< typename, typename, typename > class S >
int loadAs( const char* path, const double* x, double* y, int* rows, int* cols )
{
   Matrices::SparseMatrix< Real, Host, int, Matrices::GeneralMatrix, S > B;
   B.load( std::string( path ) );
   *rows = B.getRows();
   *cols = B.getColumns();
   if( x && y ) {
      Containers::Vector< Real, Host, int > xv( B.getColumns() ), yv( B.getRows() );
      for( int i = 0; i < B.getColumns(); i++ )
         xv[ i ] = (Real) x[ i ];
      yv = 0;
      B.vectorProduct( xv, yv );
      for( int i = 0; i < B.getRows(); i++ )
         y[ i ] = yv[ i ];
   }
   return 0;
}

// variant: 0..4 CSR with the Scalar, Vector, Hybrid, Light, Adaptive kernel, 5 Ellpack,
// 6 SlicedEllpack, 7 ChunkedEllpack, 8 BiEllpack; colmajor selects the organization of 5..8.
template< typename Real, typename F >
int byVariant( int variant, int colmajor, F&& f )
{
   switch( variant * 2 + ( colmajor ? 1 : 0 ) ) {
      case 0: case 1: return f.template operator()< Real, Seg::CSRScalar >();
      case 2: case 3: return f.template operator()< Real, Seg::CSRVector >();
      case 4: case 5: return f.template operator()< Real, Seg::CSRHybrid >();
      case 6: case 7: return f.template operator()< Real, Seg::CSRLight >();
      case 8: case 9: return f.template operator()< Real, Seg::CSRAdaptive >();
      case 10: return f.template operator()< Real, EllRow >();
      case 11: return f.template operator()< Real, EllC >();
      case 12: return f.template operator()< Real, SlR >();
      case 13: return f.template operator()< Real, SlC >();
      case 14: return f.template operator()< Real, ChR >();
      case 15: return f.template operator()< Real, ChC >();
      case 16: return f.template operator()< Real, BiR >();
      case 17: return f.template operator()< Real, BiC >();
   }
   return -2;
}

struct SaveFn
{
   int n, ncols;
   const int *rp, *cols;
   const double* vals;
   const char* path;
   template< typename Real, template< typename, typename, typename > class S >
   int operator()() const { return saveAs< Real, S >( n, ncols, rp, cols, vals, path ); }
};

struct LoadFn
{
   const char* path;
   const double* x;
   double* y;
   int *rows, *cols;
   template< typename Real, template< typename, typename, typename > class S >
   int operator()() const { return loadAs< Real, S >( path, x, y, rows, cols ); }
};

}  // namespace

// TNL's row traversals on a matrix of format S filled from CSR (the same CSR type as the source
// keeps its storage as given, explicit zeros and padding included; other types are assigned
// from it): reduceAllRows with the fetch / reduce pair of orc_reduce_rows, getRowCapacities,
// getCompressedRowLengths and getRowCapacity of every row.
template< template< typename, typename, typename > class S >
int rowReductions( int n, int ncols, const int* rp, const int* cols, const double* vals, int fk, int rk,
                   double identity, double* out, int* capacities, int* compressed, int* segmentSizes )
{
   CsrMatrix A;
   fillCSR( A, n, ncols, rp, cols, vals );
   Matrices::SparseMatrix< double, Host, int, Matrices::GeneralMatrix, S > B;
   B = A;
   auto run = [ & ]( const auto& M ) {
      auto fetch = [ = ]( int, int column, const double& value ) -> double
      {
         switch( fk ) {
            case 0: return value;
            case 1: return std::abs( value );
            case 2: return 1.0;
            case 3: return value != 0.0 ? 1.0 : 0.0;
            default: return (double) column;
         }
      };
      auto reduce = [ = ]( const double& a, const double& b ) -> double
      {
         return rk == 0 ? a + b : rk == 1 ? TNL::min( a, b ) : TNL::max( a, b );
      };
      auto keep = [ & ]( int row, const double& v ) mutable { out[ row ] = v; };
      M.reduceAllRows( fetch, reduce, keep, identity );
      IVec caps, lens;
      M.getRowCapacities( caps );
      M.getCompressedRowLengths( lens );
      for( int i = 0; i < n; i++ ) {
         capacities[ i ] = caps[ i ];
         compressed[ i ] = lens[ i ];
         segmentSizes[ i ] = M.getRowCapacity( i );
      }
   };
   if( std::is_same< S< Host, int, Alloc >, Seg::CSRScalar< Host, int, Alloc > >::value )
      run( A );
   else
      run( B );
   return 0;
}

template< Seg::ElementsOrganization O, int S >
Handle* buildSliced( const IVec& caps )
{
   auto* h = new Handle;
   PeekSliced< O, S > s;
   s.setSegmentsSizes( caps );
   h->arrays[ "sliceOffsets" ] = toStd( s.sliceOffsets );
   h->arrays[ "sliceSegmentSizes" ] = toStd( s.sliceSegmentSizes );
   h->arrays[ "sliced" ] = s.meta();
   h->storage = s.getStorageSize();
   return h;
}

template< Seg::ElementsOrganization O >
Handle* buildSlicedOf( int slice, const IVec& caps )
{
   switch( slice ) {
      case 1: return buildSliced< O, 1 >( caps );
      case 8: return buildSliced< O, 8 >( caps );
      case 16: return buildSliced< O, 16 >( caps );
      case 32: return buildSliced< O, 32 >( caps );
      case 64: return buildSliced< O, 64 >( caps );
      case 128: return buildSliced< O, 128 >( caps );
      case 256: return buildSliced< O, 256 >( caps );
      default: return nullptr;
   }
}

extern "C" {

// SlicedEllpack< Host, int, Allocator, Organization, SliceSize > for SliceSize 1, 8, ..., 256
void* ref_sliced_create( int n, const int* caps, int colmajor, int slice )
{
   IVec c( n );
   for( int i = 0; i < n; i++ )
      c[ i ] = caps[ i ];
   return colmajor ? (void*) buildSlicedOf< Seg::ColumnMajorOrder >( slice, c )
                   : (void*) buildSlicedOf< Seg::RowMajorOrder >( slice, c );
}

void* ref_segments_create( int kind, int n, const int* caps, int colmajor )
{
   IVec c( n );
   for( int i = 0; i < n; i++ )
      c[ i ] = caps[ i ];
   return colmajor ? (void*) build< Seg::ColumnMajorOrder >( kind, c ) : (void*) build< Seg::RowMajorOrder >( kind, c );
}

void ref_segments_destroy( void* h )
{
   delete (Handle*) h;
}

long long ref_segments_storage_size( void* h )
{
   return ( (Handle*) h )->storage;
}

long long ref_segments_get( void* hv, const char* name, int* out )
{
   auto* h = (Handle*) hv;
   auto it = h->arrays.find( name );
   if( it == h->arrays.end() )
      return -1;
   if( out )
      std::memcpy( out, it->second.data(), it->second.size() * sizeof( int ) );
   return (long long) it->second.size();
}

// SparseMatrix<Host, kind>::operator=( CSR ) then vectorProduct( x, y ); exports the storage.
int ref_convert_spmv( int kind, int colmajor, int n, int ncols, const int* rp, const int* cols, const double* vals,
                      const double* x, double* y, double* out_vals, int* out_cols, long long storage_cap )
{
   switch( kind * 2 + ( colmajor ? 1 : 0 ) ) {
      case 0:
      // CSRVector: a CSR type different from the source's, so the converting operator= runs
      // (same-type assignment is a plain copy that keeps explicit zeros)
      case 1: return convertSpmv< Seg::CSRVector >( n, ncols, rp, cols, vals, x, y, out_vals, out_cols, storage_cap );
      case 2: return convertSpmv< EllRow >( n, ncols, rp, cols, vals, x, y, out_vals, out_cols, storage_cap );
      case 3: return convertSpmv< EllC >( n, ncols, rp, cols, vals, x, y, out_vals, out_cols, storage_cap );
      case 4: return convertSpmv< SlR >( n, ncols, rp, cols, vals, x, y, out_vals, out_cols, storage_cap );
      case 5: return convertSpmv< SlC >( n, ncols, rp, cols, vals, x, y, out_vals, out_cols, storage_cap );
      case 6: return convertSpmv< ChR >( n, ncols, rp, cols, vals, x, y, out_vals, out_cols, storage_cap );
      case 7: return convertSpmv< ChC >( n, ncols, rp, cols, vals, x, y, out_vals, out_cols, storage_cap );
      case 8: return convertSpmv< BiR >( n, ncols, rp, cols, vals, x, y, out_vals, out_cols, storage_cap );
      case 9: return convertSpmv< BiC >( n, ncols, rp, cols, vals, x, y, out_vals, out_cols, storage_cap );
   }
   return -2;
}

int ref_row_reductions( int kind, int colmajor, int n, int ncols, const int* rp, const int* cols, const double* vals,
                        int fetch, int reduce, double identity, double* out, int* capacities, int* compressed,
                        int* segmentSizes )
{
   switch( kind * 2 + ( colmajor ? 1 : 0 ) ) {
      case 0:
      case 1:
         return rowReductions< Seg::CSRScalar >( n, ncols, rp, cols, vals, fetch, reduce, identity, out, capacities,
                                                 compressed, segmentSizes );
      case 2: return rowReductions< EllRow >( n, ncols, rp, cols, vals, fetch, reduce, identity, out, capacities, compressed, segmentSizes );
      case 3: return rowReductions< EllC >( n, ncols, rp, cols, vals, fetch, reduce, identity, out, capacities, compressed, segmentSizes );
      case 4: return rowReductions< SlR >( n, ncols, rp, cols, vals, fetch, reduce, identity, out, capacities, compressed, segmentSizes );
      case 5: return rowReductions< SlC >( n, ncols, rp, cols, vals, fetch, reduce, identity, out, capacities, compressed, segmentSizes );
      case 6: return rowReductions< ChR >( n, ncols, rp, cols, vals, fetch, reduce, identity, out, capacities, compressed, segmentSizes );
      case 7: return rowReductions< ChC >( n, ncols, rp, cols, vals, fetch, reduce, identity, out, capacities, compressed, segmentSizes );
      case 8: return rowReductions< BiR >( n, ncols, rp, cols, vals, fetch, reduce, identity, out, capacities, compressed, segmentSizes );
      case 9: return rowReductions< BiC >( n, ncols, rp, cols, vals, fetch, reduce, identity, out, capacities, compressed, segmentSizes );
   }
   return -2;
}

// CSRAdaptiveKernel::init, blocks for sizeof(Value) == 2^value_size_log; 6 ints per block.
long long ref_csr_adaptive_blocks( int n, const int* offsets, int value_size_log, int* out )
{
   struct PeekKernel : Seg::CSRAdaptiveKernel< int, Host >
   {
      using Seg::CSRAdaptiveKernel< int, Host >::blocksArray;
   };
   IVec off( n + 1 );
   for( int i = 0; i <= n; i++ )
      off[ i ] = offsets[ i ];
   PeekKernel k;
   k.init( off );
   const auto& b = k.blocksArray[ value_size_log ];
   if( out )
      for( int i = 0; i < b.getSize(); i++ ) {
         const auto d = b[ i ];
         const int type = (int) d.getType();
         const bool sized = type != 0 && i + 1 < b.getSize();
         out[ 6 * i + 0 ] = type;
         out[ 6 * i + 1 ] = d.getFirstSegment();
         out[ 6 * i + 2 ] = sized ? d.getSize() : 0;
         out[ 6 * i + 3 ] = sized ? d.getSegmentsInBlock() : 0;
         out[ 6 * i + 4 ] = d.getWarpIdx();
         out[ 6 * i + 5 ] = type == 0 ? d.getWarpsCount() : 0;  // set only for LONG blocks
      }
   return b.getSize();
}

// Timed Host CSR SpMV (OpenMP CSRScalarKernel): returns the median seconds per call.
double ref_time_spmv_csr( int n, int ncols, const int* rp, const int* cols, const double* vals, const double* x,
                          double* y, int loops, int threads )
{
   if( threads > 0 )
      omp_set_num_threads( threads );
   CsrMatrix A;
   fillCSR( A, n, ncols, rp, cols, vals );
   DVec xv( ncols ), yv( n );
   for( int i = 0; i < ncols; i++ )
      xv[ i ] = x[ i ];
   A.vectorProduct( xv, yv );
   std::vector< double > t;
   for( int l = 0; l < loops; l++ ) {
      const auto a = std::chrono::steady_clock::now();
      A.vectorProduct( xv, yv );
      const auto b = std::chrono::steady_clock::now();
      t.push_back( std::chrono::duration< double >( b - a ).count() );
   }
   for( int i = 0; i < n; i++ )
      y[ i ] = yv[ i ];
   std::sort( t.begin(), t.end() );
   return t.empty() ? 0.0 : t[ t.size() / 2 ];
}

typedef struct
{
   int max_iterations, min_iterations;
   double convergence_residue, divergence_residue;
   int gmres_variant, restarting_min, restarting_max, restarting_step_min, restarting_step_max;
   int bicgstab_exact_residue, preconditioner, bicgstab_ell;
} ref_solver_params;

// When set, the next ref_solve writes TNL's residual-history-file to this path and keeps it.
static std::string g_keepHistory;
void ref_keep_history_file( const char* path )
{
   g_keepHistory = path ? path : "";
}

// Solve on the Host with TNL's solvers; history via residual-history-file.  Returns
// 1 converged, 0 not, -1 GMRES breakdown; *seconds = wall time of solve().
int ref_solve( int method, int n, const int* rp, const int* cols, const double* vals, const double* b, double* x,
               const ref_solver_params* p, int* iterations, double* residue, double* history, int history_cap,
               int* history_len, int threads, double* seconds )
{
   if( threads > 0 )
      omp_set_num_threads( threads );
   auto A = std::make_shared< CsrMatrix >();
   fillCSR( *A, n, n, rp, cols, vals );
   DVec bv( n ), xv( n );
   for( int i = 0; i < n; i++ ) {
      bv[ i ] = b[ i ];
      xv[ i ] = x[ i ];
   }
   char tmp[] = "/tmp/tnlref_histXXXXXX";
   const bool keep = ! g_keepHistory.empty();
   if( ! keep ) {
      const int fd = mkstemp( tmp );
      if( fd >= 0 )
         close( fd );
   }
   const std::string pathStr = keep ? g_keepHistory : std::string( tmp );
   const char* path = pathStr.c_str();
   g_keepHistory.clear();
   using Pre = Solvers::Linear::Preconditioners::Diagonal< CsrMatrix >;
   using Ilu = Solvers::Linear::Preconditioners::ILU0< CsrMatrix >;
   auto pre = std::make_shared< Pre >();
   auto ilu = std::make_shared< Ilu >();
   int rc = 0;
   auto run = [ & ]( auto& solver )
   {
      solver.setMatrix( A );
      if( p->preconditioner == 1 ) {
         pre->update( A );
         solver.setPreconditioner( pre );
      }
      else if( p->preconditioner == 2 ) {
         ilu->update( A );
         solver.setPreconditioner( ilu );
      }
      solver.setMaxIterations( p->max_iterations );
      solver.setMinIterations( p->min_iterations );
      solver.setConvergenceResidue( p->convergence_residue );
      solver.setDivergenceResidue( p->divergence_residue );
      solver.openHistory( path );
      const auto t0 = std::chrono::steady_clock::now();
      try {
         rc = solver.solve( bv, xv ) ? 1 : 0;
      }
      catch( int ) {
         rc = -1;
      }
      const auto t1 = std::chrono::steady_clock::now();
      if( seconds )
         *seconds = std::chrono::duration< double >( t1 - t0 ).count();
      solver.closeHistory();
      *iterations = solver.getIterations();
      *residue = solver.getResidue();
   };
   if( method == 0 ) {
      WithHistory< Solvers::Linear::CG< CsrMatrix > > s;
      run( s );
   }
   else if( method == 1 ) {
      WithHistory< Solvers::Linear::BICGStab< CsrMatrix > > s;
      Config::ParameterContainer prm;
      prm.addParameter< bool >( "bicgstab-exact-residue", p->bicgstab_exact_residue != 0 );
      s.setup( prm );
      run( s );
   }
   else if( method == 3 ) {
      WithHistory< Solvers::Linear::TFQMR< CsrMatrix > > s;
      run( s );
   }
   else if( method == 4 ) {
      WithHistory< Solvers::Linear::BICGStabL< CsrMatrix > > s;
      Config::ParameterContainer prm;
      prm.addParameter< int >( "bicgstab-ell", p->bicgstab_ell );
      prm.addParameter< bool >( "bicgstab-exact-residue", p->bicgstab_exact_residue != 0 );
      s.setup( prm );
      try {
         run( s );
      }
      catch( const std::runtime_error& ) {
         rc = -1;  // initial residue NaN
      }
   }
   else {
      WithHistory< Solvers::Linear::GMRES< CsrMatrix > > s;
      Config::ParameterContainer prm;
      const char* names[] = { "CGS", "CGSR", "MGS", "MGSR", "CWY" };
      prm.addParameter< String >( "gmres-variant", names[ p->gmres_variant ] );
      prm.addParameter< int >( "gmres-restarting-min", p->restarting_min );
      prm.addParameter< int >( "gmres-restarting-max", p->restarting_max );
      prm.addParameter< int >( "gmres-restarting-step-min", p->restarting_step_min );
      prm.addParameter< int >( "gmres-restarting-step-max", p->restarting_step_max );
      s.setup( prm );
      run( s );
   }
   for( int i = 0; i < n; i++ )
      x[ i ] = xv[ i ];
   // parse "iter \t residue" lines
   std::ifstream in( path );
   int count = 0;
   long long itn;
   double val;
   while( in >> itn >> val ) {
      if( history && count < history_cap )
         history[ count ] = val;
      count++;
   }
   if( history_len )
      *history_len = count;
   if( ! keep )
      std::remove( path );
   return rc;
}


// MatrixReader::readMtx into a Host CSR matrix; returns 0, or -1 with the exception text in err.
int ref_read_mtx( const char* path, int* rows, int* cols, long long* storage, int* rp, int* ci, double* v, char* err,
                  int errlen )
{
   using M = CsrMatrix;
   try {
      M A;
      TNL::Matrices::MatrixReader< M >::readMtx( std::string( path ), A );
      *rows = A.getRows();
      *cols = A.getColumns();
      const auto& off = A.getSegments().getOffsets();
      *storage = A.getValues().getSize();
      if( rp )
         for( int i = 0; i <= A.getRows(); i++ )
            rp[ i ] = off.getElement( i );
      if( ci )
         for( long long k = 0; k < *storage; k++ ) {
            ci[ k ] = A.getColumnIndexes().getElement( k );
            v[ k ] = A.getValues().getElement( k );
         }
      return 0;
   }
   catch( const std::exception& e ) {
      std::snprintf( err, errlen, "%s", e.what() );
      return -1;
   }
}

// MatrixWriter::writeMtx of a Host CSR matrix given by its arrays.
int ref_write_mtx( const char* path, int rows, int cols, const int* rp, const int* ci, const double* v )
{
   using M = CsrMatrix;
   M A( rows, cols );
   IVec caps( rows );
   for( int i = 0; i < rows; i++ )
      caps[ i ] = rp[ i + 1 ] - rp[ i ];
   A.setRowCapacities( caps );
   for( int k = 0; k < rp[ rows ]; k++ ) {
      A.getColumnIndexes()[ k ] = ci[ k ];
      A.getValues()[ k ] = v[ k ];
   }
   TNL::Matrices::MatrixWriter< M >::writeMtx( std::string( path ), A );
   return 0;
}

// writeMtx of a device matrix read by readMtx (MatrixReader.hpp:19-26, MatrixWriter.hpp:35-40):
// the host-read matrix is assigned to the device matrix and that back to a host copy, each an
// operator= (stored entries per row as capacities, zeros dropped); TNL's Cuda build is absent, so
// the two assignments run between Host matrices of different CSR types, the same conversion code.
int ref_write_mtx_device( const char* path, int rows, int cols, const int* rp, const int* ci, const double* v )
{
   CsrMatrix A( rows, cols );
   IVec caps( rows );
   for( int i = 0; i < rows; i++ )
      caps[ i ] = rp[ i + 1 ] - rp[ i ];
   A.setRowCapacities( caps );
   for( int k = 0; k < rp[ rows ]; k++ ) {
      A.getColumnIndexes()[ k ] = ci[ k ];
      A.getValues()[ k ] = v[ k ];
   }
   Matrices::SparseMatrix< double, Host, int, Matrices::GeneralMatrix, Seg::CSRVector > D;
   D = A;
   CsrMatrix H;
   H = D;
   TNL::Matrices::MatrixWriter< CsrMatrix >::writeMtx( std::string( path ), H );
   return 0;
}

// TNL's binary save of a matrix converted from CSR (single = 1: float values).
int ref_save_tnl( int variant, int colmajor, int single, int n, int ncols, const int* rp, const int* cols,
                  const double* vals, const char* path, char* err, int errlen )
{
   SaveFn f{ n, ncols, rp, cols, vals, path };
   try {
      return single ? byVariant< float >( variant, colmajor, f ) : byVariant< double >( variant, colmajor, f );
   }
   catch( const std::exception& e ) {
      std::snprintf( err, errlen, "%s", e.what() );
      return -1;
   }
}

// TNL's binary load into the given matrix type, then y = A x when x and y are given.
int ref_load_tnl( int variant, int colmajor, int single, const char* path, const double* x, double* y, int* rows,
                  int* cols, char* err, int errlen )
{
   LoadFn f{ path, x, y, rows, cols };
   try {
      return single ? byVariant< float >( variant, colmajor, f ) : byVariant< double >( variant, colmajor, f );
   }
   catch( const std::exception& e ) {
      std::snprintf( err, errlen, "%s", e.what() );
      return -1;
   }
}

// TNL's ILU0< CSR >::update of a CSR matrix: the factors as CSR arrays (rows of U in TNL's
// reversed order, as stored), then solve( b, x ).  Pass NULL arrays to query the sizes.
int ref_ilu0( int n, const int* rp, const int* cols, const double* vals, int* nnzL, int* nnzU, int* Lrp, int* Lc,
              double* Lv, int* Urp, int* Uc, double* Uv, const double* b, double* x )
{
   struct Peek : Solvers::Linear::Preconditioners::ILU0< CsrMatrix >
   {
      using Solvers::Linear::Preconditioners::ILU0< CsrMatrix >::L;
      using Solvers::Linear::Preconditioners::ILU0< CsrMatrix >::U;
   };
   auto A = std::make_shared< CsrMatrix >();
   fillCSR( *A, n, n, rp, cols, vals );
   Peek ilu;
   ilu.update( A );
   *nnzL = ilu.L.getValues().getSize();
   *nnzU = ilu.U.getValues().getSize();
   auto dump = []( const auto& M, int* p, int* c, double* v ) {
      if( ! p )
         return;
      const auto& off = M.getSegments().getOffsets();
      for( int i = 0; i <= M.getRows(); i++ )
         p[ i ] = off[ i ];
      for( int k = 0; k < M.getValues().getSize(); k++ ) {
         c[ k ] = M.getColumnIndexes()[ k ];
         v[ k ] = M.getValues()[ k ];
      }
   };
   dump( ilu.L, Lrp, Lc, Lv );
   dump( ilu.U, Urp, Uc, Uv );
   if( b && x ) {
      DVec bv( n ), xv( n );
      for( int i = 0; i < n; i++ )
         bv[ i ] = b[ i ];
      ilu.solve( bv.getConstView(), xv.getView() );
      for( int i = 0; i < n; i++ )
         x[ i ] = xv[ i ];
   }
   return 0;
}

// SparseMatrix< double, Host, int, SymmetricMatrix, CSR >::vectorProduct on one thread, the
// matrix given by its lower triangle in CSR.
int ref_symmetric_spmv( int n, const int* rp, const int* cols, const double* vals, const double* x, double* y,
                        double alpha, double beta, int first, int last )
{
   omp_set_num_threads( 1 );
   Matrices::SparseMatrix< double, Host, int, Matrices::SymmetricMatrix, Seg::CSRDefault > A;
   fillCSR( A, n, n, rp, cols, vals );
   DVec xv( n ), yv( n );
   for( int i = 0; i < n; i++ ) {
      xv[ i ] = x[ i ];
      yv[ i ] = y[ i ];
   }
   A.vectorProduct( xv, yv, alpha, beta, first, last );
   for( int i = 0; i < n; i++ )
      y[ i ] = yv[ i ];
   return 0;
}

// SparseMatrix< bool, Host, int, General, CSR, int >::vectorProduct (a binary matrix), one thread.
int ref_binary_spmv( int n, int ncols, const int* rp, const int* cols, const int* x, int* y, int alpha, int beta,
                     int first, int last )
{
   omp_set_num_threads( 1 );
   Matrices::SparseMatrix< bool, Host, int, Matrices::GeneralMatrix, Seg::CSRDefault, int > A( n, ncols );
   IVec caps( n );
   for( int i = 0; i < n; i++ )
      caps[ i ] = rp[ i + 1 ] - rp[ i ];
   A.setRowCapacities( caps );
   for( int i = 0; i < n; i++ )
      for( int k = rp[ i ]; k < rp[ i + 1 ]; k++ )
         A.setElement( i, cols[ k ], true );
   IVec xv( ncols ), yv( n );
   for( int i = 0; i < ncols; i++ )
      xv[ i ] = x[ i ];
   for( int i = 0; i < n; i++ )
      yv[ i ] = y[ i ];
   A.vectorProduct( xv, yv, alpha, beta, first, last );
   for( int i = 0; i < n; i++ )
      y[ i ] = yv[ i ];
   return 0;
}

}  // extern "C"
