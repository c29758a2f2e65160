// TNL-style C++ user code on Devices::Hip, compiled with hipcc: the row traversals with device
// lambdas (reduceAllRows, forAllElements, forAllRows) and the row queries of
// src/UnitTests/Matrices/SparseMatrixTest.hpp:241-297 (getCompressedRowLengths) and
// src/UnitTests/Algorithms/Segments/SegmentsTest.hpp:101-150 (maximum in each segment), for every
// segments type; a solver that sees its matrix change after setMatrix (setElement, operator=); the
// solver monitor and the residual history file.  Prints one line per failure; exits non-zero on
// any failure.
#include <TNL/Matrices/SparseMatrix.h>
#include <TNL/Solvers/Linear/CG.h>
#include <TNL/Solvers/Linear/GMRES.h>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

using namespace TNL;
namespace Seg = TNL::Algorithms::Segments;

static int failures = 0;
#define EXPECT_EQ( a, b )                                                                                    \
   do {                                                                                                      \
      if( !( ( a ) == ( b ) ) ) {                                                                            \
         std::printf( "FAIL %s:%d: %s != %s (%g vs %g)\n", __FILE__, __LINE__, #a, #b, (double) ( a ),       \
                      (double) ( b ) );                                                                      \
         failures++;                                                                                         \
      }                                                                                                      \
   } while( 0 )

template< typename Matrix >
void test_SetRowCapacities()
{
   // SparseMatrixTest.hpp:241-297
   const int rows = 10, cols = 11;
   Matrix m( rows, cols );
   std::vector< int > rowLengths( rows, 3 );
   int rowLength = 1;
   for( int i = 2; i < rows; i++ )
      rowLengths[ i ] = rowLength++;
   m.setRowCapacities( rowLengths );
   double value = 1;
   const int fill[] = { 3, 3, 1, 2, 3, 4, 5, 6, 7, 8 };
   for( int r = 0; r < rows; r++ )
      for( int i = 0; i < fill[ r ]; i++ )
         m.setElement( r, i, value++ );
   Containers::Vector< int, Devices::Hip, int > lengths;
   m.getCompressedRowLengths( lengths );
   for( int r = 0; r < rows; r++ )
      EXPECT_EQ( lengths.getElement( r ), fill[ r ] );
   m.getRowCapacities( lengths );
   for( int r = 0; r < rows; r++ )
      EXPECT_EQ( lengths.getElement( r ), fill[ r ] );
   for( int r = 0; r < rows; r++ )
      EXPECT_EQ( m.getRowCapacity( r ) >= fill[ r ], true );
   const auto row = m.getRow( 9 );
   int stored = 0;
   for( int l = 0; l < row.getSize(); l++ )
      if( row.getColumnIndex( l ) >= 0 ) {
         EXPECT_EQ( row.getColumnIndex( l ), stored );
         stored++;
      }
   EXPECT_EQ( stored, 8 );
}

template< typename Matrix >
void test_MaximumInRows()
{
   // SegmentsTest.hpp:101-150 with a matrix: 20 rows of 5, element (i, l) = 5 i + l + 1
   const int n = 20, size = 5;
   Matrix m( n, size );
   m.setRowCapacities( std::vector< int >( n, size ) );
   for( int i = 0; i < n; i++ )
      for( int l = 0; l < size; l++ )
         m.setElement( i, l, i * size + l + 1 );
   Containers::Vector< int, Devices::Hip, int > result( n, 0 );
   int* out = result.getData();
   auto fetch = [] __device__( int, int, const double& v ) -> int { return (int) v; };
   auto reduce = [] __device__( const int& a, const int& b ) -> int { return a > b ? a : b; };
   auto keep = [ = ] __device__( int row, const int& v ) mutable { out[ row ] = v; };
   m.reduceAllRows( fetch, reduce, keep, std::numeric_limits< int >::min() );
   for( int i = 0; i < n; i++ )
      EXPECT_EQ( result.getElement( i ), ( i + 1 ) * size );
}

template< typename Matrix >
void test_ForElementsAndRows()
{
   using Vector = Containers::Vector< double, Devices::Hip, int >;
   const int n = 37;
   Matrix m( n, n );
   std::vector< int > caps( n );
   for( int i = 0; i < n; i++ )
      caps[ i ] = 1 + i % 4;
   m.setRowCapacities( caps );
   for( int i = 0; i < n; i++ )
      for( int k = 0; k < caps[ i ]; k++ )
         m.setElement( i, ( i + 3 * k ) % n, 1.0 + k );
   Vector x( n, 1.0 ), y( n, 0.0 );
   m.vectorProduct( x, y );
   std::vector< double > before = y.toHost();
   // forAllElements: scale every stored value by 2 (padding slots are visited too: skip them)
   m.forAllElements( [] __device__( int, int, int& column, double& value ) mutable {
      if( column != -1 )
         value *= 2.0;
   } );
   m.vectorProduct( x, y );
   std::vector< double > after = y.toHost();
   for( int i = 0; i < n; i++ )
      EXPECT_EQ( after[ i ], 2.0 * before[ i ] );
   // forAllRows: the first stored entry of each row becomes 100 + row
   m.forAllRows( [] __device__( auto& row ) mutable {
      if( row.getSize() > 0 && row.getColumnIndex( 0 ) != -1 )
         row.setValue( 0, 100.0 + row.getRowIndex() );
   } );
   for( int i = 0; i < n; i += 9 )
      EXPECT_EQ( m.getRow( i ).getValue( 0 ), 100.0 + i );
}

template< template< typename, typename, typename > class S >
void runFormat( const char* name )
{
   using M = Matrices::SparseMatrix< double, Devices::Hip, int, Matrices::GeneralMatrix, S >;
   const int before = failures;
   test_SetRowCapacities< M >();
   test_MaximumInRows< M >();
   test_ForElementsAndRows< M >();
   std::printf( "%s %s\n", failures == before ? "ok  " : "FAIL", name );
}

// 1D Laplacian (2 on the diagonal, -1 beside it) built element by element
template< typename M >
void laplacian1d( M& A, int n, double diagonal )
{
   A.setDimensions( n, n );
   A.setRowCapacities( std::vector< int >( n, 3 ) );
   for( int i = 0; i < n; i++ ) {
      if( i > 0 )
         A.setElement( i, i - 1, -1.0 );
      A.setElement( i, i, diagonal );
      if( i + 1 < n )
         A.setElement( i, i + 1, -1.0 );
   }
}

template< typename M, typename V >
double relativeResidual( const M& A, const V& b, const V& x )
{
   V r( b.getSize(), 0.0 );
   A.vectorProduct( x, r );
   const auto hr = r.toHost(), hb = b.toHost();
   double num = 0, den = 0;
   for( size_t i = 0; i < hr.size(); i++ ) {
      num += ( hb[ i ] - hr[ i ] ) * ( hb[ i ] - hr[ i ] );
      den += hb[ i ] * hb[ i ];
   }
   return std::sqrt( num / den );
}

void test_SolverSeesMatrixChanges()
{
   using M = Matrices::SparseMatrix< double, Devices::Hip, int >;
   using V = Containers::Vector< double, Devices::Hip, int >;
   const int n = 200;
   auto A = std::make_shared< M >();
   laplacian1d( *A, n, 2.5 );
   Solvers::Linear::CG< M > cg;
   cg.setMatrix( A );
   cg.setConvergenceResidue( 1e-12 );
   V b( n, 1.0 ), x( n, 0.0 );
   EXPECT_EQ( cg.solve( b, x ), true );
   EXPECT_EQ( relativeResidual( *A, b, x ) < 1e-10, true );
   // setElement after setMatrix: the next solve uses the changed matrix (LinearSolver.h:127-130
   // keeps the pointer, not a copy)
   for( int i = 0; i < n; i++ )
      A->setElement( i, i, 4.0 );
   x.setValue( 0.0 );
   EXPECT_EQ( cg.solve( b, x ), true );
   EXPECT_EQ( relativeResidual( *A, b, x ) < 1e-10, true );
   // operator= after setMatrix
   M other;
   laplacian1d( other, n, 3.0 );
   *A = other;
   x.setValue( 0.0 );
   EXPECT_EQ( cg.solve( b, x ), true );
   EXPECT_EQ( relativeResidual( *A, b, x ) < 1e-10, true );
   std::printf( "%s solver re-binds a changed matrix\n", failures ? "FAIL" : "ok  " );
}

void test_MonitorAndHistoryFile()
{
   using M = Matrices::SparseMatrix< double, Devices::Hip, int >;
   using V = Containers::Vector< double, Devices::Hip, int >;
   const int before = failures;
   const int n = 400;
   auto A = std::make_shared< M >();
   laplacian1d( *A, n, 2.01 );
   Solvers::Linear::GMRES< M > gmres;
   Solvers::IterativeSolverMonitor< double, int > monitor;
   monitor.setVerbose( 0 );
   gmres.setSolverMonitor( monitor );
   gmres.setMatrix( A );
   Config::ParameterContainer p;
   const std::string path = "/tmp/tnlhip_residual_history.txt";
   p.addParameter( "convergence-residue", "1e-10" );
   p.addParameter( "residual-history-file", path );
   p.addParameter( "refresh-rate", "100" );
   gmres.setup( p );
   V b( n, 1.0 ), x( n, 0.0 );
   {
      Solvers::SolverMonitorThread thread( monitor );
      EXPECT_EQ( gmres.solve( b, x ), true );
   }
   EXPECT_EQ( monitor.getUpdates() > 0, true );
   EXPECT_EQ( monitor.getIterations(), gmres.getIterations() );
   EXPECT_EQ( monitor.getResidue(), gmres.getResidue() );
   EXPECT_EQ( monitor.getRefreshRate(), 100 );
   // the file holds one line per setResidue: a blank line at iteration 0, "<it>\t<residue>"
   const auto history = gmres.getResidualHistory();
   std::ifstream f( path );
   std::string line;
   size_t k = 0;
   int blank = 0;
   while( std::getline( f, line ) ) {
      if( line.empty() ) {
         blank++;
         continue;
      }
      int it = -1;
      double r = 0;
      std::istringstream( line ) >> it >> r;
      if( k < history.size() )
         EXPECT_EQ( std::fabs( r - history[ k ] ) <= 1e-6 * std::fabs( history[ k ] ), true );
      k++;
   }
   EXPECT_EQ( k, history.size() );
   EXPECT_EQ( blank >= 1, true );
   std::printf( "%s monitor and residual history file\n", failures == before ? "ok  " : "FAIL" );
}

int main()
{
   runFormat< Seg::CSRScalar >( "CSRScalar" );
   runFormat< Seg::CSRAdaptive >( "CSRAdaptive" );
   runFormat< Seg::Ellpack >( "Ellpack" );
   runFormat< Seg::SlicedEllpack >( "SlicedEllpack" );
   runFormat< Seg::ChunkedEllpack >( "ChunkedEllpack" );
   runFormat< Seg::BiEllpack >( "BiEllpack" );
   test_SolverSeesMatrixChanges();
   test_MonitorAndHistoryFile();
   if( failures ) {
      std::printf( "%d FAILURES\n", failures );
      return 1;
   }
   std::printf( "ALL PASSED\n" );
   return 0;
}
