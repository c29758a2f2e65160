// TNL-style C++ user code on Devices::Hip: the SpMV unit tests of
// src/UnitTests/Matrices/SparseMatrixVectorProductTest.hpp (smallMatrix1, mediumSizeMatrix1,
// largeMatrix, longRowsMatrix) written against the TNL class names, for every segments type,
// and a CG / GMRES solve through LinearSolver::setMatrix / solve.  Prints one line per check;
// exits non-zero on the first failure.
#include <TNL/Matrices/SparseMatrix.h>
#include <TNL/Matrices/MatrixReader.h>
#include <TNL/Matrices/MatrixWriter.h>
#include <TNL/Solvers/Linear/CG.h>
#include <TNL/Solvers/Linear/GMRES.h>
#include <TNL/Solvers/LinearSolverTypeResolver.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

using namespace TNL;
namespace Seg = TNL::Algorithms::Segments;

static int failures = 0;
#define EXPECT_EQ( a, b )                                                                           \
   do {                                                                                             \
      if( !( ( a ) == ( b ) ) ) {                                                                   \
         std::printf( "FAIL %s:%d: %s != %s (%g vs %g)\n", __FILE__, __LINE__, #a, #b, (double) ( a ), \
                      (double) ( b ) );                                                             \
         failures++;                                                                                \
      }                                                                                             \
   } while( 0 )

template< typename Matrix >
void test_smallMatrix1()
{
   using Vector = Containers::Vector< typename Matrix::VectorReal, Devices::Hip, int >;
   Matrix m( 4, 4 );
   m.setRowCapacities( std::vector< int >{ 1, 2, 1, 1 } );
   double value = 1;
   m.setElement( 0, 0, value++ );
   m.setElement( 1, 1, value++ );
   m.setElement( 1, 3, value++ );
   m.setElement( 2, 1, value++ );
   m.setElement( 3, 2, value++ );
   Vector in( 4, 2.0 ), out( 4, 0.0 );
   m.vectorProduct( in, out );
   EXPECT_EQ( out.getElement( 0 ), 2 );
   EXPECT_EQ( out.getElement( 1 ), 10 );
   EXPECT_EQ( out.getElement( 2 ), 8 );
   EXPECT_EQ( out.getElement( 3 ), 10 );
}

template< typename Matrix >
void test_mediumSizeMatrix1()
{
   using Vector = Containers::Vector< typename Matrix::VectorReal, Devices::Hip, int >;
   Matrix m( 8, 8 );
   m.setRowCapacities( std::vector< int >{ 4, 4, 5, 4, 4, 4, 5, 5 } );
   double v = 1;
   for( int i = 0; i < 3; i++ ) m.setElement( 0, i, v++ );
   m.setElement( 0, 5, v++ );
   for( int i = 1; i < 5; i++ ) m.setElement( 1, i, v++ );
   for( int i = 0; i < 5; i++ ) m.setElement( 2, i, v++ );
   for( int i = 1; i < 5; i++ ) m.setElement( 3, i, v++ );
   for( int i = 2; i < 6; i++ ) m.setElement( 4, i, v++ );
   for( int i = 3; i < 7; i++ ) m.setElement( 5, i, v++ );
   for( int i = 0; i < 5; i++ ) m.setElement( 6, i, v++ );
   for( int i = 0; i < 5; i++ ) m.setElement( 7, i, v++ );
   Vector in( 8, 2.0 ), out( 8, 0.0 );
   m.vectorProduct( in, out );
   const double want[] = { 20, 52, 110, 124, 156, 188, 280, 330 };
   for( int i = 0; i < 8; i++ )
      EXPECT_EQ( out.getElement( i ), want[ i ] );
}

template< typename Matrix >
void test_largeMatrix()
{
   using Vector = Containers::Vector< typename Matrix::VectorReal, Devices::Hip, int >;
   const int size = 1051;
   Matrix m2( size, size );
   std::vector< int > caps( size );
   for( int i = 0; i < size; i++ ) caps[ i ] = i + 1;
   m2.setRowCapacities( caps );
   // lower triangle a_ij = j + 1, assembled as host CSR and assigned (operator= from CSR)
   std::vector< int > rp( size + 1, 0 ), cols;
   std::vector< double > vals;
   for( int i = 0; i < size; i++ ) {
      for( int j = 0; j <= i; j++ ) {
         cols.push_back( j );
         vals.push_back( j + 1 );
      }
      rp[ i + 1 ] = (int) cols.size();
   }
   m2.setFromCSR( size, size, rp.data(), cols.data(), vals.data() );
   Vector in( size, 1.0 ), out( size, 0.0 );
   m2.vectorProduct( in, out );
   auto h = out.toHost();
   for( int i = 0; i < size; i++ )
      EXPECT_EQ( h[ i ], ( i + 1 ) * ( i + 2 ) / 2 );
}

template< typename Matrix >
void test_longRowsMatrix()
{
   using Vector = Containers::Vector< typename Matrix::VectorReal, Devices::Hip, int >;
   for( int columns : { 64, 65, 128, 129, 256, 257, 512, 513, 1024, 1025, 2048, 2049, 3000 } ) {
      const int rows = 33;
      std::vector< int > rp( rows + 1, 0 ), cols;
      std::vector< double > vals;
      for( int r = 0; r < rows; r++ ) {
         for( int l = 0; l < columns; l++ ) {
            cols.push_back( l );
            vals.push_back( l + r );
         }
         rp[ r + 1 ] = (int) cols.size();
      }
      Matrix m;
      m.setFromCSR( rows, columns, rp.data(), cols.data(), vals.data() );
      Vector in( columns, 1.0 ), out( rows, 0.0 );
      m.vectorProduct( in, out );
      auto h = out.toHost();
      for( int r = 0; r < rows; r++ )
         EXPECT_EQ( h[ r ], (double) columns * ( columns - 1 ) / 2.0 + columns * r );
   }
}

template< template< typename, typename, typename > class S >
void runFormat( const char* name )
{
   using M64 = Matrices::SparseMatrix< double, Devices::Hip, int, Matrices::GeneralMatrix, S >;
   using M32 = Matrices::SparseMatrix< float, Devices::Hip, int, Matrices::GeneralMatrix, S >;
   const int before = failures;
   test_smallMatrix1< M64 >();
   test_mediumSizeMatrix1< M64 >();
   test_largeMatrix< M64 >();
   test_longRowsMatrix< M64 >();
   test_smallMatrix1< M32 >();
   test_mediumSizeMatrix1< M32 >();
   // cross-format conversion: B = A on the device
   M64 a( 4, 4 );
   a.setRowCapacities( std::vector< int >{ 1, 2, 1, 1 } );
   a.setElement( 0, 0, 1 );
   a.setElement( 1, 1, 2 );
   a.setElement( 1, 3, 3 );
   a.setElement( 2, 1, 4 );
   a.setElement( 3, 2, 5 );
   Matrices::SparseMatrix< double, Devices::Hip, int, Matrices::GeneralMatrix, Seg::BiEllpack > b;
   b = a;
   Containers::Vector< double, Devices::Hip, int > in( 4, 2.0 ), out( 4, 0.0 );
   b.vectorProduct( in, out );
   EXPECT_EQ( out.getElement( 1 ), 10 );
   std::printf( "%s %s\n", failures == before ? "PASS" : "FAIL", name );
}

int main()
{
   runFormat< Seg::CSRScalar >( "CSRScalar" );
   runFormat< Seg::CSRVector >( "CSRVector" );
   runFormat< Seg::CSRLight >( "CSRLight" );
   runFormat< Seg::CSRHybrid >( "CSRHybrid" );
   runFormat< Seg::CSRAdaptive >( "CSRAdaptive" );
   runFormat< Seg::Ellpack >( "Ellpack" );
   runFormat< Seg::SlicedEllpack >( "SlicedEllpack" );
   runFormat< Seg::ChunkedEllpack >( "ChunkedEllpack" );
   runFormat< Seg::BiEllpack >( "BiEllpack" );

   // capacity overflow is std::logic_error, as in SparseMatrixView::setElement
   {
      Matrices::SparseMatrix< double > m( 2, 2 );
      m.setRowCapacities( std::vector< int >{ 1, 1 } );
      m.setElement( 0, 0, 1.0 );
      bool thrown = false;
      try {
         m.setElement( 0, 1, 1.0 );
      }
      catch( const std::logic_error& ) {
         thrown = true;
      }
      EXPECT_EQ( thrown, true );
   }

   // Solvers: 2D 5-point Laplacian 64^2, CG and GMRES (CWY) with Jacobi to 1e-8
   {
      using Matrix = Matrices::SparseMatrix< double, Devices::Hip, int, Matrices::GeneralMatrix, Seg::CSRLight >;
      using Vector = Containers::Vector< double, Devices::Hip, int >;
      const int N = 64, n = N * N;
      auto A = std::make_shared< Matrix >( n, n );
      std::vector< int > caps( n, 5 );
      A->setRowCapacities( caps );
      for( int j = 0; j < N; j++ )
         for( int i = 0; i < N; i++ ) {
            const int r = i + N * j;
            if( j > 0 ) A->setElement( r, r - N, -1 );
            if( i > 0 ) A->setElement( r, r - 1, -1 );
            A->setElement( r, r, 4 );
            if( i < N - 1 ) A->setElement( r, r + 1, -1 );
            if( j < N - 1 ) A->setElement( r, r + N, -1 );
         }
      std::vector< double > bh( n );
      for( int i = 0; i < n; i++ ) bh[ i ] = std::sin( 0.37 * i );
      Vector b( bh ), x( n, 0.0 );
      Solvers::Linear::CG< Matrix > cg;
      cg.setMatrix( A );
      cg.setConvergenceResidue( 1e-8 );
      cg.setMaxIterations( 10000 );
      const bool ok = cg.solve( b, x );
      std::printf( "CG converged=%d iterations=%d residue=%.3e\n", (int) ok, cg.getIterations(), cg.getResidue() );
      EXPECT_EQ( ok, true );
      Vector r( n );
      A->vectorProduct( x, r );
      auto rh = r.toHost();
      double num = 0, den = 0;
      for( int i = 0; i < n; i++ ) {
         num += ( bh[ i ] - rh[ i ] ) * ( bh[ i ] - rh[ i ] );
         den += bh[ i ] * bh[ i ];
      }
      std::printf( "CG true residual %.3e\n", std::sqrt( num / den ) );
      EXPECT_EQ( std::sqrt( num / den ) < 1e-7, true );

      Solvers::Linear::GMRES< Matrix > gmres;
      gmres.setMatrix( A );
      gmres.setPreconditioner( std::make_shared< Solvers::Linear::Preconditioners::Diagonal< Matrix > >() );
      Config::ParameterContainer prm;
      prm.addParameter< std::string >( "gmres-variant", "CWY" );
      prm.addParameter< int >( "gmres-restarting-min", 30 );
      prm.addParameter< int >( "gmres-restarting-max", 30 );
      prm.addParameter< double >( "convergence-residue", 1e-8 );
      prm.addParameter< int >( "max-iterations", 20000 );
      gmres.setup( prm );
      Vector x2( n, 0.0 );
      const bool ok2 = gmres.solve( b, x2 );
      std::printf( "GMRES converged=%d iterations=%d residue=%.3e\n", (int) ok2, gmres.getIterations(),
                   gmres.getResidue() );
      EXPECT_EQ( ok2, true );

      // FP32 GMRES with the Jacobi polynomial over a BF16 copy (no TNL counterpart): fewer
      // iterations than Jacobi on the same FP32 system, same tolerance
      {
         using MatrixF = Matrices::SparseMatrix< float, Devices::Hip, int, Matrices::GeneralMatrix, Seg::CSRScalar >;
         using VectorF = Containers::Vector< float, Devices::Hip, int >;
         auto Af = std::make_shared< MatrixF >( n, n );
         Af->setRowCapacities( caps );
         for( int j = 0; j < N; j++ )
            for( int i = 0; i < N; i++ ) {
               const int r = i + N * j;
               if( j > 0 ) Af->setElement( r, r - N, -1 );
               if( i > 0 ) Af->setElement( r, r - 1, -1 );
               Af->setElement( r, r, 4 );
               if( i < N - 1 ) Af->setElement( r, r + 1, -1 );
               if( j < N - 1 ) Af->setElement( r, r + N, -1 );
            }
         std::vector< float > bf( bh.begin(), bh.end() );
         VectorF b4( bf );
         int its[ 2 ] = { 0, 0 };
         for( int which = 0; which < 2; which++ ) {
            Solvers::Linear::GMRES< MatrixF > g;
            g.setMatrix( Af );
            if( which == 0 )
               g.setPreconditioner( std::make_shared< Solvers::Linear::Preconditioners::Diagonal< MatrixF > >() );
            else {
               auto poly = std::make_shared< Solvers::Linear::Preconditioners::JacobiPolynomial< MatrixF > >(
                  4, TNLHIP_BF16 );
               poly->update( Af );
               g.setPreconditioner( poly );
            }
            Config::ParameterContainer pf;
            pf.addParameter< int >( "gmres-restarting-min", 30 );
            pf.addParameter< int >( "gmres-restarting-max", 30 );
            pf.addParameter< double >( "convergence-residue", 1e-4 );
            pf.addParameter< int >( "max-iterations", 20000 );
            g.setup( pf );
            VectorF xf( n, 0.0f );
            const bool okf = g.solve( b4, xf );
            its[ which ] = g.getIterations();
            VectorF rf( n );
            Af->vectorProduct( xf, rf );
            auto rfh = rf.toHost();
            double num2 = 0, den2 = 0;
            for( int i = 0; i < n; i++ ) {
               num2 += ( bf[ i ] - rfh[ i ] ) * ( bf[ i ] - rfh[ i ] );
               den2 += (double) bf[ i ] * bf[ i ];
            }
            std::printf( "FP32 GMRES %s converged=%d iterations=%d true residual %.3e\n",
                         which ? "JacobiPolynomial(4, BF16)" : "Jacobi", (int) okf, its[ which ],
                         std::sqrt( num2 / den2 ) );
            EXPECT_EQ( okf, true );
            EXPECT_EQ( std::sqrt( num2 / den2 ) < 1e-3, true );
         }
         EXPECT_EQ( its[ 1 ] < its[ 0 ], true );
      }

      // run-time choice by name (LinearSolverTypeResolver.h): BiCGStab with ILU(0)
      auto solver = Solvers::getLinearSolver< Matrix >( "bicgstab" );
      auto pre = Solvers::getPreconditioner< Matrix >( "ilu0" );
      EXPECT_EQ( solver != nullptr && pre != nullptr, true );
      EXPECT_EQ( Solvers::getLinearSolver< Matrix >( "sor" ) == nullptr, true );
      solver->setMatrix( A );
      solver->setPreconditioner( pre );
      solver->setConvergenceResidue( 1e-8 );
      solver->setMaxIterations( 10000 );
      Vector x3( n, 0.0 );
      const bool ok3 = solver->solve( b, x3 );
      std::printf( "BiCGStab+ILU0 converged=%d iterations=%d residue=%.3e\n", (int) ok3, solver->getIterations(),
                   solver->getResidue() );
      EXPECT_EQ( ok3, true );
      EXPECT_EQ( solver->getIterations() < cg.getIterations(), true );
   }
   // Matrix Market: readMtx into a SlicedEllpack matrix on the device, writeMtx back
   {
      using Matrix = Matrices::SparseMatrix< double, Devices::Hip, int, Matrices::GeneralMatrix, Seg::SlicedEllpack >;
      const char* path = "/tmp/tnlhip_front_test.mtx";
      {
         std::FILE* f = std::fopen( path, "w" );
         std::fprintf( f, "%%%%MatrixMarket matrix coordinate real symmetric\n%% test\n3 3 4\n1 1 2\n2 1 -1\n2 2 2\n"
                          "3 3 5\n" );
         std::fclose( f );
      }
      Matrix A;
      Matrices::MatrixReader< Matrix >::readMtx( path, A );
      EXPECT_EQ( A.getRows(), 3 );
      EXPECT_EQ( A.getColumns(), 3 );
      // A = [[2,-1,0],[-1,2,0],[0,0,5]]: A (1,2,3) = (0,3,15)
      Containers::Vector< double, Devices::Hip, int > xv( std::vector< double >{ 1, 2, 3 } ), yv( 3, 0.0 );
      A.vectorProduct( xv, yv );
      const auto yh = yv.toHost();
      EXPECT_EQ( yh[ 0 ], 0.0 );
      EXPECT_EQ( yh[ 1 ], 3.0 );
      EXPECT_EQ( yh[ 2 ], 15.0 );
      Matrices::MatrixWriter< Matrix >::writeMtx( "/tmp/tnlhip_front_test.out.mtx", A );
      std::FILE* f = std::fopen( "/tmp/tnlhip_front_test.out.mtx", "r" );
      int lines = 0;
      for( int ch; ( ch = std::fgetc( f ) ) != EOF; )
         lines += ch == '\n';
      std::fclose( f );
      EXPECT_EQ( lines, 5 + 5 );  // header block of 5 lines, 5 stored entries
      std::remove( path );
      std::remove( "/tmp/tnlhip_front_test.out.mtx" );
      std::printf( "%s MatrixMarket\n", failures ? "FAIL" : "PASS" );
      // binary save / load: the loaded matrix gives the same product
      A.save( "/tmp/tnlhip_front_test.tnl" );
      Matrix B;
      B.load( "/tmp/tnlhip_front_test.tnl" );
      Containers::Vector< double, Devices::Hip, int > zv( 3, 0.0 );
      B.vectorProduct( xv, zv );
      const auto zh = zv.toHost();
      EXPECT_EQ( zh[ 0 ], 0.0 );
      EXPECT_EQ( zh[ 1 ], 3.0 );
      EXPECT_EQ( zh[ 2 ], 15.0 );
      std::remove( "/tmp/tnlhip_front_test.tnl" );
      // a file of another segments type is refused, as SparseMatrix::load does
      A.save( "/tmp/tnlhip_front_test.tnl" );
      bool refused = false;
      try {
         Matrices::SparseMatrix< double, Devices::Hip, int, Matrices::GeneralMatrix, Seg::Ellpack > E;
         E.load( "/tmp/tnlhip_front_test.tnl" );
      }
      catch( const std::exception& ) {
         refused = true;
      }
      EXPECT_EQ( refused, true );
      std::remove( "/tmp/tnlhip_front_test.tnl" );
      std::printf( "%s save/load\n", failures ? "FAIL" : "PASS" );
   }
   // SymmetricMatrix: setElement above the diagonal lands on the mirror (SymmetricSparseMatrixTest.hpp:679-712)
   {
      using Sym = Matrices::SparseMatrix< double, Devices::Hip, int, Matrices::SymmetricMatrix, Seg::CSRDefault >;
      Sym S( 4, 4 );
      S.setRowCapacities( std::vector< int >{ 1, 1, 2, 2 } );
      S.setElement( 0, 0, 1 );
      S.setElement( 1, 1, 2 );
      S.setElement( 1, 2, 3 );  // stored as (2, 1)
      S.setElement( 3, 1, 4 );
      S.setElement( 3, 2, 5 );
      Containers::Vector< double, Devices::Hip, int > xs( 4, 2.0 ), ys( 4, 0.0 );
      S.vectorProduct( xs, ys );
      const auto yh = ys.toHost();
      EXPECT_EQ( yh[ 0 ], 2.0 );
      EXPECT_EQ( yh[ 1 ], 18.0 );
      EXPECT_EQ( yh[ 2 ], 16.0 );
      EXPECT_EQ( yh[ 3 ], 18.0 );
      S.save( "/tmp/tnlhip_front_sym.tnl" );
      Sym T;
      T.load( "/tmp/tnlhip_front_sym.tnl" );
      Containers::Vector< double, Devices::Hip, int > zs( 4, 0.0 );
      T.vectorProduct( xs, zs );
      EXPECT_EQ( zs.toHost()[ 1 ], 18.0 );
      bool refused = false;
      try {
         Matrices::SparseMatrix< double, Devices::Hip, int, Matrices::GeneralMatrix, Seg::CSRDefault > G;
         G.load( "/tmp/tnlhip_front_sym.tnl" );
      }
      catch( const std::exception& ) {
         refused = true;
      }
      EXPECT_EQ( refused, true );
      std::remove( "/tmp/tnlhip_front_sym.tnl" );
      std::printf( "%s SymmetricMatrix\n", failures ? "FAIL" : "PASS" );
   }
   // binary matrix (Real = bool, int vectors): BinarySparseMatrixTest.hpp:592-635
   {
      using Bin = Matrices::SparseMatrix< bool, Devices::Hip, int, Matrices::GeneralMatrix, Seg::Ellpack >;
      Bin B( 4, 4 );
      B.setRowCapacities( std::vector< int >{ 1, 2, 1, 1 } );
      B.setElement( 0, 0, 1 );
      B.setElement( 1, 1, 1 );
      B.setElement( 1, 3, 1 );
      B.setElement( 2, 1, 1 );
      B.setElement( 3, 2, 1 );
      Containers::Vector< int, Devices::Hip, int > xb( 4, 2 ), yb( 4, 0 );
      B.vectorProduct( xb, yb );
      const auto yh = yb.toHost();
      EXPECT_EQ( yh[ 0 ], 2 );
      EXPECT_EQ( yh[ 1 ], 4 );
      EXPECT_EQ( yh[ 2 ], 2 );
      EXPECT_EQ( yh[ 3 ], 2 );
      std::printf( "%s binary SparseMatrix\n", failures ? "FAIL" : "PASS" );
   }
   std::printf( failures ? "FAILURES %d\n" : "ALL PASSED\n", failures );
   return failures ? 1 : 0;
}
