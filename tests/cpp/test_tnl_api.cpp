// The reference's own SpMV known-answer tests and solver example, written
// against TNL's class templates with Devices::MI355X in place of
// Devices::Host (src/UnitTests/Matrices/SparseMatrixVectorProductTest.hpp,
// Documentation/Examples/Solvers/Linear/IterativeLinearSolverExample.cpp).
// The reference's own headers come first: the device is added beside them,
// and the solver example is one template instantiated for Devices::Host (the
// reference's code) and for Devices::MI355X (this library).
// Prints "ALL OK" and exits 0 when every check passes.
#include <sstream>
#include <unistd.h>

#include <TNL/Containers/Vector.h>
#include <TNL/Matrices/SparseMatrix.h>
#include <TNL/Solvers/Linear/CG.h>
#include <TNL/Solvers/Linear/Preconditioners/Diagonal.h>

#include <TNL/MI355X.h>

#include <cmath>
#include <cstdio>
#include <iostream>
#include <memory>

using namespace TNL;
using Device = Devices::MI355X;
using Vector = Containers::Vector< double, Device, int >;
namespace Seg = Algorithms::Segments;
template< typename D, typename I, typename A >
using SlicedEllpack64 = Seg::SlicedEllpack< D, I, A, Seg::ColumnMajorOrder, 64 >;

// the reference's full parameter lists, allocators spelled out, name the
// same types as its defaults
static_assert( std::is_same_v< Matrices::SparseMatrix< double, Device, int, Matrices::GeneralMatrix, Seg::CSR, double,
                                                      Allocators::MI355X< double >, Allocators::MI355X< int > >,
                               Matrices::SparseMatrix< double, Device > > );
static_assert( std::is_same_v< Containers::Vector< double, Device, int, Allocators::MI355X< double > >, Vector > );
static_assert( std::is_base_of_v< Solvers::Linear::LinearSolver< Matrices::SparseMatrix< double, Device > >,
                                  Solvers::Linear::CG< Matrices::SparseMatrix< double, Device > > > );

static int failures = 0;
#define EXPECT_EQ( a, b )                                                                                            \
   do {                                                                                                              \
      if( !( ( a ) == ( b ) ) ) {                                                                                    \
         std::cerr << __FILE__ << ":" << __LINE__ << ": " #a " == " #b " failed (" << ( a ) << " vs " << ( b ) << ")\n"; \
         failures++;                                                                                                 \
      }                                                                                                              \
   } while( 0 )

// IndexT = int or long: the reference's SpMV tests instantiate both index
// types (src/UnitTests/Matrices/SparseMatrixTest_types.h)
template< template< class, class, class > class Segments, typename IndexT = int >
void vectorProductKATs( const char* name, bool csr = false )
{
   using Matrix = Matrices::SparseMatrix< double, Device, IndexT, Matrices::GeneralMatrix, Segments >;
   std::vector< Seg::LaunchConfiguration > configs{ Seg::LaunchConfiguration() };
   if( csr ) {
      for( int t : { 2, 4, 8, 16, 32, 64, 128, 256 } )
         configs.emplace_back( Seg::ThreadsToSegmentsMapping::Fixed, t );
      configs.emplace_back( Seg::ThreadsToSegmentsMapping::Warp, 1 );
      configs.emplace_back( Seg::ThreadsToSegmentsMapping::DynamicGrouping, 1 );
   }
   for( const auto& cfg : configs ) {
      // zeroMatrix: capacities [1, 2, 1, 1], no entries
      {
         Matrix m( 4, 4 );
         m.setRowCapacities( std::vector< IndexT >{ 1, 2, 1, 1 } );
         Vector in( 4, 1 ), out( 4, -1 );
         m.vectorProduct( in, out, cfg );
         for( int i = 0; i < 4; i++ )
            EXPECT_EQ( out.getElement( i ), 0.0 );
      }
      // smallMatrix1
      {
         Matrix m( 4, 4, { { 0, 0, 1 }, { 1, 1, 2 }, { 1, 3, 3 }, { 2, 1, 4 }, { 3, 2, 5 } } );
         Vector in( 4, 2 ), out( 4, 0 );
         m.vectorProduct( in, out, cfg );
         EXPECT_EQ( out.getElement( 0 ), 2.0 );
         EXPECT_EQ( out.getElement( 1 ), 10.0 );
         EXPECT_EQ( out.getElement( 2 ), 8.0 );
         EXPECT_EQ( out.getElement( 3 ), 10.0 );
      }
      // smallMatrix3
      {
         Matrix m( 4, 4,
                   { { 0, 0, 1 }, { 0, 1, 2 }, { 0, 2, 3 }, { 1, 1, 4 }, { 1, 2, 5 }, { 1, 3, 6 },
                     { 2, 0, 7 }, { 2, 1, 8 }, { 2, 2, 9 }, { 3, 1, 10 }, { 3, 2, 11 }, { 3, 3, 12 } } );
         Vector in( 4, 2 ), out( 4, 0 );
         m.vectorProduct( in, out, cfg );
         EXPECT_EQ( out.getElement( 0 ), 12.0 );
         EXPECT_EQ( out.getElement( 1 ), 30.0 );
         EXPECT_EQ( out.getElement( 2 ), 48.0 );
         EXPECT_EQ( out.getElement( 3 ), 66.0 );
      }
      // mediumSizeMatrix1
      {
         Matrix m( 8, 8,
                   { { 0, 0, 1 },  { 0, 1, 2 },  { 0, 2, 3 },  { 0, 5, 4 },  { 1, 1, 5 },  { 1, 2, 6 },  { 1, 3, 7 },
                     { 1, 4, 8 },  { 2, 0, 9 },  { 2, 1, 10 }, { 2, 2, 11 }, { 2, 3, 12 }, { 2, 4, 13 }, { 3, 1, 14 },
                     { 3, 2, 15 }, { 3, 3, 16 }, { 3, 4, 17 }, { 4, 2, 18 }, { 4, 3, 19 }, { 4, 4, 20 }, { 4, 5, 21 },
                     { 5, 3, 22 }, { 5, 4, 23 }, { 5, 5, 24 }, { 5, 6, 25 }, { 6, 0, 26 }, { 6, 1, 27 }, { 6, 2, 28 },
                     { 6, 3, 29 }, { 6, 4, 30 }, { 7, 0, 31 }, { 7, 1, 32 }, { 7, 2, 33 }, { 7, 3, 34 }, { 7, 4, 35 } } );
         Vector in( 8, 2 ), out( 8, 0 );
         m.vectorProduct( in, out, cfg );
         const double y[ 8 ] = { 20, 52, 110, 124, 156, 188, 280, 330 };
         for( int i = 0; i < 8; i++ )
            EXPECT_EQ( out.getElement( i ), y[ i ] );
      }
      // largeMatrix: lower triangular, a_ij = j + 1
      {
         const int n = 1051;
         Matrix m( n, n );
         std::vector< IndexT > off( n + 1, 0 ), cols;
         std::vector< double > vals;
         for( int i = 0; i < n; i++ ) {
            for( int j = 0; j <= i; j++ ) {
               cols.push_back( j );
               vals.push_back( j + 1 );
            }
            off[ i + 1 ] = (IndexT) cols.size();
         }
         m.setFromCSR( off, cols, vals );
         Vector in( n, 1.0 ), out( n, 0.0 );
         m.vectorProduct( in, out, cfg );
         const auto h = out.toHost();
         for( int i = 0; i < n; i++ )
            EXPECT_EQ( h[ i ], double( ( i + 1 ) * ( i + 2 ) / 2 ) );
      }
   }
   // alpha / beta / row range (SparseMatrixBase.hpp:346-389)
   {
      Matrix m( 4, 4, { { 0, 0, 1 }, { 1, 1, 2 }, { 1, 3, 3 }, { 2, 1, 4 }, { 3, 2, 5 } } );
      Vector in( 4, 2 ), out( 4, 1 );
      m.vectorProduct( in, out, 2.0, 3.0, 1, 3 );
      EXPECT_EQ( out.getElement( 0 ), 1.0 );
      EXPECT_EQ( out.getElement( 1 ), 2.0 * 10 + 3.0 );
      EXPECT_EQ( out.getElement( 2 ), 2.0 * 8 + 3.0 );
      EXPECT_EQ( out.getElement( 3 ), 1.0 );
   }
   std::cout << name << ( std::is_same_v< IndexT, int > ? "" : " (Index = long)" ) << " KATs done\n";
}

// The solver example written once for any device: the reference's code for
// Devices::Host, this library for Devices::MI355X.  Returns the iterations.
template< typename Solver, typename Matrix >
int solveExample( const std::shared_ptr< Matrix >& A, const char* name, bool diag = false )
{
   using V = Containers::Vector< double, typename Matrix::DeviceType, typename Matrix::IndexType >;
   V x( 5, 1.0 ), b( 5 );
   A->vectorProduct( x, b );
   x = 0.0;
   Solver s;
   s.setMatrix( A );
   if( diag ) {
      auto p = std::make_shared< Solvers::Linear::Preconditioners::Diagonal< Matrix > >();
      p->update( A );
      s.setPreconditioner( p );
   }
   s.setConvergenceResidue( 1e-12 );
   const bool ok = s.solve( b, x );
   EXPECT_EQ( ok, true );
   for( int i = 0; i < 5; i++ )
      EXPECT_EQ( std::fabs( x.getElement( i ) - 1.0 ) < 1e-10, true );
   EXPECT_EQ( Solvers::Linear::LinearResidueGetter::getResidue( *A, x, b ) < 1e-12, true );
   const char* device = std::is_same_v< typename Matrix::DeviceType, Devices::Host > ? "Host" : "MI355X";
   std::cout << name << " (" << device << "): " << s.getIterations()
             << " iterations, residue " << s.getResidue() << '\n';
   return s.getIterations();
}

// IterativeLinearSolverExample.cpp's matrix, tridiag(-1, 2.5, -1), on any device
template< typename D, typename IndexT = int >
auto exampleMatrix()
{
   using M = Matrices::SparseMatrix< double, D, IndexT >;
   return std::make_shared< M >( 5, 5,
                                 std::initializer_list< std::tuple< IndexT, IndexT, double > >{
                                    { 0, 0, 2.5 }, { 0, 1, -1 }, { 1, 0, -1 }, { 1, 1, 2.5 }, { 1, 2, -1 },
                                    { 2, 1, -1 },  { 2, 2, 2.5 }, { 2, 3, -1 }, { 3, 2, -1 }, { 3, 3, 2.5 },
                                    { 3, 4, -1 },  { 4, 3, -1 },  { 4, 4, 2.5 } } );
}

int main()
{
   vectorProductKATs< Seg::CSR >( "CSR", true );
   vectorProductKATs< Seg::AdaptiveCSR >( "AdaptiveCSR" );
   vectorProductKATs< Seg::Ellpack >( "Ellpack" );
   vectorProductKATs< Seg::SlicedEllpack >( "SlicedEllpack" );
   vectorProductKATs< SlicedEllpack64 >( "SlicedEllpack<64>" );
   vectorProductKATs< Seg::BiEllpack >( "BiEllpack" );
   vectorProductKATs< Seg::ChunkedEllpack >( "ChunkedEllpack" );
   vectorProductKATs< Seg::SortedCSR >( "SortedCSR" );
   vectorProductKATs< Seg::CSR, long >( "CSR", true );
   vectorProductKATs< Seg::Ellpack, long >( "Ellpack" );
   vectorProductKATs< Seg::SlicedEllpack, long >( "SlicedEllpack" );
   vectorProductKATs< Seg::BiEllpack, long >( "BiEllpack" );
   vectorProductKATs< Seg::ChunkedEllpack, long >( "ChunkedEllpack" );

   // SparseMatrixExample_addElement.cpp: diagonal i, then addElement( i, j, 1, 5 )
   // everywhere: 5 i + 1 on the diagonal, 1 elsewhere (inserted in column order)
   {
      Matrices::SparseMatrix< double, Device > M( { 5, 5, 5, 5, 5 }, 5 );
      for( int i = 0; i < 5; i++ )
         M.setElement( i, i, i );
      for( int i = 0; i < 5; i++ )
         for( int j = 0; j < 5; j++ )
            M.addElement( i, j, 1.0, 5.0 );
      for( int i = 0; i < 5; i++ )
         for( int j = 0; j < 5; j++ )
            EXPECT_EQ( M.getElement( i, j ), i == j ? 5.0 * i + 1.0 : 1.0 );
      Vector x( 5, 1.0 ), y( 5 );
      M.vectorProduct( x, y );
      for( int i = 0; i < 5; i++ )
         EXPECT_EQ( y.getElement( i ), 5.0 * i + 5.0 );
      bool thrown = false;
      try {
         Matrices::SparseMatrix< double, Device > F( { 1, 1 }, 3 );
         F.setElement( 0, 2, 1.0 );
         F.setElement( 0, 1, 2.0 );  // full row: the reference shifts, the last slot falls off
         EXPECT_EQ( F.getElement( 0, 1 ), 2.0 );
         EXPECT_EQ( F.getElement( 0, 2 ), 0.0 );
         F.setElement( 0, 2, 3.0 );  // full row, no larger column
      }
      catch( const std::logic_error& e ) {
         thrown = std::string( e.what() ) == "The capacity of the sparse matrix row number 0 was exceeded.";
      }
      EXPECT_EQ( thrown, true );
      // a symmetric matrix stores (j, i) for i < j and reads it back from either side
      Matrices::SparseMatrix< double, Device, int, Matrices::SymmetricMatrix > S( { 1, 2, 2 }, 3 );
      S.setElement( 0, 0, 4.0 );
      S.setElement( 0, 1, -1.0 );
      S.setElement( 1, 1, 4.0 );
      S.setElement( 1, 2, -1.0 );
      S.setElement( 2, 2, 4.0 );
      EXPECT_EQ( S.getElement( 1, 0 ), -1.0 );
      EXPECT_EQ( S.getElement( 1, 2 ), -1.0 );
      EXPECT_EQ( S.getElement( 0, 2 ), 0.0 );
      Vector s1( 3, 1.0 ), s2( 3 );
      S.vectorProduct( s1, s2 );
      EXPECT_EQ( s2.getElement( 0 ), 3.0 );
      EXPECT_EQ( s2.getElement( 1 ), 2.0 );
      EXPECT_EQ( s2.getElement( 2 ), 3.0 );
      EXPECT_EQ( M.getNonzeroElementsCount(), (int64_t) 25 );
      EXPECT_EQ( S.getNonzeroElementsCount(), (int64_t) 7 );  // 3 diagonal + 2 x 2 off-diagonal
      Matrices::SparseMatrix< double, Device, int, Matrices::GeneralMatrix, Seg::Ellpack > E( { 1, 3 }, 4 );
      EXPECT_EQ( E.getRowCapacity( 0 ), 3 );  // an Ellpack row has the full width
      E.setElement( 0, 1, 0.0 );
      E.setElement( 1, 2, 5.0 );
      Matrices::SparseMatrix< double, Device >::RowCapacitiesType caps, lens;
      E.getRowCapacities( caps );
      E.getCompressedRowLengths( lens );
      EXPECT_EQ( caps.getElement( 0 ), 1 );
      EXPECT_EQ( lens.getElement( 0 ), 0 );  // a stored zero
      EXPECT_EQ( lens.getElement( 1 ), 1 );
      std::cout << "setElement / addElement / getElement: ok\n";
   }

   // IterativeLinearSolverExample.cpp: tridiag(-1, 2.5, -1), b = A * 1
   using Matrix = Matrices::SparseMatrix< double, Device, int >;
   auto A = std::make_shared< Matrix >( 5, 5,
                                        std::initializer_list< std::tuple< int, int, double > >{
                                           { 0, 0, 2.5 },  { 0, 1, -1 },  { 1, 0, -1 },  { 1, 1, 2.5 }, { 1, 2, -1 },
                                           { 2, 1, -1 },   { 2, 2, 2.5 }, { 2, 3, -1 },  { 3, 2, -1 },  { 3, 3, 2.5 },
                                           { 3, 4, -1 },   { 4, 3, -1 },  { 4, 4, 2.5 } } );
   // the same code on the reference's host device and on this one: the same
   // iterations
   {
      using HostMatrix = Matrices::SparseMatrix< double, Devices::Host, int >;
      auto H = exampleMatrix< Devices::Host >();
      auto G = exampleMatrix< Device >();
      EXPECT_EQ( solveExample< Solvers::Linear::CG< HostMatrix > >( H, "CG" ),
                 solveExample< Solvers::Linear::CG< Matrix > >( G, "CG" ) );
      EXPECT_EQ( solveExample< Solvers::Linear::CG< HostMatrix > >( H, "CG + Diagonal", true ),
                 solveExample< Solvers::Linear::CG< Matrix > >( G, "CG + Diagonal", true ) );
      // Index = long on both devices
      using HostMatrixL = Matrices::SparseMatrix< double, Devices::Host, long >;
      using MatrixL = Matrices::SparseMatrix< double, Device, long >;
      auto HL = exampleMatrix< Devices::Host, long >();
      auto GL = exampleMatrix< Device, long >();
      EXPECT_EQ( solveExample< Solvers::Linear::CG< HostMatrixL > >( HL, "CG (Index = long)" ),
                 solveExample< Solvers::Linear::CG< MatrixL > >( GL, "CG (Index = long)" ) );
      EXPECT_EQ( solveExample< Solvers::Linear::GMRES< HostMatrixL > >( HL, "GMRES + Diagonal (Index = long)", true ),
                 solveExample< Solvers::Linear::GMRES< MatrixL > >( GL, "GMRES + Diagonal (Index = long)", true ) );
      {
         Containers::Vector< long, Device, long > caps;
         GL->getRowCapacities( caps );
         const auto hc = caps.toHost();
         EXPECT_EQ( hc[ 0 ], 2L );
         EXPECT_EQ( hc[ 2 ], 3L );
         EXPECT_EQ( GL->getNonzeroElementsCount(), 13L );
         EXPECT_EQ( GL->getElement( 3, 4 ), -1.0 );
      }
      // cross-device assignment of the reference's host matrix and vector
      Matrix D;
      D = *H;
      Containers::Vector< double, Devices::Host, int > hx( 5 );
      for( int i = 0; i < 5; i++ )
         hx[ i ] = 1.0 + i;
      Vector dx, y1( 5 ), y2( 5 );
      dx = hx;
      D.vectorProduct( dx, y1 );
      G->vectorProduct( dx, y2 );
      EXPECT_EQ( y1 == y2, true );
      EXPECT_EQ( y1.getElement( 4 ), 2.5 * 5 - 4 );
      EXPECT_EQ( ( dx, dx ), 55.0 );
      EXPECT_EQ( Containers::l2Norm( dx - dx ), 0.0 );
   }
   solveExample< Solvers::Linear::CG< Matrix > >( A, "CG" );
   solveExample< Solvers::Linear::CG< Matrix > >( A, "CG + Diagonal", true );
   // GMRES with the class default orthogonalisation (CWY, GMRES.h:212)
   solveExample< Solvers::Linear::GMRES< Matrix > >( A, "GMRES (default CWY)" );
   solveExample< Solvers::Linear::GMRES< Matrix > >( A, "GMRES (default CWY) + Diagonal", true );
   {
      Solvers::Linear::GMRES< Matrix > g;
      Config::ParameterContainer p;
      p.addParameter< std::string >( "gmres-variant", "MGSR" );
      EXPECT_EQ( g.setup( p ), true );
      Vector x( 5, 1.0 ), b( 5 );
      A->vectorProduct( x, b );
      x = 0.0;
      g.setMatrix( A );
      g.setConvergenceResidue( 1e-12 );
      EXPECT_EQ( g.solve( b, x ), true );
      std::cout << "GMRES(MGSR): " << g.getIterations() << " iterations\n";
   }
   solveExample< Solvers::Linear::Jacobi< Matrix > >( A, "Jacobi" );
   solveExample< Solvers::Linear::BICGStab< Matrix > >( A, "BICGStab" );
   solveExample< Solvers::Linear::BICGStab< Matrix > >( A, "BICGStab + Diagonal", true );
   solveExample< Solvers::Linear::TFQMR< Matrix > >( A, "TFQMR" );
   solveExample< Solvers::Linear::TFQMR< Matrix > >( A, "TFQMR + Diagonal", true );
   solveExample< Solvers::Linear::BICGStabL< Matrix > >( A, "BICGStabL" );
   solveExample< Solvers::Linear::BICGStabL< Matrix > >( A, "BICGStabL + Diagonal", true );
   solveExample< Solvers::Linear::IDRs< Matrix > >( A, "IDRs" );
   solveExample< Solvers::Linear::IDRs< Matrix > >( A, "IDRs + Diagonal", true );

   // copySparseMatrix into BiEllpack and ChunkedEllpack (column-major) and back to CSR
   {
      using BMatrix = Matrices::SparseMatrix< double, Device, int, Matrices::GeneralMatrix, Seg::BiEllpack >;
      using CMatrix = Matrices::SparseMatrix< double, Device, int, Matrices::GeneralMatrix, Seg::ChunkedEllpack >;
      BMatrix Bm;
      CMatrix Cm;
      Matrix Am;
      Matrices::copySparseMatrix( Bm, *A );
      Matrices::copySparseMatrix( Cm, Bm );
      Matrices::copySparseMatrix( Am, Cm );
      Vector x( 5 ), y( 5 ), z( 5 ), w( 5 );
      for( int i = 0; i < 5; i++ )
         x.setElement( i, 1.0 + i );
      A->vectorProduct( x, y );
      Cm.vectorProduct( x, z );
      Am.vectorProduct( x, w );
      EXPECT_EQ( y == z, true );
      EXPECT_EQ( y == w, true );
      EXPECT_EQ( Cm.getElement( 2, 3 ), -1.0 );
      // the chain's row lengths count each padded segment's slots (SparseOperations.hpp:
      // 544-597 via row.getSize()): TNL's own Host chain gives row 0 a capacity of 55 too
      // (3 in BiEllpack, then 55 in ChunkedEllpack, kept by the copy back to CSR)
      EXPECT_EQ( Bm.getRowCapacity( 0 ), 3 );
      EXPECT_EQ( Cm.getRowCapacity( 0 ), 55 );
      EXPECT_EQ( Am.getRowCapacity( 0 ), 55 );
      std::cout << "copySparseMatrix CSR -> BiEllpack -> ChunkedEllpack -> CSR done\n";
   }

   // MatrixReader::readMtx: the same matrix as a symmetric MatrixMarket file
   // (lower triangle, one repeated entry whose last value counts), into
   // SlicedEllpack, then the CG example on it
   {
      using SMatrix = Matrices::SparseMatrix< double, Device, int, Matrices::GeneralMatrix, Seg::SlicedEllpack >;
      std::stringstream mtx( "%%MatrixMarket matrix coordinate real symmetric\n% tridiag(-1, 2.5, -1)\n5 5 10\n"
                             "1 1 2.5\n2 1 -1\n2 2 2.5\n3 2 -1\n3 3 9\n3 3 2.5\n4 3 -1\n4 4 2.5\n5 4 -1\n"
                             "5 5 2.5\n" );
      auto B = std::make_shared< SMatrix >();
      Matrices::MatrixReader< SMatrix >::readMtx( mtx, *B );
      EXPECT_EQ( B->getRows(), 5 );
      EXPECT_EQ( B->getColumns(), 5 );
      Vector x( 5, 1.0 ), y( 5 ), z( 5 );
      B->vectorProduct( x, y );
      A->vectorProduct( x, z );
      EXPECT_EQ( y == z, true );
      solveExample< Solvers::Linear::CG< SMatrix > >( B, "CG on readMtx (SlicedEllpack)" );
      bool thrown = false;
      try {
         std::stringstream bad( "%%MatrixMarket matrix coordinate complex general\n2 2 1\n1 1 1 0\n" );
         Matrices::MatrixReader< SMatrix >::readMtx( bad, *B );
      }
      catch( const std::runtime_error& ) {
         thrown = true;
      }
      EXPECT_EQ( thrown, true );
   }

   // SparseMatrix::save / load: BiEllpack written, read back into a new
   // matrix, the same product; a file of another type is refused
   {
      using BMatrix = Matrices::SparseMatrix< double, Device, int, Matrices::GeneralMatrix, Seg::BiEllpack >;
      BMatrix Bm;
      Matrices::copySparseMatrix( Bm, *A );
      const std::string path = "/tmp/tnlmi_test_tnl_api_" + std::to_string( ::getpid() ) + ".tnl";
      Bm.save( path );
      BMatrix Lm;
      Lm.load( path );
      EXPECT_EQ( Lm.getRows(), 5 );
      EXPECT_EQ( Lm.getElement( 2, 3 ), -1.0 );
      const std::string type = BMatrix::getSerializationType();
      if( type.find( "BiEllpack< int, ColumnMajorOrder, 32 >" ) == std::string::npos )
         std::cerr << "serialization type: " << type << "\n";
      EXPECT_EQ( type.find( "BiEllpack< int, ColumnMajorOrder, 32 >" ) != std::string::npos, true );
      Vector x( 5 ), y( 5 ), z( 5 );
      for( int i = 0; i < 5; i++ )
         x.setElement( i, 1.0 + i );
      Bm.vectorProduct( x, y );
      Lm.vectorProduct( x, z );
      EXPECT_EQ( y == z, true );
      bool thrown = false;
      try {
         Matrix Cm;
         Cm.load( path );
      }
      catch( const std::runtime_error& e ) {
         thrown = std::string( e.what() ).find( "object type does not match" ) != std::string::npos;
      }
      EXPECT_EQ( thrown, true );
      std::remove( path.c_str() );
      std::cout << "SparseMatrix save / load (BiEllpack) done\n";
   }

   // SymmetricMatrix: the same tridiagonal matrix from its complete element
   // list (the upper entries checked against the lower, then dropped), the
   // same product and the CG example; transposedVectorProduct of A is A's
   // product (A is symmetric)
   {
      using SymMatrix = Matrices::SparseMatrix< double, Device, int, Matrices::SymmetricMatrix >;
      auto S = std::make_shared< SymMatrix >(
         5, 5,
         std::initializer_list< std::tuple< int, int, double > >{
            { 0, 0, 2.5 }, { 0, 1, -1 }, { 1, 0, -1 }, { 1, 1, 2.5 }, { 1, 2, -1 }, { 2, 1, -1 }, { 2, 2, 2.5 },
            { 2, 3, -1 }, { 3, 2, -1 }, { 3, 3, 2.5 }, { 3, 4, -1 }, { 4, 3, -1 }, { 4, 4, 2.5 } } );
      EXPECT_EQ( S->getRowCapacity( 0 ), 1 );
      EXPECT_EQ( S->getRowCapacity( 2 ), 2 );
      EXPECT_EQ( S->getElement( 2, 3 ), -1.0 );
      Vector x( 5 ), y( 5 ), z( 5 ), w( 5 );
      for( int i = 0; i < 5; i++ )
         x.setElement( i, 1.0 + i );
      S->vectorProduct( x, y );
      A->vectorProduct( x, z );
      A->transposedVectorProduct( x, w );
      EXPECT_EQ( y == z, true );
      EXPECT_EQ( w == z, true );
      solveExample< Solvers::Linear::CG< SymMatrix > >( S, "CG on SymmetricMatrix" );
      bool thrown = false;
      try {
         SymMatrix bad( 2, 2, std::initializer_list< std::tuple< int, int, double > >{ { 0, 1, 1.0 }, { 1, 0, 2.0 } } );
      }
      catch( const std::logic_error& ) {
         thrown = true;
      }
      EXPECT_EQ( thrown, true );
      std::cout << "SymmetricMatrix product, CG and transposedVectorProduct done\n";
   }

   // error behaviour of the reference
   {
      Vector in( 3, 1.0 ), out( 5 );
      bool thrown = false;
      try {
         A->vectorProduct( in, out );
      }
      catch( const std::invalid_argument& ) {
         thrown = true;
      }
      EXPECT_EQ( thrown, true );
      thrown = false;
      try {
         Seg::LaunchConfiguration cfg( Seg::ThreadsToSegmentsMapping::Fixed, 3 );
      }
      catch( const std::invalid_argument& ) {
         thrown = true;
      }
      EXPECT_EQ( thrown, true );
      thrown = false;
      try {
         Solvers::Linear::CG< Matrix > cg;
         cg.solve( out, out );
      }
      catch( const std::logic_error& ) {
         thrown = true;
      }
      EXPECT_EQ( thrown, true );
   }
   if( failures ) {
      std::cout << failures << " FAILURES\n";
      return 1;
   }
   std::cout << "ALL OK\n";
   return 0;
}
