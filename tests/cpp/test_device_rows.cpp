// Rows and elements of a Devices::MI355X matrix walked on the GPU from code
// compiled by hipcc, as the reference's own examples do on its GPU device
// (Documentation/Examples/Matrices/SparseMatrix/SparseMatrixViewExample_getRow.cpp,
// SparseMatrixExample_forRows.cpp, SparseMatrixExample_forElements.cpp):
// a 2D Laplacian assembled row by row inside Algorithms::parallelFor through
// getView().getRow( i ).setElement, then scaled through forAllRows and
// forAllElements, and read back through the const traversals -- every format,
// Index = int and long.  Each result is compared bit for bit with the same
// matrix built from the host side through setElement.
// Prints "ALL OK" and exits 0 when every check passes.
#include <TNL/Algorithms/parallelFor.h>
#include <TNL/Containers/Vector.h>
#include <TNL/Matrices/SparseMatrix.h>

#include <TNL/MI355X.h>

#include <cstdio>
#include <iostream>
#include <map>
#include <vector>

using namespace TNL;
using Device = Devices::MI355X;
namespace Seg = Algorithms::Segments;
template< typename D, typename I, typename A >
using SlicedEllpack64 = Seg::SlicedEllpack< D, I, A, Seg::ColumnMajorOrder, 64 >;
template< typename D, typename I, typename A >
using EllpackRowMajor = Seg::Ellpack< D, I, A, Seg::RowMajorOrder >;
template< typename D, typename I, typename A >
using SortedCSR = Seg::SortedSegments< Seg::CSR< D, I, A > >;

static int failures = 0;
#define EXPECT( cond, what )                                                                 \
   do {                                                                                      \
      if( !( cond ) ) {                                                                      \
         std::cerr << __FILE__ << ":" << __LINE__ << ": " << what << " failed: " #cond "\n"; \
         failures++;                                                                         \
      }                                                                                      \
   } while( 0 )

template< template< class, class, class > class Segments, typename IndexT >
void laplacianRows( const char* name )
{
   using Matrix = Matrices::SparseMatrix< double, Device, IndexT, Matrices::GeneralMatrix, Segments >;
   using Vector = Containers::Vector< double, Device, IndexT >;
   const IndexT m = 37, n = m * m;  // 1,369 rows: not a multiple of a slice or a strip
   // the same matrix from the host side
   std::map< std::pair< IndexT, IndexT >, double > entries;
   for( IndexT i = 0; i < n; i++ ) {
      const IndexT gx = i % m, gy = i / m;
      if( gy > 0 )
         entries[ { i, i - m } ] = -1.0;
      if( gx > 0 )
         entries[ { i, i - 1 } ] = -1.0;
      entries[ { i, i } ] = 4.0;
      if( gx < m - 1 )
         entries[ { i, i + 1 } ] = -1.0;
      if( gy < m - 1 )
         entries[ { i, i + m } ] = -1.0;
   }
   // (the same row capacities as below: a format's slot layout, and with it
   // a ChunkedEllpack row's summation order, depends on them)
   Matrix B( n, n );
   B.setRowCapacities( std::vector< IndexT >( n, 5 ) );
   for( const auto& [ rc, v ] : entries )
      B.setElement( rc.first, rc.second, v );

   // assembled on the device: five slots per row, filled in column order
   Matrix A( n, n );
   A.setRowCapacities( std::vector< IndexT >( n, 5 ) );
   auto view = A.getView();
   auto assemble = [ = ] __cuda_callable__( IndexT i ) mutable
   {
      auto row = view.getRow( i );
      const IndexT gx = i % m, gy = i / m;
      IndexT k = 0;
      if( gy > 0 )
         row.setElement( k++, i - m, -1.0 );
      if( gx > 0 )
         row.setElement( k++, i - 1, -1.0 );
      row.setElement( k++, i, 4.0 );
      if( gx < m - 1 )
         row.setElement( k++, i + 1, -1.0 );
      if( gy < m - 1 )
         row.setElement( k++, i + m, -1.0 );
   };
   Algorithms::parallelFor< Device >( (IndexT) 0, n, assemble );

   std::vector< double > hx( n );
   for( IndexT i = 0; i < n; i++ )
      hx[ i ] = std::sin( 0.37 * i ) + 0.01 * ( i % 7 );
   Vector x;
   x = hx;
   Vector ya( n ), yb( n );
   auto product = [ & ]( const Matrix& M, Vector& y ) {
      y.setValue( 0.0 );
      M.vectorProduct( x, y );
      return y.toHost();
   };
   const std::vector< double > want = product( B, yb );
   EXPECT( product( A, ya ) == want, std::string( name ) + ": getRow().setElement assembly" );
   EXPECT( A.getNonzeroElementsCount() == B.getNonzeroElementsCount(), std::string( name ) + ": nonzero count" );

   // forAllRows: every stored value doubled through its RowView
   A.forAllRows(
      [] __cuda_callable__( typename Matrix::RowView & row )
      {
         for( IndexT k = 0; k < row.getSize(); k++ )
            if( row.getColumnIndex( k ) >= 0 )
               row.setValue( k, 2.0 * row.getValue( k ) );
      } );
   std::vector< double > twice = product( A, ya );
   bool ok = true;
   for( IndexT i = 0; i < n; i++ )
      ok = ok && twice[ i ] == 2.0 * want[ i ];
   EXPECT( ok, std::string( name ) + ": forAllRows setValue" );

   // forAllElements: halved again through the value reference; a column
   // written through the reference is stored (moved and moved back)
   A.forAllElements(
      [] __cuda_callable__( IndexT rowIdx, IndexT localIdx, IndexT & column, double& value )
      {
         if( column >= 0 )
            value *= 0.5;
      } );
   EXPECT( product( A, ya ) == want, std::string( name ) + ": forAllElements value reference" );
   A.forElements( (IndexT) 0, n,
                  [] __cuda_callable__( IndexT rowIdx, IndexT localIdx, IndexT & column, double& value )
                  {
                     if( column == rowIdx )
                        column = -2 - rowIdx;  // off the matrix: the diagonal no longer counts
                  } );
   A.forElements( (IndexT) 0, n,
                  [] __cuda_callable__( IndexT rowIdx, IndexT localIdx, IndexT & column, double& value )
                  {
                     if( column == -2 - rowIdx )
                        column = rowIdx;
                  } );
   EXPECT( product( A, ya ) == want, std::string( name ) + ": forElements column reference" );

   // const traversals: row sums through ConstRowView, slot counts through
   // the const element walk, into device arrays
   Vector sums( n ), filled( n );
   double* ps = sums.getData();
   double* pf = filled.getData();
   const Matrix& C = A;
   C.forAllRows(
      [ = ] __cuda_callable__( const typename Matrix::ConstRowView& row )
      {
         double s = 0;
         for( IndexT k = 0; k < row.getSize(); k++ )
            if( row.getColumnIndex( k ) >= 0 )
               s += row.getValue( k );
         ps[ row.getRowIndex() ] = s;
      } );
   filled.setValue( 0.0 );
   C.forElements( (IndexT) 0, n,
                  [ = ] __cuda_callable__( IndexT rowIdx, IndexT localIdx, const IndexT& column, const double& value )
                  {
                     if( column >= 0 )
                        pf[ rowIdx ] += 1.0;
                  } );
   const std::vector< double > hs = sums.toHost(), hf = filled.toHost();
   ok = true;
   for( IndexT i = 0; i < n; i++ ) {
      const IndexT gx = i % m, gy = i / m;
      const double neighbours = ( gy > 0 ) + ( gx > 0 ) + ( gx < m - 1 ) + ( gy < m - 1 );
      ok = ok && hs[ i ] == 4.0 - neighbours && hf[ i ] == 1.0 + neighbours;
   }
   EXPECT( ok, std::string( name ) + ": const forAllRows / forElements" );

   // a row range outside the matrix is std::invalid_argument, as the reference's range checks
   bool threw = false;
   try {
      A.forRows( (IndexT) 0, n + 1, [] __cuda_callable__( typename Matrix::RowView & row ) {} );
   }
   catch( const std::invalid_argument& ) {
      threw = true;
   }
   EXPECT( threw, std::string( name ) + ": row range check" );
   std::printf( "%s: ok\n", name );
}

int main()
{
   laplacianRows< Seg::CSR, int >( "CSR<int>" );
   laplacianRows< Seg::CSR, long >( "CSR<long>" );
   laplacianRows< Seg::Ellpack, int >( "Ellpack<int>" );
   laplacianRows< EllpackRowMajor, long >( "Ellpack row-major<long>" );
   laplacianRows< SlicedEllpack64, int >( "SlicedEllpack64<int>" );
   laplacianRows< Seg::BiEllpack, int >( "BiEllpack<int>" );
   laplacianRows< Seg::ChunkedEllpack, long >( "ChunkedEllpack<long>" );
   laplacianRows< SortedCSR, int >( "SortedSegments<CSR><int>" );
   if( failures ) {
      std::printf( "%d FAILED\n", failures );
      return 1;
   }
   std::printf( "ALL OK\n" );
   return 0;
}
