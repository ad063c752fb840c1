// mtx.cpp -- MatrixMarket coordinate files -> host CSR, the input side of the
// SpMV path: MatrixReader< Matrix, Host >::readMtx (Matrices/MatrixReader.hpp:
// 52-66 header and dimensions, :158-209 header checks, :212-285 elements,
// :288-306 element lines), then SparseMatrix::setElements of the resulting
// map (row capacities = entries per row, columns ascending).
//
// Same rules as the reference: tokens are separated by spaces; lines starting
// with '%' are skipped; the first non-comment line holds exactly three tokens
// (rows, columns, entries -- the entry count is not used); an element line
// holds row, column and value (pattern files: no value, value 1); indexes are
// 1-based; a repeated (row, column) keeps the last value; a symmetric file's
// off-diagonal entries are mirrored.  Errors throw std::runtime_error with the
// reference's messages (TNLMI_RUNTIME_ERROR through the C ABI).  Indexes
// below 1 are rejected too (the reference does not check them).
//
// Parsing is a single pass over the file held in memory; the entries are then
// sorted by (row, column, position in the file), which gives the map's order
// and its last-value-wins rule without a tree.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"

struct tnlmi_host_csr
{
   int64_t rows = 0, columns = 0;
   std::vector< int32_t > offsets, cols;
   std::vector< double > values;
};

namespace tnlmi {
namespace {

// the tokens of one line, split on ' ' with empty tokens skipped
void split( const char* b, const char* e, std::vector< std::string >& out )
{
   out.clear();
   while( b < e ) {
      while( b < e && *b == ' ' )
         b++;
      const char* t = b;
      while( b < e && *b != ' ' )
         b++;
      if( b > t )
         out.emplace_back( t, b );
   }
}

void runtime( const std::string& msg )
{
   fail( TNLMI_RUNTIME_ERROR, msg );
}

void check_header( const std::string& header, bool& symmetric, bool& pattern )
{
   std::vector< std::string > p;
   split( header.data(), header.data() + header.size(), p );
   if( p.size() < 5 || p[ 0 ] != "%%MatrixMarket" )
      runtime( "Unknown format of the source file. We expect line like this: %%MatrixMarket matrix coordinate real "
               "general" );
   if( p[ 1 ] != "matrix" )
      runtime( "Keyword 'matrix' is expected in the header line: " + header );
   if( p[ 2 ] != "coordinates" && p[ 2 ] != "coordinate" )
      runtime( "Error: Only 'coordinates' format is supported now, not " + p[ 2 ] );
   if( p[ 3 ] != "real" && p[ 3 ] != "integer" && p[ 3 ] != "pattern" )
      runtime( "Only 'real' and 'integer' matrices are supported, not " + p[ 3 ] );
   pattern = p[ 3 ] == "pattern";
   if( p[ 4 ] != "general" ) {
      if( p[ 4 ] == "symmetric" )
         symmetric = true;
      else
         runtime( "Only 'general' matrices are supported, not " + p[ 4 ] );
   }
}

struct Entry
{
   int64_t key;  // row << 32 | column (0-based)
   int64_t seq;  // position in the file
   double value;
};

tnlmi_host_csr* parse( const char* text, int64_t length )
{
   const char* end = text + length;
   const char* cur = text;
   auto next_line = [ & ]( const char*& b, const char*& e ) {
      if( cur >= end )
         return false;
      b = cur;
      const char* nl = (const char*) std::memchr( cur, '\n', end - cur );
      e = nl ? nl : end;
      cur = nl ? nl + 1 : end;
      return true;
   };
   const char *b, *e;
   // the header: the first line
   if( ! next_line( b, e ) )
      runtime( "Unknown format of the source file. We expect line like this: %%MatrixMarket matrix coordinate real "
               "general" );
   std::string header( b, e );
   if( ! header.empty() && header.back() == '\r' )
      header.pop_back();
   bool symmetric = false, pattern = false;
   check_header( header, symmetric, pattern );
   // the dimensions: the first line that does not start with '%'
   std::vector< std::string > tok;
   int64_t rows = 0, columns = 0;
   bool have_dims = false;
   while( next_line( b, e ) ) {
      if( b == e || *b == '%' )
         continue;
      split( b, e, tok );
      if( tok.size() != 3 )
         runtime( "Wrong number of parameters in the matrix header - should be 3." );
      rows = std::atoi( tok[ 0 ].c_str() );
      columns = std::atoi( tok[ 1 ].c_str() );
      if( rows <= 0 || columns <= 0 )
         runtime( "Row or column index is negative." );
      have_dims = true;
      break;
   }
   if( ! have_dims )
      runtime( "Wrong number of parameters in the matrix header - should be 3." );
   // the elements
   std::vector< Entry > ent;
   int64_t seq = 0;
   const size_t want = pattern ? 2 : 3;
   while( next_line( b, e ) ) {
      if( b == e || *b == '%' )
         continue;
      split( b, e, tok );
      if( tok.size() != want )
         runtime( "Wrong number of parameters in the matrix row at line:" + std::string( b, e ) );
      const int64_t row = std::atoi( tok[ 0 ].c_str() );
      const int64_t col = std::atoi( tok[ 1 ].c_str() );
      const double value = pattern ? 1.0 : std::atof( tok[ 2 ].c_str() );
      if( col > columns || row > rows ) {
         std::stringstream s;
         s << "There is an element at position " << row << ", " << col << " out of the matrix dimensions " << rows
           << " x " << columns << ".";
         runtime( s.str() );
      }
      if( row < 1 || col < 1 ) {
         std::stringstream s;
         s << "There is an element at position " << row << ", " << col << " below the first row or column.";
         runtime( s.str() );
      }
      ent.push_back( { ( ( row - 1 ) << 32 ) | ( col - 1 ), seq++, value } );
      if( symmetric && row != col )
         ent.push_back( { ( ( col - 1 ) << 32 ) | ( row - 1 ), seq++, value } );
   }
   std::sort( ent.begin(), ent.end(), []( const Entry& x, const Entry& y ) {
      return x.key != y.key ? x.key < y.key : x.seq < y.seq;
   } );
   auto* M = new tnlmi_host_csr;
   M->rows = rows;
   M->columns = columns;
   M->offsets.assign( rows + 1, 0 );
   for( size_t i = 0; i < ent.size(); i++ ) {
      if( i + 1 < ent.size() && ent[ i + 1 ].key == ent[ i ].key )
         continue;  // a later value of the same position wins
      const int64_t r = ent[ i ].key >> 32;
      M->offsets[ r + 1 ]++;
      M->cols.push_back( (int32_t) ( ent[ i ].key & 0xffffffff ) );
      M->values.push_back( ent[ i ].value );
   }
   for( int64_t r = 0; r < rows; r++ )
      M->offsets[ r + 1 ] += M->offsets[ r ];
   return M;
}

}  // namespace
}  // namespace tnlmi

using namespace tnlmi;

extern "C" {

tnlmi_status tnlmi_mtx_read( const char* path, tnlmi_host_csr** out )
{
   return guarded( [ & ] {
      if( ! path || ! out )
         fail( TNLMI_INVALID_ARGUMENT, "mtx_read: null argument" );
      *out = nullptr;
      std::ifstream f( path, std::ios::binary );
      if( ! f )
         runtime( std::string( "I am not able to open the file " ) + path );
      std::string text( ( std::istreambuf_iterator< char >( f ) ), std::istreambuf_iterator< char >() );
      *out = parse( text.data(), (int64_t) text.size() );
   } );
}

tnlmi_status tnlmi_mtx_parse( const char* text, int64_t length, tnlmi_host_csr** out )
{
   return guarded( [ & ] {
      if( ! text || ! out || length < 0 )
         fail( TNLMI_INVALID_ARGUMENT, "mtx_parse: null argument or negative length" );
      *out = nullptr;
      *out = parse( text, length );
   } );
}

tnlmi_status tnlmi_host_csr_get( const tnlmi_host_csr* m, int64_t* rows, int64_t* columns, int64_t* nnz,
                                 const int32_t** offsets, const int32_t** cols, const double** values )
{
   return guarded( [ & ] {
      if( ! m )
         fail( TNLMI_INVALID_ARGUMENT, "host_csr_get: null matrix" );
      if( rows )
         *rows = m->rows;
      if( columns )
         *columns = m->columns;
      if( nnz )
         *nnz = (int64_t) m->cols.size();
      if( offsets )
         *offsets = m->offsets.data();
      if( cols )
         *cols = m->cols.data();
      if( values )
         *values = m->values.data();
   } );
}

void tnlmi_host_csr_destroy( tnlmi_host_csr* m )
{
   delete m;
}

}  // extern "C"
