// matrix_file.hip -- SparseMatrix::save / load in TNL's binary file format
// (Matrices/SparseMatrix.hpp:642-694, SparseMatrixBase.hpp:1149-1159), so
// that a file written by TNL loads here and one written here loads in TNL.
//
// Layout (File.hpp: raw little-endian values; Object.hpp:88-92 object types):
//   "TNLMN" + int32 length + type string of the matrix, e.g.
//     "Matrices::SparseMatrix< double, CSR< int >, [any_device], int, General,
//      [any_allocator], [any_allocator] >"
//   size_t rows, size_t columns
//   the segments' own fields and arrays (SegmentsFileData, per format)
//   the values, then the column indexes, each an array:
//   "TNLMN" + "TNL::Containers::Array< <element> >" + size_t count + raw data
//     (Containers/Array.hpp:583-620).
// Errors carry the reference's messages (Exceptions/FileDeserializationError.h,
// File.hpp:30-45) as TNLMI_RUNTIME_ERROR.
//
// A file is read into host memory whole (tnlmi_matrix_file) and checked before
// any of it reaches the device: every index its segments define must lie in
// the storage and every column index in [-1, columns).
#include <algorithm>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "segments.hpp"

struct tnlmi_matrix_file
{
   std::string type;
   int64_t rows = 0, columns = 0;
   tnlmi_format format = TNLMI_CSR;
   tnlmi_organization org = TNLMI_COLUMN_MAJOR;
   int32_t param = 0;
   tnlmi_dtype value_type = TNLMI_F64;
   tnlmi_matrix_type matrix_type = TNLMI_GENERAL;
   tnlmi::SegmentsFileData seg;
   std::vector< char > values;
   std::vector< int32_t > cols;
};

namespace tnlmi {
namespace {

constexpr char kMagic[ 5 ] = { 'T', 'N', 'L', 'M', 'N' };
const char* const kIntArray = "int";
const char* const kBlockType = "TNL::Algorithms::Segments::detail::CSRAdaptiveKernelBlockDescriptor<int>";
const char* const kSliceType = "TNL::Algorithms::Segments::detail::ChunkedEllpackSliceInfo<int>";
constexpr int kBlockBytes = 20;  // uint8 type, 3 padding, int32 first, size, segments, uint8 warp, warps, 2 padding

std::string organization_name( tnlmi_organization o )
{
   return o == TNLMI_ROW_MAJOR ? "RowMajorOrder" : "ColumnMajorOrder";  // ElementsOrganization.h:36-42
}

// <Segments>View::getSerializationType() of each format
std::string segments_type( tnlmi_format f, tnlmi_organization o, int32_t param )
{
   switch( f ) {
      case TNLMI_CSR:
         return "CSR< int >";  // CSRBase.hpp:40-43
      case TNLMI_ADAPTIVE_CSR:
         return "AdaptiveCSR< int >";  // AdaptiveCSRView.hpp:84-87
      case TNLMI_ELLPACK:
         return "Ellpack< int, " + organization_name( o ) + ", 32 >";  // EllpackBase.hpp:38-42
      case TNLMI_SLICED_ELLPACK:
         return "SlicedEllpack< int, " + organization_name( o ) + ", " + std::to_string( param ) + " >";
      case TNLMI_BIELLPACK:
         return "BiEllpack< int, " + organization_name( o ) + ", " + std::to_string( param ) + " >";
      case TNLMI_CHUNKED_ELLPACK:
         return "ChunkedEllpack< int, " + organization_name( o ) + " >";  // ChunkedEllpackBase.hpp:66-69
      case TNLMI_SORTED_CSR:
         return "SortedSegments< CSR< int > >";  // SortedSegmentsBase.hpp:15-18
   }
   return "";
}

std::string value_name( tnlmi_dtype t )
{
   return t == TNLMI_F64 ? "double" : "float";
}

// SparseMatrixBase::getSerializationType (SparseMatrixBase.hpp:46-53;
// MatrixType.h: "General" / "Symmetric")
std::string matrix_type( tnlmi_dtype t, tnlmi_format f, tnlmi_organization o, int32_t param, tnlmi_matrix_type mt )
{
   return "Matrices::SparseMatrix< " + value_name( t ) + ", " + segments_type( f, o, param ) + ", [any_device], int, "
        + ( mt == TNLMI_SYMMETRIC ? "Symmetric" : "General" ) + ", [any_allocator], [any_allocator] >";
}

std::string array_type( const std::string& element )
{
   return "TNL::Containers::Array< " + element + " >";
}

struct Writer
{
   std::ofstream out;
   std::string name;
   explicit Writer( const std::string& path ) : name( path )
   {
      out.open( path, std::ios::out | std::ios::binary | std::ios::trunc );
      if( ! out )
         fail( TNLMI_RUNTIME_ERROR, "Unable to open file " + path + "  for writing." );
   }
   void bytes( const void* p, size_t n )
   {
      if( n )
         out.write( (const char*) p, (std::streamsize) n );
      if( ! out )
         fail( TNLMI_RUNTIME_ERROR,
               "Failed to serialize an object into the file '" + name + "': unable to write " + std::to_string( n )
                  + " bytes." );
   }
   template< typename T >
   void scalar( T v )
   {
      bytes( &v, sizeof( v ) );
   }
   void object_type( const std::string& t )
   {
      bytes( kMagic, sizeof( kMagic ) );
      scalar< int32_t >( (int32_t) t.size() );
      bytes( t.data(), t.size() );
   }
   void array( const std::string& element, const void* data, uint64_t count, size_t element_bytes )
   {
      object_type( array_type( element ) );
      scalar< uint64_t >( count );
      bytes( data, count * element_bytes );
   }
   void ints( const std::vector< int32_t >& a )
   {
      array( kIntArray, a.data(), a.size(), sizeof( int32_t ) );
   }
};

struct Reader
{
   std::ifstream in;
   std::string name;
   uint64_t left = 0;
   explicit Reader( const std::string& path ) : name( path )
   {
      in.open( path, std::ios::in | std::ios::binary );
      if( ! in )
         fail( TNLMI_RUNTIME_ERROR, "Unable to open file " + path + "  for reading." );
      in.seekg( 0, std::ios::end );
      left = (uint64_t) in.tellg();
      in.seekg( 0, std::ios::beg );
   }
   [[noreturn]] void error( const std::string& details )
   {
      fail( TNLMI_RUNTIME_ERROR, "Failed to deserialize an object from the file '" + name + "': " + details );
   }
   void bytes( void* p, uint64_t n, const char* what )
   {
      if( n > left )
         error( std::string( "unable to read " ) + what + " (the file ends early)." );
      if( n )
         in.read( (char*) p, (std::streamsize) n );
      if( ! in )
         error( std::string( "unable to read " ) + what + "." );
      left -= n;
   }
   template< typename T >
   T scalar( const char* what )
   {
      T v;
      bytes( &v, sizeof( v ), what );
      return v;
   }
   std::string object_type()
   {
      char mn[ 5 ];
      bytes( mn, 5, "the magic number" );
      if( std::memcmp( mn, kMagic, 5 ) != 0 )
         error( "wrong magic number - file is not in a TNL-compatible format." );
      const int32_t len = scalar< int32_t >( "string length" );
      if( len < 0 || (uint64_t) len > left )
         error( "unable to read a C-string." );
      std::string t( (size_t) len, '\0' );
      bytes( t.data(), (uint64_t) len, "a C-string" );
      return t;
   }
   // an array's type check and count; returns the count
   uint64_t array_head( const std::string& element, size_t element_bytes )
   {
      const std::string t = object_type(), want = array_type( element );
      if( t != want )
         error( "object type does not match (expected " + want + ", found " + t + ")." );
      const uint64_t count = scalar< uint64_t >( "the array size" );
      if( count > left / element_bytes )
         error( "unable to read the array data (the file ends early)." );
      return count;
   }
   std::vector< int32_t > ints()
   {
      std::vector< int32_t > a( array_head( kIntArray, sizeof( int32_t ) ) );
      bytes( a.data(), a.size() * sizeof( int32_t ), "the array data" );
      return a;
   }
};

// the format, organization, parameter and value type named by a matrix type
// string (every type this library can hold; others are refused)
bool parse_matrix_type( const std::string& t, tnlmi_matrix_file& m )
{
   for( tnlmi_matrix_type mt : { TNLMI_GENERAL, TNLMI_SYMMETRIC } )
   for( tnlmi_dtype v : { TNLMI_F64, TNLMI_F32 } )
      for( int f = TNLMI_CSR; f <= TNLMI_SORTED_CSR; f++ )
         for( tnlmi_organization o : { TNLMI_COLUMN_MAJOR, TNLMI_ROW_MAJOR } ) {
            std::vector< int32_t > params{ 0 };
            if( f == TNLMI_SLICED_ELLPACK ) {
               params.clear();
               for( int32_t p = 1; p <= ( 1 << 20 ); p *= 2 )
                  params.push_back( p );
            }
            else if( f == TNLMI_BIELLPACK )
               params = { 32, 64 };
            for( int32_t p : params )
               if( t == matrix_type( v, (tnlmi_format) f, o, p, mt ) ) {
                  m.matrix_type = mt;
                  m.value_type = v;
                  m.format = (tnlmi_format) f;
                  m.org = o;
                  m.param = p;
                  return true;
               }
         }
   return false;
}

// <Segments>::save of each format
void save_segments( Writer& w, tnlmi_format format, int64_t n, const SegmentsFileData& f )
{
   switch( format ) {
      case TNLMI_CSR:  // CSR.hpp:116-121
         w.ints( f.offsets );
         break;
      case TNLMI_ADAPTIVE_CSR: {  // AdaptiveCSR.hpp:146-153: the CSR offsets, then 6 descriptor arrays
         w.ints( f.offsets );
         for( int i = 0; i < 6; i++ ) {
            const auto& b = f.blocks[ i ];
            const size_t count = b.size() / 6;
            std::vector< unsigned char > raw( count * kBlockBytes, 0 );
            for( size_t j = 0; j < count; j++ ) {
               unsigned char* d = raw.data() + j * kBlockBytes;
               const int32_t* e = b.data() + 6 * j;
               d[ 0 ] = (unsigned char) e[ 0 ];
               std::memcpy( d + 4, e + 1, 3 * sizeof( int32_t ) );
               d[ 16 ] = (unsigned char) e[ 4 ];
               d[ 17 ] = (unsigned char) e[ 5 ];
            }
            w.array( kBlockType, raw.data(), count, kBlockBytes );
         }
         break;
      }
      case TNLMI_ELLPACK:  // Ellpack.hpp:105-111
         w.scalar< int32_t >( f.ell_width );
         w.scalar< int32_t >( f.ell_count );
         w.scalar< int32_t >( f.ell_aligned );
         break;
      case TNLMI_SLICED_ELLPACK:  // SlicedEllpack.hpp:166-175
         w.scalar< int32_t >( (int32_t) f.elements );
         w.scalar< int32_t >( (int32_t) f.storage );
         w.scalar< int32_t >( (int32_t) n );
         w.ints( f.slice_offsets );
         w.ints( f.slice_seg_sizes );
         break;
      case TNLMI_BIELLPACK:  // BiEllpack.hpp:222-229
         w.scalar< int32_t >( (int32_t) f.elements );
         w.scalar< int32_t >( (int32_t) f.storage );
         w.ints( f.perm );
         w.ints( f.group_ptrs );
         break;
      case TNLMI_CHUNKED_ELLPACK:  // ChunkedEllpack.hpp:240-252
         w.scalar< int32_t >( (int32_t) f.elements );
         w.scalar< int32_t >( (int32_t) f.storage );
         w.scalar< int32_t >( f.ch_nslices );
         w.scalar< int32_t >( f.chunks_in_slice );
         w.scalar< int32_t >( f.desired_chunk_size );
         w.ints( f.seg_to_chunk );
         w.ints( f.seg_to_slice );
         w.ints( f.chunks_to_seg );
         w.ints( f.seg_ptrs );
         w.array( kSliceType, f.ch_slices.data(), f.ch_slices.size() / 4, 4 * sizeof( int32_t ) );
         break;
      case TNLMI_SORTED_CSR:  // SortedSegments.hpp:242-248: the embedded CSR, then both permutations
         w.ints( f.offsets );
         w.ints( f.perm );
         w.ints( f.inv_perm );
         break;
   }
}

// <Segments>::load of each format (the storage size of the formats that do
// not save one follows from their fields)
void load_segments( Reader& r, tnlmi_format format, SegmentsFileData& f )
{
   switch( format ) {
      case TNLMI_CSR:
         f.offsets = r.ints();
         f.storage = f.offsets.empty() ? 0 : f.offsets.back();
         break;
      case TNLMI_ADAPTIVE_CSR:
         f.offsets = r.ints();
         f.storage = f.offsets.empty() ? 0 : f.offsets.back();
         for( int i = 0; i < 6; i++ ) {  // descriptors of TNL's own kernel: skipped, rebuilt for this device
            const uint64_t count = r.array_head( kBlockType, kBlockBytes );
            std::vector< unsigned char > raw( count * kBlockBytes );
            r.bytes( raw.data(), raw.size(), "the array data" );
         }
         break;
      case TNLMI_ELLPACK:
         f.ell_width = r.scalar< int32_t >( "the segment size" );
         f.ell_count = r.scalar< int32_t >( "the segments count" );
         f.ell_aligned = r.scalar< int32_t >( "the aligned size" );
         f.storage = (int64_t) f.ell_aligned * f.ell_width;
         break;
      case TNLMI_SLICED_ELLPACK:
         f.elements = r.scalar< int32_t >( "the size" );
         f.storage = r.scalar< int32_t >( "the storage size" );
         f.ell_count = r.scalar< int32_t >( "the segments count" );
         f.slice_offsets = r.ints();
         f.slice_seg_sizes = r.ints();
         break;
      case TNLMI_BIELLPACK:
         f.elements = r.scalar< int32_t >( "the size" );
         f.storage = r.scalar< int32_t >( "the storage size" );
         f.perm = r.ints();
         f.group_ptrs = r.ints();
         break;
      case TNLMI_CHUNKED_ELLPACK: {
         f.elements = r.scalar< int32_t >( "the size" );
         f.storage = r.scalar< int32_t >( "the storage size" );
         f.ch_nslices = r.scalar< int32_t >( "the number of slices" );
         f.chunks_in_slice = r.scalar< int32_t >( "the chunks in a slice" );
         f.desired_chunk_size = r.scalar< int32_t >( "the desired chunk size" );
         f.seg_to_chunk = r.ints();
         f.seg_to_slice = r.ints();
         f.chunks_to_seg = r.ints();
         f.seg_ptrs = r.ints();
         const uint64_t count = r.array_head( kSliceType, 4 * sizeof( int32_t ) );
         f.ch_slices.resize( 4 * count );
         r.bytes( f.ch_slices.data(), 4 * count * sizeof( int32_t ), "the array data" );
         break;
      }
      case TNLMI_SORTED_CSR:
         f.offsets = r.ints();
         f.perm = r.ints();
         f.inv_perm = r.ints();
         f.storage = f.offsets.empty() ? 0 : f.offsets.back();
         break;
   }
}

size_t value_bytes( tnlmi_dtype t )
{
   return t == TNLMI_F64 ? 8 : t == TNLMI_F32 ? 4 : 2;
}

}  // namespace
}  // namespace tnlmi

using namespace tnlmi;

extern "C" {

tnlmi_status tnlmi_matrix_serialization_type( const tnlmi_matrix* A, char* buffer, int64_t capacity, int64_t* length )
{
   return guarded( [ & ] {
      if( ! A || ! A->segments )
         fail( TNLMI_INVALID_ARGUMENT, "matrix_serialization_type: null argument" );
      if( A->value_type == TNLMI_BF16 )
         fail( TNLMI_NOT_IMPLEMENTED, "matrix serialization: bf16 values have no TNL serialization type" );
      const tnlmi_segments& s = *A->segments;
      const std::string t = matrix_type( A->value_type, s.format, s.org, s.param, (tnlmi_matrix_type) A->matrix_type );
      if( length )
         *length = (int64_t) t.size();
      if( buffer && capacity > 0 ) {
         const size_t k = std::min< size_t >( t.size(), (size_t) capacity - 1 );
         std::memcpy( buffer, t.data(), k );
         buffer[ k ] = '\0';
      }
   } );
}

tnlmi_status tnlmi_matrix_save( const tnlmi_matrix* A, const char* path, tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( ! A || ! A->segments || ! path )
         fail( TNLMI_INVALID_ARGUMENT, "matrix_save: null argument" );
      if( A->value_type == TNLMI_BF16 )
         fail( TNLMI_NOT_IMPLEMENTED, "matrix serialization: bf16 values have no TNL serialization type" );
      const tnlmi_segments& seg = *A->segments;
      if( A->rows != seg.n )
         fail( TNLMI_INVALID_ARGUMENT, "matrix_save: matrix rows do not match the segments count" );
      if( seg.storage > 0 && ( ! A->values || ! A->column_indexes ) )
         fail( TNLMI_INVALID_ARGUMENT, "matrix_save: the storage arrays are NULL" );
      hipStream_t s = as_stream( stream );
      SegmentsFileData f;
      segments_export( seg, f, s );
      const size_t vb = value_bytes( A->value_type );
      std::vector< char > vals( (size_t) seg.storage * vb );
      std::vector< int32_t > cols( (size_t) seg.storage );
      if( seg.storage > 0 ) {
         TNLMI_HIP( hipMemcpyAsync( vals.data(), A->values, vals.size(), hipMemcpyDeviceToHost, s ) );
         TNLMI_HIP( hipMemcpyAsync( cols.data(), A->column_indexes, cols.size() * sizeof( int32_t ),
                                    hipMemcpyDeviceToHost, s ) );
      }
      TNLMI_HIP( hipStreamSynchronize( s ) );
      Writer w( path );
      w.object_type( matrix_type( A->value_type, seg.format, seg.org, seg.param, (tnlmi_matrix_type) A->matrix_type ) );
      w.scalar< uint64_t >( (uint64_t) A->rows );
      w.scalar< uint64_t >( (uint64_t) A->columns );
      save_segments( w, seg.format, seg.n, f );
      w.array( value_name( A->value_type ), vals.data(), (uint64_t) seg.storage, vb );
      w.ints( cols );
      w.out.close();
      if( ! w.out )
         fail( TNLMI_RUNTIME_ERROR, "Unable to close file " + std::string( path ) + "." );
   } );
}

tnlmi_status tnlmi_matrix_file_read( const char* path, tnlmi_matrix_file** out )
{
   return guarded( [ & ] {
      if( ! path || ! out )
         fail( TNLMI_INVALID_ARGUMENT, "matrix_file_read: null argument" );
      *out = nullptr;
      auto m = std::make_unique< tnlmi_matrix_file >();
      Reader r( path );
      m->type = r.object_type();
      if( ! parse_matrix_type( m->type, *m ) )
         fail( TNLMI_NOT_IMPLEMENTED, "matrix file '" + std::string( path ) + "': unsupported object type " + m->type );
      const uint64_t rows = r.scalar< uint64_t >( "the rows" ), columns = r.scalar< uint64_t >( "the columns" );
      if( rows > (uint64_t) INT32_MAX || columns > (uint64_t) INT32_MAX )
         r.error( "matrix dimensions out of the int32 index range." );
      m->rows = (int64_t) rows;
      m->columns = (int64_t) columns;
      load_segments( r, m->format, m->seg );
      if( m->format == TNLMI_SLICED_ELLPACK && m->seg.ell_count != m->rows )
         r.error( "the segments count differs from the matrix rows." );
      const size_t vb = value_bytes( m->value_type );
      const uint64_t nv = r.array_head( value_name( m->value_type ), vb );
      m->values.resize( nv * vb );
      r.bytes( m->values.data(), m->values.size(), "the array data" );
      m->cols = r.ints();
      if( (int64_t) nv != m->seg.storage || (int64_t) m->cols.size() != m->seg.storage )
         r.error( "the values or column indexes do not match the storage size of the segments." );
      for( int32_t c : m->cols )
         if( c < -1 || c >= m->columns )
            r.error( "a column index lies outside [-1, columns)." );
      *out = m.release();
   } );
}

tnlmi_status tnlmi_matrix_file_info( const tnlmi_matrix_file* m, int64_t* rows, int64_t* columns,
                                     tnlmi_format* format, tnlmi_organization* organization, int32_t* param,
                                     tnlmi_dtype* value_type, int64_t* storage, tnlmi_matrix_type* matrix_type )
{
   return guarded( [ & ] {
      if( ! m )
         fail( TNLMI_INVALID_ARGUMENT, "matrix_file_info: null argument" );
      if( rows )
         *rows = m->rows;
      if( columns )
         *columns = m->columns;
      if( format )
         *format = m->format;
      if( organization )
         *organization = m->org;
      if( param )
         *param = m->param;
      if( value_type )
         *value_type = m->value_type;
      if( storage )
         *storage = m->seg.storage;
      if( matrix_type )
         *matrix_type = m->matrix_type;
   } );
}

const char* tnlmi_matrix_file_type( const tnlmi_matrix_file* m )
{
   return m ? m->type.c_str() : "";
}

tnlmi_status tnlmi_matrix_file_segments( const tnlmi_matrix_file* m, tnlmi_stream stream, tnlmi_segments** out )
{
   return guarded( [ & ] {
      if( ! m || ! out )
         fail( TNLMI_INVALID_ARGUMENT, "matrix_file_segments: null argument" );
      *out = nullptr;
      auto seg = std::make_unique< tnlmi_segments >();
      seg->format = m->format;
      seg->org = m->org;
      seg->param = m->param;
      seg->n = m->rows;
      segments_import( *seg, m->seg, as_stream( stream ) );
      *out = seg.release();
   } );
}

tnlmi_status tnlmi_matrix_file_upload( const tnlmi_matrix_file* m, void* d_values, int32_t* d_column_indexes,
                                       tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( ! m || ( m->seg.storage > 0 && ( ! d_values || ! d_column_indexes ) ) )
         fail( TNLMI_INVALID_ARGUMENT, "matrix_file_upload: null argument" );
      hipStream_t s = as_stream( stream );
      if( m->seg.storage > 0 ) {
         TNLMI_HIP( hipMemcpyAsync( d_values, m->values.data(), m->values.size(), hipMemcpyHostToDevice, s ) );
         TNLMI_HIP( hipMemcpyAsync( d_column_indexes, m->cols.data(), m->cols.size() * sizeof( int32_t ),
                                    hipMemcpyHostToDevice, s ) );
      }
      TNLMI_HIP( hipStreamSynchronize( s ) );
   } );
}

void tnlmi_matrix_file_destroy( tnlmi_matrix_file* m )
{
   delete m;
}

}  // extern "C"
