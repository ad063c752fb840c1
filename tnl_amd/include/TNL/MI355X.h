// TNL/MI355X.h -- the MI355X device for TNL's SparseMatrix / Segments /
// Solvers::Linear API, over the C ABI of libtnlmi.so (include/tnlmi.h).
//
// A caller of the reference writes
//     TNL::Matrices::SparseMatrix< double, TNL::Devices::Host, int > A( ... );
//     TNL::Solvers::Linear::CG< decltype( A ) > cg;  cg.setMatrix( ptr ); cg.solve( b, x );
// and switches to the GPU by replacing Devices::Host with Devices::MI355X.
// This header provides that device's specialisations of the class templates
// the hot path uses, with the reference's names, argument meaning and
// exceptions (std::invalid_argument, std::overflow_error, std::runtime_error,
// std::logic_error, Exceptions::NotImplementedError / BackendRuntimeError).
// It is self-contained (it does not include the reference); inside a TNL tree
// the same specialisations are added next to Devices::Host.
//
// Every call is synchronous on return, as the reference's reduceSegments is
// (Algorithms/Segments/detail/ReducingOperations_CSR.h:327).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <initializer_list>
#include <iostream>
#include <iterator>
#include <limits>
#include <map>
#include <memory>
#include <ostream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <variant>
#include <vector>

#include "tnlmi.h"

namespace TNL {

namespace Exceptions {
struct NotImplementedError : std::runtime_error
{
   using std::runtime_error::runtime_error;
};
struct BackendRuntimeError : std::runtime_error
{
   using std::runtime_error::runtime_error;
};
struct BackendBadAlloc : std::bad_alloc
{
   const char* what() const noexcept override { return "BackendBadAlloc: device memory allocation failed"; }
};
}  // namespace Exceptions

namespace detail {
// status -> the exception the reference throws in that case
inline void check( tnlmi_status s )
{
   if( s == TNLMI_OK )
      return;
   const std::string m = tnlmi_last_error();
   switch( s ) {
      case TNLMI_INVALID_ARGUMENT:
         throw std::invalid_argument( m );
      case TNLMI_OVERFLOW_ERROR:
         throw std::overflow_error( m );
      case TNLMI_NOT_IMPLEMENTED:
         throw Exceptions::NotImplementedError( m );
      case TNLMI_BACKEND_ERROR:
         throw Exceptions::BackendRuntimeError( m );
      case TNLMI_BAD_ALLOC:
         throw Exceptions::BackendBadAlloc();
      case TNLMI_LOGIC_ERROR:
         throw std::logic_error( m );
      default:
         throw std::runtime_error( m );
   }
}
inline void hip( hipError_t e )
{
   if( e == hipErrorOutOfMemory )
      throw Exceptions::BackendBadAlloc();
   if( e != hipSuccess )
      throw Exceptions::BackendRuntimeError( hipGetErrorString( e ) );
}
template< typename T >
constexpr tnlmi_dtype dtype()
{
   static_assert( std::is_same_v< T, double > || std::is_same_v< T, float >, "FP64 or FP32 values" );
   return std::is_same_v< T, double > ? TNLMI_F64 : TNLMI_F32;
}
}  // namespace detail

namespace Devices {
//! The MI355X GPU (gfx950, wave64).
struct MI355X
{
   struct LaunchConfiguration
   {};
   static std::string getDeviceType() { return "Devices::MI355X"; }
};
struct Host;
}  // namespace Devices

// ------------------------------------------------------------------ vectors
namespace Containers {

template< typename Real, typename Device, typename Index = int >
class Vector;

//! Containers::Vector on the device: an owning, contiguous device array.
template< typename Real, typename Index >
class Vector< Real, Devices::MI355X, Index >
{
public:
   using RealType = Real;
   using ValueType = Real;
   using DeviceType = Devices::MI355X;
   using IndexType = Index;

   Vector() = default;
   explicit Vector( Index size ) { setSize( size ); }
   Vector( Index size, Real value )
   {
      setSize( size );
      setValue( value );
   }
   Vector( std::initializer_list< Real > list ) { *this = std::vector< Real >( list ); }
   Vector( const Vector& o ) { *this = o; }
   Vector( Vector&& o ) noexcept { swap( o ); }
   ~Vector() { reset(); }

   Vector& operator=( const Vector& o )
   {
      if( this != &o ) {
         setSize( o.size_ );
         if( size_ )
            detail::hip( hipMemcpy( data_, o.data_, sizeof( Real ) * size_, hipMemcpyDeviceToDevice ) );
      }
      return *this;
   }
   Vector& operator=( Vector&& o ) noexcept
   {
      swap( o );
      return *this;
   }
   Vector& operator=( const std::vector< Real >& h )
   {
      setSize( (Index) h.size() );
      if( size_ )
         detail::hip( hipMemcpy( data_, h.data(), sizeof( Real ) * size_, hipMemcpyHostToDevice ) );
      return *this;
   }
   Vector& operator=( Real value )
   {
      setValue( value );
      return *this;
   }

   void setSize( Index size )
   {
      if( size < 0 )
         throw std::invalid_argument( "Vector size must be non-negative" );
      if( size == size_ )
         return;
      reset();
      if( size > 0 )
         detail::hip( hipMalloc( (void**) &data_, sizeof( Real ) * size ) );
      size_ = size;
   }
   template< typename V >
   void setLike( const V& v )
   {
      setSize( v.getSize() );
   }
   void reset()
   {
      if( data_ )
         (void) hipFree( data_ );
      data_ = nullptr;
      size_ = 0;
   }
   Index getSize() const { return size_; }
   Real* getData() { return data_; }
   const Real* getData() const { return data_; }

   void setValue( Real value )
   {
      std::vector< Real > h( size_, value );
      if( size_ )
         detail::hip( hipMemcpy( data_, h.data(), sizeof( Real ) * size_, hipMemcpyHostToDevice ) );
   }
   Real getElement( Index i ) const
   {
      Real v;
      detail::hip( hipMemcpy( &v, data_ + i, sizeof( Real ), hipMemcpyDeviceToHost ) );
      return v;
   }
   void setElement( Index i, Real v ) { detail::hip( hipMemcpy( data_ + i, &v, sizeof( Real ), hipMemcpyHostToDevice ) ); }
   std::vector< Real > toHost() const
   {
      std::vector< Real > h( size_ );
      if( size_ )
         detail::hip( hipMemcpy( h.data(), data_, sizeof( Real ) * size_, hipMemcpyDeviceToHost ) );
      return h;
   }
   bool operator==( const Vector& o ) const { return toHost() == o.toHost(); }
   void swap( Vector& o ) noexcept
   {
      std::swap( data_, o.data_ );
      std::swap( size_, o.size_ );
   }

private:
   Real* data_ = nullptr;
   Index size_ = 0;
};

template< typename Real, typename Index >
std::ostream& operator<<( std::ostream& s, const Vector< Real, Devices::MI355X, Index >& v )
{
   const auto h = v.toHost();
   s << "[ ";
   for( size_t i = 0; i < h.size(); i++ )
      s << ( i ? ", " : "" ) << h[ i ];
   return s << " ]";
}

//! (a, b): the scalar product of Containers/Expressions/ExpressionTemplates.h:372-384
template< typename Index >
double operator,( const Vector< double, Devices::MI355X, Index >& a, const Vector< double, Devices::MI355X, Index >& b )
{
   double* d;
   detail::hip( hipMalloc( (void**) &d, sizeof( double ) ) );
   const tnlmi_status st = tnlmi_dot( a.getSize(), a.getData(), b.getData(), d, nullptr );
   double h = 0;
   if( st == TNLMI_OK )
      detail::hip( hipMemcpy( &h, d, sizeof( double ), hipMemcpyDeviceToHost ) );
   (void) hipFree( d );
   detail::check( st );
   return h;
}

template< typename Index >
double l2Norm( const Vector< double, Devices::MI355X, Index >& a )
{
   double* d;
   detail::hip( hipMalloc( (void**) &d, sizeof( double ) ) );
   const tnlmi_status st = tnlmi_nrm2( a.getSize(), a.getData(), d, nullptr );
   double h = 0;
   if( st == TNLMI_OK )
      detail::hip( hipMemcpy( &h, d, sizeof( double ), hipMemcpyDeviceToHost ) );
   (void) hipFree( d );
   detail::check( st );
   return h;
}

template< typename Index >
double lpNorm( const Vector< double, Devices::MI355X, Index >& a, double p )
{
   if( p != 2.0 )
      throw Exceptions::NotImplementedError( "lpNorm on MI355X is implemented for p = 2" );
   return l2Norm( a );
}

}  // namespace Containers

// ----------------------------------------------------------------- segments
namespace Algorithms::Segments {

enum ElementsOrganization : std::uint8_t { ColumnMajorOrder = 0, RowMajorOrder };

enum class ThreadsToSegmentsMapping : std::uint8_t { Fixed, Warp, Block, BlockMerged, DynamicGrouping };

//! Segments/LaunchConfiguration.h:18-91; the default-constructed
//! configuration selects the MI355X-tuned kernel of each format.
struct LaunchConfiguration
{
   LaunchConfiguration() = default;
   LaunchConfiguration( ThreadsToSegmentsMapping mapping, int threadsPerSegmentCount ) : mapping_( mapping ), set_( true )
   {
      setThreadsPerSegmentCount( threadsPerSegmentCount );
   }
   void setThreadsPerSegmentCount( int t )
   {
      if( t > 0 && ( t & ( t - 1 ) ) == 0 )
         tps_ = t;
      else
         throw std::invalid_argument( "Threads per segment count must be a positive power of two." );
      set_ = true;
   }
   void setThreadsToSegmentsMapping( ThreadsToSegmentsMapping m )
   {
      mapping_ = m;
      set_ = true;
   }
   int getThreadsPerSegmentCount() const { return tps_; }
   ThreadsToSegmentsMapping getThreadsToSegmentsMapping() const { return mapping_; }
   tnlmi_launch toABI() const
   {
      if( ! set_ )
         return { TNLMI_MAP_AUTO, 1 };
      return { (int32_t) mapping_, tps_ };
   }

private:
   ThreadsToSegmentsMapping mapping_ = ThreadsToSegmentsMapping::Fixed;
   int tps_ = 1;
   bool set_ = false;
};

// Segment format tags: the class templates of Algorithms/Segments with the
// reference's template parameters; each names its C ABI format.
template< typename Device, typename Index, typename IndexAllocator = void >
struct CSR
{
   static constexpr tnlmi_format format = TNLMI_CSR;
   static constexpr int param = 0;
   static constexpr tnlmi_organization organization = TNLMI_COLUMN_MAJOR;
};
template< typename Device, typename Index, typename IndexAllocator = void >
struct AdaptiveCSR
{
   static constexpr tnlmi_format format = TNLMI_ADAPTIVE_CSR;
   static constexpr int param = 0;
   static constexpr tnlmi_organization organization = TNLMI_COLUMN_MAJOR;
};
template< typename Device, typename Index, typename IndexAllocator = void,
          ElementsOrganization Organization = ColumnMajorOrder, int Alignment = 32 >
struct Ellpack
{
   static constexpr tnlmi_format format = TNLMI_ELLPACK;
   static constexpr int param = 0;
   static constexpr tnlmi_organization organization = (tnlmi_organization) Organization;
};
template< typename Device, typename Index, typename IndexAllocator = void,
          ElementsOrganization Organization = ColumnMajorOrder, int SliceSize = 32 >
struct SlicedEllpack
{
   static constexpr tnlmi_format format = TNLMI_SLICED_ELLPACK;
   static constexpr int param = SliceSize;
   static constexpr tnlmi_organization organization = (tnlmi_organization) Organization;
};
template< typename Device, typename Index, typename IndexAllocator = void,
          ElementsOrganization Organization = ColumnMajorOrder, int WarpSize = 32 >
struct BiEllpack
{
   static constexpr tnlmi_format format = TNLMI_BIELLPACK;
   static constexpr int param = WarpSize;
   static constexpr tnlmi_organization organization = (tnlmi_organization) Organization;
};
template< typename Device, typename Index, typename IndexAllocator = void,
          ElementsOrganization Organization = ColumnMajorOrder >
struct ChunkedEllpack
{
   static constexpr tnlmi_format format = TNLMI_CHUNKED_ELLPACK;
   static constexpr int param = 0;
   static constexpr tnlmi_organization organization = (tnlmi_organization) Organization;
};
template< typename Device, typename Index, typename IndexAllocator = void >
struct SortedCSR
{
   static constexpr tnlmi_format format = TNLMI_SORTED_CSR;
   static constexpr int param = 0;
   static constexpr tnlmi_organization organization = TNLMI_COLUMN_MAJOR;
};
// three-parameter aliases for SparseMatrix< ..., Segments > (the form the
// reference's template template parameter takes)
template< typename D, typename I, typename A >
using SlicedEllpack64 = SlicedEllpack< D, I, A, ColumnMajorOrder, 64 >;

}  // namespace Algorithms::Segments

// ------------------------------------------------------------------- matrix
namespace Matrices {

struct GeneralMatrix
{
   static constexpr bool isSymmetric() { return false; }
};
//! MatrixType.h: the storage holds the lower triangle and the diagonal
struct SymmetricMatrix
{
   static constexpr bool isSymmetric() { return true; }
};

template< typename Real, typename Device, typename Index = int, typename MatrixType = GeneralMatrix,
          template< typename, typename, typename > class Segments = Algorithms::Segments::CSR,
          typename ComputeReal = Real >
class SparseMatrix
{
   // the device kernels take int32 offsets and column indexes (TNL's default
   // Index = int); 64-bit indices are not provided
   static_assert( std::is_same_v< Index, int >, "SparseMatrix on Devices::MI355X supports Index = int only" );
};

//! Matrices::SparseMatrix on the MI355X (SparseMatrix.h:41-56), general
//! matrix, int32 indices, FP64 or FP32 values.
template< typename Real, typename MatrixType, template< typename, typename, typename > class Segments,
          typename ComputeReal >
class SparseMatrix< Real, Devices::MI355X, int, MatrixType, Segments, ComputeReal >
{
   using SegTag = Segments< Devices::MI355X, int, void >;

public:
   using RealType = Real;
   using DeviceType = Devices::MI355X;
   using IndexType = int;
   using ComputeRealType = ComputeReal;
   using VectorType = Containers::Vector< Real, Devices::MI355X, int >;
   using RowCapacitiesType = Containers::Vector< int, Devices::MI355X, int >;

   SparseMatrix() = default;
   SparseMatrix( int rows, int columns ) { setDimensions( rows, columns ); }
   //! SparseMatrix( rows, columns, { { row, column, value }, ... } ) (SparseMatrix.hpp:336-367)
   SparseMatrix( int rows, int columns, const std::initializer_list< std::tuple< int, int, Real > >& data )
   {
      setDimensions( rows, columns );
      std::map< std::pair< int, int >, Real > m;
      for( const auto& [ r, c, v ] : data )
         m[ { r, c } ] = v;
      setElements( m );
   }
   //! SparseMatrix( rowCapacities, columns ) (SparseMatrix.h): empty rows of the
   //! given capacities (setRowCapacities)
   SparseMatrix( const std::initializer_list< int >& rowCapacities, int columns )
   {
      setDimensions( (int) rowCapacities.size(), columns );
      setRowCapacities( std::vector< int >( rowCapacities ) );
   }
   SparseMatrix( const SparseMatrix& ) = delete;
   SparseMatrix& operator=( const SparseMatrix& ) = delete;
   ~SparseMatrix() { release(); }

   //! copySparseMatrix( *this, B ) (Matrices/SparseOperations.hpp:600-690) from
   //! a matrix of any format on the device: row capacities = B's row lengths
   //! (tnlmi_matrix_row_lengths), then element k of every row copied
   //! (tnlmi_matrix_copy)
   template< typename Matrix2 >
   void copyFrom( const Matrix2& B )
   {
      static_assert( std::is_same_v< Real, typename Matrix2::RealType >, "The matrices must have the same RealType." );
      static_assert( std::is_same_v< typename Matrix2::DeviceType, Devices::MI355X >,
                     "The matrices must be allocated on the same device." );
      release();
      rows_ = B.getRows();
      columns_ = B.getColumns();
      const tnlmi_matrix src = B.abi();
      int* d_len = nullptr;
      detail::hip( hipMalloc( (void**) &d_len, std::max( 1, rows_ ) * sizeof( int ) ) );
      std::unique_ptr< int, void ( * )( int* ) > own( d_len, []( int* p ) { (void) hipFree( p ); } );
      detail::check( tnlmi_matrix_row_lengths( &src, d_len, nullptr ) );
      detail::check( tnlmi_segments_create( SegTag::format, SegTag::organization, SegTag::param, d_len, rows_, nullptr,
                                            &segments_ ) );
      const int64_t storage = std::max< int64_t >( 1, tnlmi_segments_storage_size( segments_ ) );
      detail::hip( hipMalloc( (void**) &values_, storage * sizeof( Real ) ) );
      detail::hip( hipMalloc( (void**) &columnIndexes_, storage * sizeof( int ) ) );
      tnlmi_matrix dst = abi();
      detail::check( tnlmi_matrix_copy( &src, &dst, d_len, nullptr ) );
      mirror();
   }
   template< typename, typename, typename, typename, template< typename, typename, typename > class, typename >
   friend class SparseMatrix;

   void setDimensions( int rows, int columns )
   {
      if( rows < 0 || columns < 0 )
         throw std::invalid_argument( "setDimensions: negative dimensions" );
      rows_ = rows;
      columns_ = columns;
      std::vector< int > zero( rows, 0 );
      setRowCapacities( zero );
   }
   int getRows() const { return rows_; }
   int getColumns() const { return columns_; }

   //! setRowCapacities (SparseMatrix.hpp:301-327): builds the format's
   //! segments; values = 0, column indexes = -1 (padding)
   void setRowCapacities( const std::vector< int >& capacities )
   {
      if( (int) capacities.size() != rows_ )
         throw std::invalid_argument( "setRowCapacities: the number of row capacities differs from the rows" );
      std::vector< int > off( rows_ + 1, 0 );
      for( int r = 0; r < rows_; r++ )
         off[ r + 1 ] = off[ r ] + capacities[ r ];
      std::vector< int > cols( off[ rows_ ], -1 );
      std::vector< Real > vals( off[ rows_ ], Real( 0 ) );
      assemble( capacities, off, cols, vals );
   }
   void setRowCapacities( const RowCapacitiesType& capacities ) { setRowCapacities( capacities.toHost() ); }

   //! setElements( map ) (SparseMatrix.hpp:380-470, Complete encoding):
   //! capacities = entries per row; a symmetric matrix keeps the entries on
   //! and below the diagonal and checks that the others mirror them
   void setElements( const std::map< std::pair< int, int >, Real >& m )
   {
      std::vector< int > caps( rows_, 0 );
      for( const auto& [ rc, v ] : m ) {
         if( rc.first < 0 || rc.first >= rows_ || rc.second < 0 || rc.second >= columns_ )
            throw std::invalid_argument( "setElements: element out of the matrix" );
         if constexpr( MatrixType::isSymmetric() ) {
            const auto q = m.find( { rc.second, rc.first } );
            if( q != m.end() && q->second != v )
               throw std::logic_error( "SparseMatrix is configured as symmetric, but the input data is not symmetric. "
                                       "The matrix elements at position ("
                                       + std::to_string( rc.first ) + ", " + std::to_string( rc.second )
                                       + ") do not match." );
            if( rc.first < rc.second )
               continue;
         }
         caps[ rc.first ]++;
      }
      std::vector< int > off( rows_ + 1, 0 );
      for( int r = 0; r < rows_; r++ )
         off[ r + 1 ] = off[ r ] + caps[ r ];
      std::vector< int > cols( off[ rows_ ] ), pos( off.begin(), off.end() - 1 );
      std::vector< Real > vals( off[ rows_ ] );
      for( const auto& [ rc, v ] : m ) {  // std::map is ordered: columns ascend within a row
         if( MatrixType::isSymmetric() && rc.first < rc.second )
            continue;
         cols[ pos[ rc.first ] ] = rc.second;
         vals[ pos[ rc.first ]++ ] = v;
      }
      assemble( caps, off, cols, vals );
   }

   //! Assembly from CSR-ordered host arrays with row capacities taken from
   //! `offsets` (the device build of TNL fills a host matrix and copies it,
   //! SparseMatrix.hpp:360-367).
   void setFromCSR( const std::vector< int >& offsets, const std::vector< int >& columns, const std::vector< Real >& values )
   {
      std::vector< int > caps( rows_ );
      for( int r = 0; r < rows_; r++ )
         caps[ r ] = offsets[ r + 1 ] - offsets[ r ];
      assemble( caps, offsets, columns, values );
   }

   //! vectorProduct (SparseMatrixBase.hpp:267-391): out = a*A*in + b*out on [begin, end)
   template< typename InVector, typename OutVector >
   void vectorProduct( const InVector& in, OutVector& out, ComputeReal matrixMultiplicator = 1.0,
                       ComputeReal outVectorMultiplicator = 0.0, int begin = 0, int end = 0,
                       const Algorithms::Segments::LaunchConfiguration& launchConfig = {} ) const
   {
      if( getColumns() != in.getSize() )
         throw std::invalid_argument( "vectorProduct: size of the input vector does not match the number of matrix columns" );
      if( getRows() != out.getSize() )
         throw std::invalid_argument( "vectorProduct: size of the output vector does not match the number of matrix rows" );
      using OutReal = typename OutVector::RealType;
      const tnlmi_matrix m = abi();
      const tnlmi_launch cfg = launchConfig.toABI();
      detail::check( tnlmi_spmv( &m, detail::dtype< OutReal >(), in.getData(), out.getData(),
                                 detail::dtype< ComputeReal >(), matrixMultiplicator, outVectorMultiplicator, begin, end,
                                 &cfg, nullptr ) );
      detail::hip( hipDeviceSynchronize() );
   }
   template< typename InVector, typename OutVector >
   void vectorProduct( const InVector& in, OutVector& out, const Algorithms::Segments::LaunchConfiguration& cfg ) const
   {
      vectorProduct( in, out, 1.0, 0.0, 0, 0, cfg );
   }

   //! getElement( row, column ) (SparseMatrixBase.hpp:231-265), 0 if absent;
   //! a symmetric matrix reads (column, row) for row < column
   Real getElement( int row, int column ) const
   {
      const tnlmi_matrix m = abi();
      double v = 0;
      detail::check( tnlmi_matrix_get_element( &m, row, column, &v, nullptr ) );
      return (Real) v;
   }
   //! setElement( row, column, value ) (SparseMatrixBase.hpp:152-161)
   void setElement( int row, int column, const Real& value ) { addElement( row, column, value, Real( 0 ) ); }
   //! addElement( row, column, value, thisElementMultiplicator ) (SparseMatrixBase.hpp:166-231):
   //! an existing element becomes multiplicator * old + value; a new one is
   //! inserted in column order into the row's capacity (std::logic_error
   //! "The capacity of the sparse matrix row number r was exceeded." if full)
   void addElement( int row, int column, const Real& value, const Real& thisElementMultiplicator = 1.0 )
   {
      tnlmi_matrix m = abi();
      detail::check( tnlmi_matrix_add_element( &m, row, column, value, thisElementMultiplicator, nullptr ) );
   }
   //! getRowCapacity( row ): the row's segment size (Ellpack: the width)
   int getRowCapacity( int row ) const { return hoff_[ row + 1 ] - hoff_[ row ]; }
   //! getRowCapacities (SparseMatrixBase.hpp:75-92): stored (non-padding) slots per row
   void getRowCapacities( RowCapacitiesType& out ) const { rowCounts( 1, out ); }
   //! getCompressedRowLengths (SparseMatrixBase.hpp:55-73): stored nonzero values per row
   void getCompressedRowLengths( RowCapacitiesType& out ) const { rowCounts( 2, out ); }
   //! getNonzeroElementsCount (SparseMatrixBase.hpp:102-132): a symmetric matrix
   //! counts its stored off-diagonal entries twice
   int64_t getNonzeroElementsCount() const
   {
      RowCapacitiesType c;
      rowCounts( MatrixType::isSymmetric() ? 3 : 2, c );
      int64_t total = 0;
      for( int v : c.toHost() )
         total += v;
      return total;
   }
   int64_t getStorageSize() const { return segments_ ? tnlmi_segments_storage_size( segments_ ) : 0; }

   //! getSerializationType (SparseMatrixBase.hpp:46-53)
   static std::string getSerializationType()
   {
      tnlmi_segments* empty = nullptr;
      detail::check( tnlmi_segments_create( SegTag::format, SegTag::organization, SegTag::param, nullptr, 0, nullptr,
                                            &empty ) );
      std::unique_ptr< tnlmi_segments, void ( * )( tnlmi_segments* ) > own( empty, tnlmi_segments_destroy );
      const tnlmi_matrix m{ 0, 0, empty, detail::dtype< Real >(), nullptr, nullptr, matrixType() };
      int64_t n = 0;
      detail::check( tnlmi_matrix_serialization_type( &m, nullptr, 0, &n ) );
      std::string t( (size_t) n + 1, '\0' );
      detail::check( tnlmi_matrix_serialization_type( &m, t.data(), n + 1, &n ) );
      t.resize( (size_t) n );
      return t;
   }

   //! save( fileName ) (SparseMatrix.hpp:640-646): TNL's binary format
   void save( const std::string& fileName ) const
   {
      const tnlmi_matrix m = abi();
      detail::check( tnlmi_matrix_save( &m, fileName.c_str(), nullptr ) );
   }

   //! load( fileName ) (SparseMatrix.hpp:656-694): a file of this matrix type
   //! (other types: "object type does not match"), its segments arrays as saved
   void load( const std::string& fileName )
   {
      tnlmi_matrix_file* f = nullptr;
      detail::check( tnlmi_matrix_file_read( fileName.c_str(), &f ) );
      std::unique_ptr< tnlmi_matrix_file, void ( * )( tnlmi_matrix_file* ) > own( f, tnlmi_matrix_file_destroy );
      const std::string found = tnlmi_matrix_file_type( f ), expected = getSerializationType();
      if( found != expected )
         throw std::runtime_error( "Failed to deserialize an object from the file '" + fileName
                                   + "': object type does not match (expected " + expected + ", found " + found + ")." );
      int64_t rows = 0, columns = 0, storage = 0;
      detail::check( tnlmi_matrix_file_info( f, &rows, &columns, nullptr, nullptr, nullptr, nullptr, &storage,
                                                     nullptr ) );
      release();
      rows_ = (int) rows;
      columns_ = (int) columns;
      detail::check( tnlmi_matrix_file_segments( f, nullptr, &segments_ ) );
      detail::hip( hipMalloc( (void**) &values_, std::max< int64_t >( 1, storage ) * sizeof( Real ) ) );
      detail::hip( hipMalloc( (void**) &columnIndexes_, std::max< int64_t >( 1, storage ) * sizeof( int ) ) );
      detail::check( tnlmi_matrix_file_upload( f, values_, columnIndexes_, nullptr ) );
      mirror();
   }

   //! The C ABI description of this matrix.
   tnlmi_matrix abi() const
   {
      return { rows_, columns_, segments_, detail::dtype< Real >(), values_, columnIndexes_, matrixType() };
   }
   static constexpr bool isSymmetric() { return MatrixType::isSymmetric(); }
   static constexpr int32_t matrixType() { return MatrixType::isSymmetric() ? TNLMI_SYMMETRIC : TNLMI_GENERAL; }

   //! transposedVectorProduct (SparseMatrixBase.hpp:541-588): out = a*A^T*in + b*out on columns [begin, end)
   template< typename InVector, typename OutVector >
   void transposedVectorProduct( const InVector& in, OutVector& out, ComputeReal matrixMultiplicator = 1.0,
                                 ComputeReal outVectorMultiplicator = 0.0, int begin = 0, int end = 0 ) const
   {
      if( getRows() != in.getSize() )
         throw std::invalid_argument(
            "transposedVectorProduct: size of the input vector does not match the number of matrix rows" );
      if( getColumns() != out.getSize() )
         throw std::invalid_argument(
            "transposedVectorProduct: size of the output vector does not match the number of matrix columns" );
      using OutReal = typename OutVector::RealType;
      const tnlmi_matrix m = abi();
      detail::check( tnlmi_spmv_transposed( &m, detail::dtype< OutReal >(), in.getData(), out.getData(),
                                            matrixMultiplicator, outVectorMultiplicator, begin, end, nullptr ) );
      detail::hip( hipDeviceSynchronize() );
   }

private:
   void rowCounts( int kind, RowCapacitiesType& out ) const
   {
      out.setSize( rows_ );
      const tnlmi_matrix m = abi();
      detail::check( tnlmi_matrix_row_counts( &m, kind, out.getData(), nullptr ) );
   }
   // the rows' segment sizes on the host (getRowCapacity)
   void mirror()
   {
      RowCapacitiesType c;
      rowCounts( 0, c );
      const std::vector< int > sizes = c.toHost();
      hoff_.assign( rows_ + 1, 0 );
      for( int r = 0; r < rows_; r++ )
         hoff_[ r + 1 ] = hoff_[ r ] + sizes[ r ];
   }

   void release()
   {
      if( segments_ )
         tnlmi_segments_destroy( segments_ );
      for( void* p : { (void*) values_, (void*) columnIndexes_ } )
         if( p )
            (void) hipFree( p );
      segments_ = nullptr;
      values_ = nullptr;
      columnIndexes_ = nullptr;
   }
   void assemble( const std::vector< int >& caps, const std::vector< int >& off, const std::vector< int >& cols,
                  const std::vector< Real >& vals )
   {
      release();
      int *d_caps = nullptr, *d_off = nullptr, *d_cols = nullptr;
      Real* d_vals = nullptr;
      auto up = [ & ]( auto*& d, const auto& h ) {
         const size_t bytes = std::max< size_t >( 1, h.size() ) * sizeof( h[ 0 ] );
         detail::hip( hipMalloc( (void**) &d, bytes ) );
         if( ! h.empty() )
            detail::hip( hipMemcpy( d, h.data(), h.size() * sizeof( h[ 0 ] ), hipMemcpyHostToDevice ) );
      };
      up( d_caps, caps );
      up( d_off, off );
      up( d_cols, cols );
      up( d_vals, vals );
      tnlmi_status st = tnlmi_segments_create( SegTag::format, SegTag::organization, SegTag::param, d_caps, rows_,
                                               nullptr, &segments_ );
      if( st == TNLMI_OK ) {
         const int64_t storage = std::max< int64_t >( 1, tnlmi_segments_storage_size( segments_ ) );
         detail::hip( hipMalloc( (void**) &values_, storage * sizeof( Real ) ) );
         detail::hip( hipMalloc( (void**) &columnIndexes_, storage * sizeof( int ) ) );
         st = tnlmi_matrix_fill( segments_, d_off, d_cols, d_vals, detail::dtype< Real >(), values_,
                                 detail::dtype< Real >(), columnIndexes_, nullptr );
         detail::hip( hipDeviceSynchronize() );
      }
      for( void* p : { (void*) d_caps, (void*) d_off, (void*) d_cols, (void*) d_vals } )
         (void) hipFree( p );
      detail::check( st );
      mirror();
   }

   int rows_ = 0, columns_ = 0;
   tnlmi_segments* segments_ = nullptr;
   Real* values_ = nullptr;
   int* columnIndexes_ = nullptr;
   std::vector< int > hoff_;  // prefix sums of the row capacities
};

//! copySparseMatrix (Matrices/SparseOperations.hpp:716-721) between MI355X
//! matrices of any formats
template< typename Matrix1, typename Matrix2 >
void copySparseMatrix( Matrix1& A, const Matrix2& B )
{
   A.copyFrom( B );
}

template< typename Matrix, typename Device = typename Matrix::DeviceType >
class MatrixReader;

//! MatrixReader::readMtx (Matrices/MatrixReader.hpp:16-32): a MatrixMarket
//! coordinate file (real / integer / pattern, general / symmetric) into the
//! matrix -- the reference reads into a host matrix and copies it; here the
//! library parses (tnlmi_mtx_read) and the matrix is assembled on the GPU
//! from the CSR arrays, with the same dimensions, entries and exceptions
//! (std::runtime_error).
template< typename Matrix >
class MatrixReader< Matrix, Devices::MI355X >
{
public:
   static void readMtx( const std::string& fileName, Matrix& matrix, bool verbose = false )
   {
      tnlmi_host_csr* h = nullptr;
      detail::check( tnlmi_mtx_read( fileName.c_str(), &h ) );
      load( h, matrix, verbose );
   }
   static void readMtx( std::istream& str, Matrix& matrix, bool verbose = false )
   {
      const std::string text( ( std::istreambuf_iterator< char >( str ) ), std::istreambuf_iterator< char >() );
      tnlmi_host_csr* h = nullptr;
      detail::check( tnlmi_mtx_parse( text.data(), (int64_t) text.size(), &h ) );
      load( h, matrix, verbose );
   }

private:
   static void load( tnlmi_host_csr* h, Matrix& matrix, bool verbose )
   {
      std::unique_ptr< tnlmi_host_csr, void ( * )( tnlmi_host_csr* ) > own( h, tnlmi_host_csr_destroy );
      int64_t rows = 0, columns = 0, nnz = 0;
      const int32_t *off = nullptr, *cols = nullptr;
      const double* vals = nullptr;
      detail::check( tnlmi_host_csr_get( h, &rows, &columns, &nnz, &off, &cols, &vals ) );
      if( verbose )
         std::cout << "Matrix dimensions are " << rows << " x " << columns << '\n';
      matrix.setDimensions( (int) rows, (int) columns );
      using Real = typename Matrix::RealType;
      matrix.setFromCSR( std::vector< int >( off, off + rows + 1 ), std::vector< int >( cols, cols + nnz ),
                         std::vector< Real >( vals, vals + nnz ) );
   }
};

}  // namespace Matrices

// ------------------------------------------------------------------ solvers
namespace Config {
//! The subset of Config::ParameterContainer the solvers' setup() reads.
class ParameterContainer
{
public:
   using Value = std::variant< int, double, std::string >;
   template< typename T >
   void addParameter( const std::string& name, const T& value )
   {
      if constexpr( std::is_same_v< T, const char* > || std::is_convertible_v< T, std::string > )
         values_[ name ] = std::string( value );
      else if constexpr( std::is_integral_v< T > )
         values_[ name ] = (int) value;
      else
         values_[ name ] = (double) value;
   }
   bool checkParameter( const std::string& name ) const { return values_.count( name ) > 0; }
   template< typename T >
   T getParameter( const std::string& name ) const
   {
      const Value& v = values_.at( name );
      if constexpr( std::is_same_v< T, std::string > )
         return std::get< std::string >( v );
      else if( std::holds_alternative< int >( v ) )
         return (T) std::get< int >( v );
      else
         return (T) std::get< double >( v );
   }

private:
   std::map< std::string, Value > values_;
};
}  // namespace Config

namespace Solvers::Linear {

namespace Preconditioners {

template< typename Matrix >
class Preconditioner
{
public:
   using MatrixPointer = std::shared_ptr< const Matrix >;
   using VectorType = typename Matrix::VectorType;
   virtual ~Preconditioner()
   {
      if( handle_ )
         tnlmi_preconditioner_destroy( handle_ );
   }
   virtual void update( const MatrixPointer& matrix ) = 0;
   void solve( const VectorType& b, VectorType& x ) const
   {
      if( ! handle_ )
         throw std::logic_error( "Preconditioner::solve: update() was not called" );
      detail::check( tnlmi_preconditioner_solve( handle_, b.getData(), x.getData(), nullptr ) );
      detail::hip( hipDeviceSynchronize() );
   }
   const tnlmi_preconditioner* handle() const { return handle_; }

protected:
   void build( tnlmi_precond_kind kind, const MatrixPointer& matrix )
   {
      if( handle_ )
         tnlmi_preconditioner_destroy( handle_ );
      handle_ = nullptr;
      const tnlmi_matrix m = matrix->abi();
      detail::check( tnlmi_preconditioner_create( kind, &m, nullptr, &handle_ ) );
   }
   tnlmi_preconditioner* handle_ = nullptr;
};

//! Preconditioners::Diagonal (Diagonal.hpp:12-43)
template< typename Matrix >
class Diagonal : public Preconditioner< Matrix >
{
public:
   void update( const typename Preconditioner< Matrix >::MatrixPointer& m ) override
   {
      this->build( TNLMI_PRECOND_DIAGONAL, m );
   }
};

//! Preconditioners::ILU0 (ILU0.hpp:12-139); the reference's GPU version throws
//! NotImplementedError, this device factorises on the host and solves on the GPU
template< typename Matrix >
class ILU0 : public Preconditioner< Matrix >
{
public:
   void update( const typename Preconditioner< Matrix >::MatrixPointer& m ) override
   {
      this->build( TNLMI_PRECOND_ILU0, m );
   }
};

}  // namespace Preconditioners

//! LinearSolver + IterativeSolver (LinearSolver.h:37-167, IterativeSolver.h:226-235)
template< typename Matrix >
class LinearSolver
{
public:
   using MatrixPointer = std::shared_ptr< const Matrix >;
   using PreconditionerPointer = std::shared_ptr< const Preconditioners::Preconditioner< Matrix > >;
   using VectorType = typename Matrix::VectorType;
   using RealType = typename Matrix::RealType;
   virtual ~LinearSolver() = default;

   void setMatrix( const MatrixPointer& m ) { matrix_ = m; }
   void setPreconditioner( const PreconditionerPointer& p ) { preconditioner_ = p; }
   virtual bool setup( const Config::ParameterContainer& p, const std::string& prefix = "" )
   {
      if( p.checkParameter( prefix + "max-iterations" ) )
         max_iterations_ = p.getParameter< long >( prefix + "max-iterations" );
      if( p.checkParameter( prefix + "min-iterations" ) )
         min_iterations_ = p.getParameter< long >( prefix + "min-iterations" );
      if( p.checkParameter( prefix + "convergence-residue" ) )
         convergence_residue_ = p.getParameter< double >( prefix + "convergence-residue" );
      if( p.checkParameter( prefix + "divergence-residue" ) )
         divergence_residue_ = p.getParameter< double >( prefix + "divergence-residue" );
      return true;
   }
   void setMaxIterations( long n ) { max_iterations_ = n; }
   void setMinIterations( long n ) { min_iterations_ = n; }
   void setConvergenceResidue( double r ) { convergence_residue_ = r; }
   void setDivergenceResidue( double r ) { divergence_residue_ = r; }
   long getIterations() const { return iterations_; }
   double getResidue() const { return residue_; }
   virtual bool solve( const VectorType& b, VectorType& x ) = 0;

protected:
   tnlmi_solver_state state() const
   {
      return { max_iterations_, min_iterations_, convergence_residue_, divergence_residue_, 0, 0.0, 0 };
   }
   bool finish( const tnlmi_solver_state& st )
   {
      iterations_ = st.iterations;
      residue_ = st.residue;
      return st.converged != 0;
   }
   const tnlmi_preconditioner* pre() const { return preconditioner_ ? preconditioner_->handle() : nullptr; }
   tnlmi_matrix mat( const char* who ) const
   {
      if( ! matrix_ )
         throw std::logic_error( std::string( "No matrix was set in " ) + who + ". Call setMatrix() before solve()." );
      return matrix_->abi();
   }
   MatrixPointer matrix_;
   PreconditionerPointer preconditioner_;
   long max_iterations_ = 1000000000, min_iterations_ = 0;
   double convergence_residue_ = 1e-6, divergence_residue_ = std::numeric_limits< double >::max();
   long iterations_ = 0;
   double residue_ = 0;
};

//! CG (CG.hpp:10-108)
template< typename Matrix >
class CG : public LinearSolver< Matrix >
{
public:
   bool solve( const typename LinearSolver< Matrix >::VectorType& b,
               typename LinearSolver< Matrix >::VectorType& x ) override
   {
      const tnlmi_matrix m = this->mat( "CG" );
      tnlmi_solver_state st = this->state();
      detail::check( tnlmi_cg_solve( &m, nullptr, this->pre(), b.getData(), x.getData(), &st, nullptr ) );
      return this->finish( st );
   }
};

//! GMRES (GMRES.hpp:18-180): variant and restarting parameters via setup()
template< typename Matrix >
class GMRES : public LinearSolver< Matrix >
{
public:
   bool setup( const Config::ParameterContainer& p, const std::string& prefix = "" ) override
   {
      if( p.checkParameter( prefix + "gmres-variant" ) ) {
         const std::string v = p.getParameter< std::string >( prefix + "gmres-variant" );
         const std::map< std::string, tnlmi_gmres_variant > names{ { "CGS", TNLMI_GMRES_CGS },
                                                                   { "CGSR", TNLMI_GMRES_CGSR },
                                                                   { "MGS", TNLMI_GMRES_MGS },
                                                                   { "MGSR", TNLMI_GMRES_MGSR },
                                                                   { "CWY", TNLMI_GMRES_CWY } };
         if( ! names.count( v ) )
            return false;
         variant_ = names.at( v );
      }
      auto get = [ & ]( const char* k, int& dst ) {
         if( p.checkParameter( prefix + k ) )
            dst = p.getParameter< int >( prefix + k );
      };
      get( "gmres-restarting-min", rmin_ );
      get( "gmres-restarting-max", rmax_ );
      get( "gmres-restarting-step-min", dmin_ );
      get( "gmres-restarting-step-max", dmax_ );
      return LinearSolver< Matrix >::setup( p, prefix );
   }
   bool solve( const typename LinearSolver< Matrix >::VectorType& b,
               typename LinearSolver< Matrix >::VectorType& x ) override
   {
      const tnlmi_matrix m = this->mat( "GMRES" );
      tnlmi_solver_state st = this->state();
      detail::check( tnlmi_gmres_solve( &m, nullptr, this->pre(), variant_, rmin_, rmax_, dmin_, dmax_, b.getData(),
                                        x.getData(), &st, nullptr ) );
      return this->finish( st );
   }

private:
   tnlmi_gmres_variant variant_ = TNLMI_GMRES_CWY;  // class default, GMRES.h:212
   int rmin_ = 10, rmax_ = 10, dmin_ = 3, dmax_ = 3;
};

//! BICGStab (BICGStab.hpp:12-127): bicgstab-exact-residue via setup()
template< typename Matrix >
class BICGStab : public LinearSolver< Matrix >
{
public:
   bool setup( const Config::ParameterContainer& p, const std::string& prefix = "" ) override
   {
      if( p.checkParameter( prefix + "bicgstab-exact-residue" ) )
         exact_residue_ = p.getParameter< bool >( prefix + "bicgstab-exact-residue" );
      return LinearSolver< Matrix >::setup( p, prefix );
   }
   bool solve( const typename LinearSolver< Matrix >::VectorType& b,
               typename LinearSolver< Matrix >::VectorType& x ) override
   {
      const tnlmi_matrix m = this->mat( "BICGStab" );
      tnlmi_solver_state st = this->state();
      detail::check( tnlmi_bicgstab_solve( &m, nullptr, this->pre(), exact_residue_ ? 1 : 0, b.getData(), x.getData(),
                                           &st, nullptr ) );
      return this->finish( st );
   }

private:
   bool exact_residue_ = false;  // configSetup default, BICGStab.hpp:17-20
};

//! TFQMR (TFQMR.hpp:12-121)
template< typename Matrix >
class TFQMR : public LinearSolver< Matrix >
{
public:
   bool solve( const typename LinearSolver< Matrix >::VectorType& b,
               typename LinearSolver< Matrix >::VectorType& x ) override
   {
      const tnlmi_matrix m = this->mat( "TFQMR" );
      tnlmi_solver_state st = this->state();
      detail::check( tnlmi_tfqmr_solve( &m, nullptr, this->pre(), b.getData(), x.getData(), &st, nullptr ) );
      return this->finish( st );
   }
};

//! BICGStabL (BICGStabL.hpp:12-257): bicgstab-ell and bicgstab-exact-residue via setup()
template< typename Matrix >
class BICGStabL : public LinearSolver< Matrix >
{
public:
   bool setup( const Config::ParameterContainer& p, const std::string& prefix = "" ) override
   {
      if( p.checkParameter( prefix + "bicgstab-ell" ) )
         ell_ = p.getParameter< int >( prefix + "bicgstab-ell" );
      if( p.checkParameter( prefix + "bicgstab-exact-residue" ) )
         exact_residue_ = p.getParameter< bool >( prefix + "bicgstab-exact-residue" );
      return LinearSolver< Matrix >::setup( p, prefix );
   }
   bool solve( const typename LinearSolver< Matrix >::VectorType& b,
               typename LinearSolver< Matrix >::VectorType& x ) override
   {
      const tnlmi_matrix m = this->mat( "BICGStabL" );
      tnlmi_solver_state st = this->state();
      detail::check( tnlmi_bicgstabl_solve( &m, nullptr, this->pre(), ell_, exact_residue_ ? 1 : 0, b.getData(),
                                            x.getData(), &st, nullptr ) );
      return this->finish( st );
   }

private:
   int ell_ = 1;                 // configSetup default, BICGStabL.hpp:17
   bool exact_residue_ = false;  // BICGStabL.hpp:18-21
};

//! IDRs (IDRs.hpp:12-386): idr-s and idr-residual-smoothing via setup() or the setters
template< typename Matrix >
class IDRs : public LinearSolver< Matrix >
{
public:
   bool setup( const Config::ParameterContainer& p, const std::string& prefix = "" ) override
   {
      if( p.checkParameter( prefix + "idr-s" ) )
         s_ = p.getParameter< int >( prefix + "idr-s" );
      if( p.checkParameter( prefix + "idr-residual-smoothing" ) )
         smoothing_ = p.getParameter< bool >( prefix + "idr-residual-smoothing" );
      return LinearSolver< Matrix >::setup( p, prefix );
   }
   void setShadowSpaceDimension( int s ) { s_ = s; }
   void setResidualSmoothing( bool smoothing ) { smoothing_ = smoothing; }
   bool solve( const typename LinearSolver< Matrix >::VectorType& b,
               typename LinearSolver< Matrix >::VectorType& x ) override
   {
      const tnlmi_matrix m = this->mat( "IDRs" );
      tnlmi_solver_state st = this->state();
      detail::check(
         tnlmi_idrs_solve( &m, nullptr, this->pre(), s_, smoothing_ ? 1 : 0, b.getData(), x.getData(), &st, nullptr ) );
      return this->finish( st );
   }

private:
   int s_ = 4;               // IDRs.h default shadow space dimension
   bool smoothing_ = false;  // IDRs.h default
};

//! Jacobi (Jacobi.hpp:65-125)
template< typename Matrix >
class Jacobi : public LinearSolver< Matrix >
{
public:
   void setOmega( double w ) { omega_ = w; }
   void setResiduePeriod( long p ) { period_ = p; }
   bool solve( const typename LinearSolver< Matrix >::VectorType& b,
               typename LinearSolver< Matrix >::VectorType& x ) override
   {
      const tnlmi_matrix m = this->mat( "Jacobi" );
      tnlmi_solver_state st = this->state();
      detail::check( tnlmi_jacobi_solve( &m, nullptr, omega_, period_, b.getData(), x.getData(), &st, nullptr ) );
      return this->finish( st );
   }

private:
   double omega_ = 1.0;
   long period_ = 4;
};

//! LinearResidueGetter (Utils/LinearResidueGetter.hpp:11-23)
struct LinearResidueGetter
{
   template< typename Matrix, typename Vector >
   static double getResidue( const Matrix& A, const Vector& x, const Vector& b, double = 0.0 )
   {
      const tnlmi_matrix m = A.abi();
      double r = 0;
      detail::check( tnlmi_true_residual( &m, x.getData(), b.getData(), &r, nullptr ) );
      return r;
   }
};

}  // namespace Solvers::Linear
}  // namespace TNL
