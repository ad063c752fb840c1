// TNL/MI355X.h -- the MI355X device for TNL's own class templates, over the C
// ABI of libtnlmi.so (include/tnlmi.h).
//
// A caller of the reference writes
//     TNL::Matrices::SparseMatrix< double, TNL::Devices::Host, int > A( ... );
//     TNL::Solvers::Linear::CG< decltype( A ) > cg;  cg.setMatrix( ptr ); cg.solve( b, x );
// and moves to the GPU by replacing Devices::Host with Devices::MI355X.  This
// header includes the reference's headers and adds the device the way the
// reference adds its own: a new Devices tag, Allocators::Default for it, and
// partial specialisations of the reference's primary templates for that tag
// --
//   Containers::Vector< Real, MI355X, Index, Allocator >, VectorView,
//   Matrices::SparseMatrix< Real, MI355X, Index, MatrixType, Segments,
//                           ComputeReal, RealAllocator, IndexAllocator >,
//   Matrices::MatrixReader< Matrix, MI355X >,
//   Solvers::Linear::{CG, GMRES, BICGStab, BICGStabL, IDRs, TFQMR, Jacobi}
//      for those matrices (deriving from the reference's LinearSolver, so
//      setup / setMatrix / setPreconditioner / setConvergenceResidue /
//      getIterations / getResidue are the reference's own),
//   Preconditioners::Diagonal, and ILU0 through the reference's ILU0_impl
//      device hook,
// with the reference's names, argument meaning and exceptions.  Any segments
// template the reference defines (CSR, AdaptiveCSR, Ellpack, SlicedEllpack,
// BiEllpack, ChunkedEllpack, SortedSegments over CSR) selects its format.
// The reference's LinearResidueGetter and copySparseMatrix work unchanged on
// these types (the latter through an overload for two MI355X matrices).
//
// Every call is synchronous on return, as the reference's reduceSegments is
// (Algorithms/Segments/detail/ReducingOperations_CSR.h:327).
#pragma once

#include <hip/hip_runtime_api.h>

#include <climits>
#include <cmath>
#include <cstdint>
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
#include <vector>

#include <TNL/Algorithms/Segments/AdaptiveCSR.h>
#include <TNL/Algorithms/Segments/BiEllpack.h>
#include <TNL/Algorithms/Segments/CSR.h>
#include <TNL/Algorithms/Segments/ChunkedEllpack.h>
#include <TNL/Algorithms/Segments/Ellpack.h>
#include <TNL/Algorithms/Segments/LaunchConfiguration.h>
#include <TNL/Algorithms/Segments/SlicedEllpack.h>
#include <TNL/Algorithms/Segments/SortedSegments.h>
#include <TNL/Allocators/Default.h>
#include <TNL/Config/ParameterContainer.h>
#include <TNL/Containers/Vector.h>
#include <TNL/Containers/VectorView.h>
#include <TNL/Exceptions/BackendBadAlloc.h>
#include <TNL/Exceptions/BackendRuntimeError.h>
#include <TNL/Exceptions/NotImplementedError.h>
#include <TNL/Matrices/MatrixReader.h>
#include <TNL/Matrices/MatrixType.h>
#include <TNL/Matrices/SparseMatrix.h>
#include <TNL/Matrices/SparseOperations.h>
#include <TNL/Solvers/Linear/BICGStab.h>
#include <TNL/Solvers/Linear/BICGStabL.h>
#include <TNL/Solvers/Linear/CG.h>
#include <TNL/Solvers/Linear/GMRES.h>
#include <TNL/Solvers/Linear/IDRs.h>
#include <TNL/Solvers/Linear/Jacobi.h>
#include <TNL/Solvers/Linear/LinearSolver.h>
#include <TNL/Solvers/Linear/Preconditioners/Diagonal.h>
#include <TNL/Solvers/Linear/Preconditioners/ILU0.h>
#include <TNL/Solvers/Linear/TFQMR.h>
#include <TNL/Solvers/Linear/Utils/LinearResidueGetter.h>

#include <TNL/Algorithms/parallelFor.h>

#include "tnlmi.h"
#if defined( __HIPCC__ ) || defined( __HIP__ )
   #include "tnlmi_segments_view.h"
#endif

namespace TNL {

// ------------------------------------------------------------------ device
namespace Devices {
//! The MI355X GPU (gfx950, wave64), driven through libtnlmi.so.
struct MI355X
{
   static std::string getDeviceType() { return "Devices::MI355X"; }
   //! Algorithms::parallelFor< MI355X > takes it (ParallelFor1D.h); nothing to configure
   struct LaunchConfiguration
   {};
};
}  // namespace Devices

namespace MI355XDetail {
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
         throw Exceptions::BackendRuntimeError( Backend::error_t( 1 ), m );
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
      throw Exceptions::BackendRuntimeError( Backend::error_t( e ), hipGetErrorString( e ) );
}
template< typename T >
constexpr tnlmi_dtype dtype()
{
   static_assert( std::is_same_v< std::remove_const_t< T >, double > || std::is_same_v< std::remove_const_t< T >, float >,
                  "Devices::MI355X vectors and matrices hold double or float values" );
   return std::is_same_v< std::remove_const_t< T >, double > ? TNLMI_F64 : TNLMI_F32;
}
}  // namespace MI355XDetail

// --------------------------------------------------------------- allocator
namespace Allocators {
//! Device memory of the MI355X (hipMalloc / hipFree), the std::allocator
//! interface of Allocators::Hip.
template< class T >
struct MI355X
{
   using value_type = T;
   using size_type = std::size_t;
   using difference_type = std::ptrdiff_t;
   MI355X() = default;
   template< class U >
   MI355X( const MI355X< U >& ) noexcept
   {}
   [[nodiscard]] T* allocate( size_type n )
   {
      T* p = nullptr;
      MI355XDetail::hip( hipMalloc( (void**) &p, std::max< size_type >( 1, n ) * sizeof( T ) ) );
      return p;
   }
   void deallocate( T* p, size_type ) noexcept { (void) hipFree( (void*) p ); }
};
template< class T1, class T2 >
bool operator==( const MI355X< T1 >&, const MI355X< T2 >& )
{
   return true;
}
template< class T1, class T2 >
bool operator!=( const MI355X< T1 >&, const MI355X< T2 >& )
{
   return false;
}

//! Allocators/Default.h: the default allocator of Devices::MI355X
template<>
struct Default< Devices::MI355X >
{
   template< typename T >
   using Allocator = Allocators::MI355X< T >;
};
}  // namespace Allocators

// ------------------------------------------------------------------ vectors
namespace Containers {

template< typename Real, typename Index, typename Allocator >
class Vector< Real, Devices::MI355X, Index, Allocator >;

//! Containers::VectorView on the MI355X: a non-owning (data, size) pair.
//! Copy construction binds; assignment copies elements (ArrayView semantics).
template< typename Real, typename Index >
class VectorView< Real, Devices::MI355X, Index >
{
   using NonConstReal = std::remove_const_t< Real >;

public:
   using RealType = Real;
   using ValueType = Real;
   using DeviceType = Devices::MI355X;
   using IndexType = Index;
   using ViewType = VectorView< Real, Devices::MI355X, Index >;
   using ConstViewType = VectorView< std::add_const_t< Real >, Devices::MI355X, Index >;

   VectorView() = default;
   VectorView( Real* data, Index size ) : data_( data ), size_( size ) {}
   VectorView( const VectorView& ) = default;
   //! a view of const values from a view of non-const ones
   template< typename R, std::enable_if_t< std::is_same_v< std::add_const_t< R >, Real > && ! std::is_same_v< R, Real >,
                                           bool > = true >
   VectorView( const VectorView< R, Devices::MI355X, Index >& v ) : data_( v.getData() ), size_( v.getSize() )
   {}
   //! the view of a vector (Array::getView / getConstView)
   template< typename Allocator >
   VectorView( Vector< NonConstReal, Devices::MI355X, Index, Allocator >& v ) : data_( v.getData() ), size_( v.getSize() )
   {}
   template< typename Allocator, typename R = Real, std::enable_if_t< std::is_const_v< R >, bool > = true >
   VectorView( const Vector< NonConstReal, Devices::MI355X, Index, Allocator >& v )
   : data_( v.getData() ), size_( v.getSize() )
   {}

   void bind( Real* data, Index size )
   {
      data_ = data;
      size_ = size;
   }
   //! element-wise copy from a view of the same size (ArrayView::operator=)
   VectorView& operator=( const VectorView& v )
   {
      copyFrom( v.getData(), v.getSize() );
      return *this;
   }
   template< typename R >
   VectorView& operator=( const VectorView< R, Devices::MI355X, Index >& v )
   {
      copyFrom( v.getData(), v.getSize() );
      return *this;
   }
   VectorView& operator=( NonConstReal value )
   {
      setValue( value );
      return *this;
   }

   Index getSize() const { return size_; }
   Real* getData() const { return data_; }
   ViewType getView() const { return *this; }
   ConstViewType getConstView() const { return ConstViewType( data_, size_ ); }

   void setValue( NonConstReal value )
   {
      std::vector< NonConstReal > h( size_, value );
      if( size_ )
         MI355XDetail::hip( hipMemcpy( (void*) data_, h.data(), sizeof( Real ) * size_, hipMemcpyHostToDevice ) );
   }
   NonConstReal getElement( Index i ) const
   {
      NonConstReal v;
      MI355XDetail::hip( hipMemcpy( &v, data_ + i, sizeof( Real ), hipMemcpyDeviceToHost ) );
      return v;
   }
   void setElement( Index i, NonConstReal v ) const
   {
      MI355XDetail::hip( hipMemcpy( (void*) ( data_ + i ), &v, sizeof( Real ), hipMemcpyHostToDevice ) );
   }
   //! the elements on the host (an extension of this device)
   std::vector< NonConstReal > toHost() const
   {
      std::vector< NonConstReal > h( size_ );
      if( size_ )
         MI355XDetail::hip( hipMemcpy( h.data(), data_, sizeof( Real ) * size_, hipMemcpyDeviceToHost ) );
      return h;
   }
   template< typename R >
   bool operator==( const VectorView< R, Devices::MI355X, Index >& o ) const
   {
      return toHost() == o.toHost();
   }

private:
   void copyFrom( const void* src, Index size )
   {
      if( size != size_ )
         throw std::invalid_argument( "VectorView::operator=: the sizes of the views differ" );
      if( size_ && src != data_ )
         MI355XDetail::hip( hipMemcpy( (void*) data_, src, sizeof( Real ) * size_, hipMemcpyDeviceToDevice ) );
   }

   Real* data_ = nullptr;
   Index size_ = 0;
};

//! Containers::Vector on the MI355X: an owning, contiguous device array
//! allocated through Allocator (Allocators::MI355X by default).
template< typename Real, typename Index, typename Allocator >
class Vector< Real, Devices::MI355X, Index, Allocator >
{
public:
   using RealType = Real;
   using ValueType = Real;
   using DeviceType = Devices::MI355X;
   using IndexType = Index;
   using AllocatorType = Allocator;
   using ViewType = VectorView< Real, Devices::MI355X, Index >;
   using ConstViewType = VectorView< std::add_const_t< Real >, Devices::MI355X, Index >;
   template< typename R, typename D = Devices::MI355X, typename I = Index,
             typename A = typename Allocators::Default< D >::template Allocator< R > >
   using Self = Vector< R, D, I, A >;

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
            MI355XDetail::hip( hipMemcpy( data_, o.data_, sizeof( Real ) * size_, hipMemcpyDeviceToDevice ) );
      }
      return *this;
   }
   Vector& operator=( Vector&& o ) noexcept
   {
      swap( o );
      return *this;
   }
   //! a copy of a view's elements
   template< typename R >
   Vector& operator=( const VectorView< R, Devices::MI355X, Index >& v )
   {
      setSize( v.getSize() );
      if( size_ && (const void*) v.getData() != (const void*) data_ )
         MI355XDetail::hip( hipMemcpy( data_, v.getData(), sizeof( Real ) * size_, hipMemcpyDeviceToDevice ) );
      return *this;
   }
   //! cross-device assignment from the reference's host vector
   template< typename HostAllocator >
   Vector& operator=( const Vector< Real, Devices::Host, Index, HostAllocator >& h )
   {
      setSize( h.getSize() );
      if( size_ )
         MI355XDetail::hip( hipMemcpy( data_, h.getData(), sizeof( Real ) * size_, hipMemcpyHostToDevice ) );
      return *this;
   }
   //! from host elements (an extension of this device)
   Vector& operator=( const std::vector< Real >& h )
   {
      setSize( (Index) h.size() );
      if( size_ )
         MI355XDetail::hip( hipMemcpy( data_, h.data(), sizeof( Real ) * size_, hipMemcpyHostToDevice ) );
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
         throw std::invalid_argument( "Array size must be non-negative." );
      if( size == size_ )
         return;
      reset();
      if( size > 0 )
         data_ = Allocator().allocate( (std::size_t) size );
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
         Allocator().deallocate( data_, (std::size_t) size_ );
      data_ = nullptr;
      size_ = 0;
   }
   Index getSize() const { return size_; }
   Real* getData() { return data_; }
   const Real* getData() const { return data_; }
   ViewType getView() { return ViewType( data_, size_ ); }
   ConstViewType getView() const { return ConstViewType( data_, size_ ); }
   ConstViewType getConstView() const { return ConstViewType( data_, size_ ); }
   operator ViewType() { return getView(); }
   operator ConstViewType() const { return getConstView(); }

   void setValue( Real value ) { getView().setValue( value ); }
   Real getElement( Index i ) const { return getConstView().getElement( i ); }
   void setElement( Index i, Real v ) { getView().setElement( i, v ); }
   //! the elements on the host (an extension of this device)
   std::vector< Real > toHost() const { return getConstView().toHost(); }
   template< typename A >
   bool operator==( const Vector< Real, Devices::MI355X, Index, A >& o ) const
   {
      return toHost() == o.toHost();
   }
   void swap( Vector& o ) noexcept
   {
      std::swap( data_, o.data_ );
      std::swap( size_, o.size_ );
   }

private:
   Real* data_ = nullptr;
   Index size_ = 0;
};

// The reference's expression templates evaluate through Algorithms on the
// vector's device; this device provides its operations below instead.
namespace Expressions {
template< typename Real, typename Index, typename Allocator >
struct HasEnabledExpressionTemplates< Vector< Real, Devices::MI355X, Index, Allocator > > : std::false_type
{};
template< typename Real, typename Index >
struct HasEnabledExpressionTemplates< VectorView< Real, Devices::MI355X, Index > > : std::false_type
{};
}  // namespace Expressions

template< typename Real, typename Index >
std::ostream& operator<<( std::ostream& s, const VectorView< Real, Devices::MI355X, Index >& v )
{
   const auto h = v.toHost();
   s << "[ ";
   for( std::size_t i = 0; i < h.size(); i++ )
      s << ( i ? ", " : "" ) << h[ i ];
   return s << " ]";
}
template< typename Real, typename Index, typename Allocator >
std::ostream& operator<<( std::ostream& s, const Vector< Real, Devices::MI355X, Index, Allocator >& v )
{
   return s << v.getConstView();
}

namespace detail_mi355x {
// a library reduction into one device double, read back
template< typename F >
double device_scalar( F&& f )
{
   Vector< double, Devices::MI355X, int > d( 1 );
   MI355XDetail::check( f( d.getData() ) );
   return d.getElement( 0 );
}
template< typename Index >
double dot( VectorView< const double, Devices::MI355X, Index > a, VectorView< const double, Devices::MI355X, Index > b )
{
   if( a.getSize() != b.getSize() )
      throw std::invalid_argument( "the vectors of a scalar product differ in size" );
   return device_scalar( [ & ]( double* d ) { return tnlmi_dot( a.getSize(), a.getData(), b.getData(), d, nullptr ); } );
}
template< typename Index >
double nrm2( VectorView< const double, Devices::MI355X, Index > a )
{
   return device_scalar( [ & ]( double* d ) { return tnlmi_nrm2( a.getSize(), a.getData(), d, nullptr ); } );
}
template< typename Index >
Vector< double, Devices::MI355X, Index > minus( VectorView< const double, Devices::MI355X, Index > a,
                                                VectorView< const double, Devices::MI355X, Index > b )
{
   if( a.getSize() != b.getSize() )
      throw std::invalid_argument( "the vectors of a difference differ in size" );
   Vector< double, Devices::MI355X, Index > out;
   out = a;
   MI355XDetail::check( tnlmi_axpby( out.getSize(), -1.0, b.getData(), 1.0, out.getData(), nullptr ) );
   MI355XDetail::hip( hipDeviceSynchronize() );
   return out;
}
}  // namespace detail_mi355x

//! (a, b): the scalar product of Containers/Expressions/ExpressionTemplates.h:372-384
template< typename I, typename A1, typename A2 >
double operator,( const Vector< double, Devices::MI355X, I, A1 >& a, const Vector< double, Devices::MI355X, I, A2 >& b )
{
   return detail_mi355x::dot< I >( a.getConstView(), b.getConstView() );
}
template< typename R1, typename R2, typename I >
double operator,( const VectorView< R1, Devices::MI355X, I >& a, const VectorView< R2, Devices::MI355X, I >& b )
{
   return detail_mi355x::dot< I >( a.getConstView(), b.getConstView() );
}
//! l2Norm / lpNorm( ., 2 ) (ExpressionTemplates.h:442-460)
template< typename I, typename A >
double l2Norm( const Vector< double, Devices::MI355X, I, A >& a )
{
   return detail_mi355x::nrm2< I >( a.getConstView() );
}
template< typename R, typename I >
double l2Norm( const VectorView< R, Devices::MI355X, I >& a )
{
   return detail_mi355x::nrm2< I >( a.getConstView() );
}
template< typename I, typename A >
double lpNorm( const Vector< double, Devices::MI355X, I, A >& a, double p )
{
   if( p != 2.0 )
      throw Exceptions::NotImplementedError( "lpNorm on Devices::MI355X is implemented for p = 2" );
   return l2Norm( a );
}
template< typename R, typename I >
double lpNorm( const VectorView< R, Devices::MI355X, I >& a, double p )
{
   if( p != 2.0 )
      throw Exceptions::NotImplementedError( "lpNorm on Devices::MI355X is implemented for p = 2" );
   return l2Norm( a );
}
//! a - b into a new vector (the expression LinearResidueGetter.hpp:21 evaluates)
template< typename I, typename A1, typename A2 >
Vector< double, Devices::MI355X, I > operator-( const Vector< double, Devices::MI355X, I, A1 >& a,
                                                 const Vector< double, Devices::MI355X, I, A2 >& b )
{
   return detail_mi355x::minus< I >( a.getConstView(), b.getConstView() );
}
template< typename R, typename I, typename A >
Vector< double, Devices::MI355X, I > operator-( const Vector< double, Devices::MI355X, I, A >& a,
                                                 const VectorView< R, Devices::MI355X, I >& b )
{
   return detail_mi355x::minus< I >( a.getConstView(), b.getConstView() );
}
template< typename R, typename I, typename A >
Vector< double, Devices::MI355X, I > operator-( const VectorView< R, Devices::MI355X, I >& a,
                                                 const Vector< double, Devices::MI355X, I, A >& b )
{
   return detail_mi355x::minus< I >( a.getConstView(), b.getConstView() );
}
template< typename R1, typename R2, typename I >
Vector< double, Devices::MI355X, I > operator-( const VectorView< R1, Devices::MI355X, I >& a,
                                                 const VectorView< R2, Devices::MI355X, I >& b )
{
   return detail_mi355x::minus< I >( a.getConstView(), b.getConstView() );
}

}  // namespace Containers

// ----------------------------------------------------------------- segments
namespace MI355XDetail {

//! The C ABI format of each segments template of the reference (instantiated
//! on the host device only to name its parameters).
template< typename S >
struct SegmentsFormat;
template< typename D, typename I, typename A >
struct SegmentsFormat< Algorithms::Segments::CSR< D, I, A > >
{
   static constexpr tnlmi_format format = TNLMI_CSR;
   static constexpr int param = 0;
   static constexpr tnlmi_organization organization = TNLMI_COLUMN_MAJOR;
};
template< typename D, typename I, typename A >
struct SegmentsFormat< Algorithms::Segments::AdaptiveCSR< D, I, A > >
{
   static constexpr tnlmi_format format = TNLMI_ADAPTIVE_CSR;
   static constexpr int param = 0;
   static constexpr tnlmi_organization organization = TNLMI_COLUMN_MAJOR;
};
template< typename D, typename I, typename A, Algorithms::Segments::ElementsOrganization O, int Alignment >
struct SegmentsFormat< Algorithms::Segments::Ellpack< D, I, A, O, Alignment > >
{
   static_assert( Alignment == 32, "Ellpack on Devices::MI355X keeps the reference's default alignment 32" );
   static constexpr tnlmi_format format = TNLMI_ELLPACK;
   static constexpr int param = 0;
   static constexpr tnlmi_organization organization = (tnlmi_organization) O;
};
template< typename D, typename I, typename A, Algorithms::Segments::ElementsOrganization O, int SliceSize >
struct SegmentsFormat< Algorithms::Segments::SlicedEllpack< D, I, A, O, SliceSize > >
{
   static constexpr tnlmi_format format = TNLMI_SLICED_ELLPACK;
   static constexpr int param = SliceSize;
   static constexpr tnlmi_organization organization = (tnlmi_organization) O;
};
template< typename D, typename I, typename A, Algorithms::Segments::ElementsOrganization O, int WarpSize >
struct SegmentsFormat< Algorithms::Segments::BiEllpack< D, I, A, O, WarpSize > >
{
   static constexpr tnlmi_format format = TNLMI_BIELLPACK;
   static constexpr int param = WarpSize;
   static constexpr tnlmi_organization organization = (tnlmi_organization) O;
};
template< typename D, typename I, typename A, Algorithms::Segments::ElementsOrganization O >
struct SegmentsFormat< Algorithms::Segments::ChunkedEllpack< D, I, A, O > >
{
   static constexpr tnlmi_format format = TNLMI_CHUNKED_ELLPACK;
   static constexpr int param = 0;
   static constexpr tnlmi_organization organization = (tnlmi_organization) O;
};
template< typename D, typename I, typename A, typename A2 >
struct SegmentsFormat< Algorithms::Segments::SortedSegments< Algorithms::Segments::CSR< D, I, A >, A2 > >
{
   static constexpr tnlmi_format format = TNLMI_SORTED_CSR;
   static constexpr int param = 0;
   static constexpr tnlmi_organization organization = TNLMI_COLUMN_MAJOR;
};

//! LaunchConfiguration.h:18-91 -> the C ABI.  The reference's
//! default-constructed configuration (Fixed, 1) asks for no particular kernel
//! and selects this device's tuned product (TNLMI_MAP_AUTO); every other
//! configuration runs the reference's kernel of that mapping.
inline tnlmi_launch launch( const Algorithms::Segments::LaunchConfiguration& c )
{
   using M = Algorithms::Segments::ThreadsToSegmentsMapping;
   if( c.getThreadsToSegmentsMapping() == M::Fixed && c.getThreadsPerSegmentCount() == 1 )
      return { TNLMI_MAP_AUTO, 1 };
   return { (int32_t) c.getThreadsToSegmentsMapping(), c.getThreadsPerSegmentCount() };
}
}  // namespace MI355XDetail

#if defined( __HIPCC__ ) || defined( __HIP__ )
// ------------------------------------------- device-side rows (hipcc only)
// Code compiled by a HIP compiler walks a matrix's rows on the GPU the way the
// reference's does: SparseMatrix::getView().getRow( i ) inside
// Algorithms::parallelFor< Devices::MI355X >, forRows / forAllRows with a
// RowView, and forElements / forAllElements with (row, local index, column,
// value) (Matrices/SparseMatrixBase.h:239-257, 552-606, 858-925;
// SparseMatrixRowView.h).  The slots are those of the matrix's segments
// (tnlmi_segments_view.h: getGlobalIndex and getSegmentSize of every format),
// column -1 marking a free slot, as in the reference.
namespace MI355XDetail {

//! A row of a matrix on the device (SparseMatrixRowView.h).  The column
//! indexes are 32-bit on the device, so getColumnIndex returns a value rather
//! than a reference to the stored index.
template< typename Real, typename Index, bool Const >
class DeviceRowView
{
public:
   using RealType = Real;
   using IndexType = Index;
   using ValuePointer = std::conditional_t< Const, const Real*, Real* >;
   using ColumnPointer = std::conditional_t< Const, const int32_t*, int32_t* >;

   __host__ __device__ DeviceRowView( const tnlmi_segments_view& v, Index row, ValuePointer values,
                                      ColumnPointer columns )
   : v_( v ), row_( row ), values_( values ), columns_( columns )
   {}
   //! getSegmentSize of the row: its slots (the capacity the format gives it)
   __host__ __device__ Index getSize() const { return (Index) tnlmi_segment_size( &v_, row_ ); }
   __host__ __device__ Index getRowIndex() const { return row_; }
   __host__ __device__ Index getColumnIndex( Index localIdx ) const { return columns_[ slot( localIdx ) ]; }
   __host__ __device__ auto& getValue( Index localIdx ) const { return values_[ slot( localIdx ) ]; }
   template< bool C = Const, std::enable_if_t< ! C, bool > = true >
   __host__ __device__ void setValue( Index localIdx, const Real& value ) const
   {
      values_[ slot( localIdx ) ] = value;
   }
   template< bool C = Const, std::enable_if_t< ! C, bool > = true >
   __host__ __device__ void setColumnIndex( Index localIdx, Index column ) const
   {
      columns_[ slot( localIdx ) ] = (int32_t) column;
   }
   template< bool C = Const, std::enable_if_t< ! C, bool > = true >
   __host__ __device__ void setElement( Index localIdx, Index column, const Real& value ) const
   {
      const int64_t g = slot( localIdx );
      columns_[ g ] = (int32_t) column;
      values_[ g ] = value;
   }
   static constexpr bool isBinary() { return false; }

private:
   __host__ __device__ int64_t slot( Index localIdx ) const
   {
      return tnlmi_global_index( &v_, row_, (int32_t) localIdx );
   }
   tnlmi_segments_view v_;
   Index row_;
   ValuePointer values_;
   ColumnPointer columns_;
};

//! The device view of a matrix (SparseMatrixView.h): plain values, passed to
//! kernels by copy.  getRow of a non-const view gives a RowView, of a const
//! one a ConstRowView, as the reference's.
template< typename Real, typename Index >
class DeviceMatrixView
{
public:
   using RealType = Real;
   using IndexType = Index;
   using RowView = DeviceRowView< Real, Index, false >;
   using ConstRowView = DeviceRowView< Real, Index, true >;

   DeviceMatrixView() = default;
   DeviceMatrixView( const tnlmi_segments_view& v, Index rows, Index columns, Real* values, int32_t* cols )
   : v_( v ), rows_( rows ), columns_( columns ), values_( values ), cols_( cols )
   {}
   __host__ __device__ Index getRows() const { return rows_; }
   __host__ __device__ Index getColumns() const { return columns_; }
   __host__ __device__ RowView getRow( Index r ) { return RowView( v_, r, values_, cols_ ); }
   __host__ __device__ ConstRowView getRow( Index r ) const { return ConstRowView( v_, r, values_, cols_ ); }
   __host__ __device__ Real* getValuesData() const { return values_; }
   __host__ __device__ int32_t* getColumnIndexesData() const { return cols_; }

private:
   tnlmi_segments_view v_{};
   Index rows_ = 0, columns_ = 0;
   Real* values_ = nullptr;
   int32_t* cols_ = nullptr;
};

// one thread per row of [begin, end), grid-stride
template< typename View, bool Const, typename Function >
__global__ void for_rows_kernel( View view, int64_t begin, int64_t end, Function f )
{
   using Index = typename View::IndexType;
   for( int64_t r = begin + (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < end;
        r += (int64_t) gridDim.x * blockDim.x ) {
      if constexpr( Const ) {
         const View& cv = view;
         auto row = cv.getRow( (Index) r );
         f( row );
      }
      else {
         auto row = view.getRow( (Index) r );
         f( row );
      }
   }
}

// one thread per row of [begin, end); the row's slots in order, each as
// ( row, local index, column, value ) -- by reference unless Const; a column
// index written by the function is stored back as 32 bits
template< typename View, bool Const, typename Function >
__global__ void for_elements_kernel( View view, int64_t begin, int64_t end, Function f )
{
   using Index = typename View::IndexType;
   using Real = typename View::RealType;
   for( int64_t r = begin + (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < end;
        r += (int64_t) gridDim.x * blockDim.x ) {
      auto row = view.getRow( (Index) r );
      const Index size = row.getSize();
      for( Index k = 0; k < size; k++ ) {
         if constexpr( Const ) {
            const Index c = row.getColumnIndex( k );
            const Real& v = row.getValue( k );
            f( (Index) r, k, c, v );
         }
         else {
            Index c = row.getColumnIndex( k );
            const Index c0 = c;
            f( (Index) r, k, c, row.getValue( k ) );
            if( c != c0 )
               row.setColumnIndex( k, c );
         }
      }
   }
}

template< typename Index, typename Function, typename... Args >
__global__ void parallel_for_kernel( Index begin, Index end, Function f, Args... args )
{
   for( Index i = begin + (Index) ( blockIdx.x * blockDim.x + threadIdx.x ); i < end;
        i += (Index) ( gridDim.x * blockDim.x ) )
      f( i, args... );
}

inline dim3 rows_grid( int64_t count )
{
   const int64_t blocks = ( count + 255 ) / 256;
   return dim3( (unsigned) std::max< int64_t >( 1, std::min< int64_t >( blocks, 65536 ) ) );
}

}  // namespace MI355XDetail

namespace Algorithms::detail {
//! parallelFor< Devices::MI355X >( begin, end, f, args... ): f( i, args... )
//! for every i of [begin, end) in a HIP kernel, synchronous on return as the
//! reference's GPU launch by default (ParallelFor1D.h:69-120)
template<>
struct ParallelFor1D< Devices::MI355X >
{
   template< typename Index, typename Function, typename... FunctionArgs >
   static void exec( Index begin, Index end, typename Devices::MI355X::LaunchConfiguration, Function f,
                     FunctionArgs... args )
   {
      if( end <= begin )
         return;
      hipLaunchKernelGGL( ( MI355XDetail::parallel_for_kernel< Index, Function, FunctionArgs... > ),
                          MI355XDetail::rows_grid( (int64_t) ( end - begin ) ), dim3( 256 ), 0, 0, begin, end, f,
                          args... );
      MI355XDetail::hip( hipGetLastError() );
      MI355XDetail::hip( hipDeviceSynchronize() );
   }
};
}  // namespace Algorithms::detail
#else
namespace Algorithms::detail {
//! Without a HIP compiler there is no device code for the loop body: a
//! parallelFor on Devices::MI355X does not compile (rather than run a device
//! lambda on the host)
template<>
struct ParallelFor1D< Devices::MI355X >
{
   template< typename Index, typename Function, typename... FunctionArgs >
   static void exec( Index, Index, typename Devices::MI355X::LaunchConfiguration, Function, FunctionArgs... )
   {
      static_assert( sizeof( Function ) == 0, "parallelFor< Devices::MI355X > needs code compiled by hipcc" );
   }
};
}  // namespace Algorithms::detail
#endif

// ------------------------------------------------------------------- matrix
namespace Matrices {

//! Matrices::SparseMatrix on the MI355X (SparseMatrix.h:41-56): general or
//! symmetric (lower triangle stored), any segments template of the
//! reference, FP64 or FP32 values, int indices.  Values and column indexes
//! live in device memory from RealAllocator / IndexAllocator.
template< typename Real, typename Index, typename MatrixType, template< typename, typename, typename > class Segments,
          typename ComputeReal, typename RealAllocator, typename IndexAllocator >
class SparseMatrix< Real, Devices::MI355X, Index, MatrixType, Segments, ComputeReal, RealAllocator, IndexAllocator >
{
   // Index = int or long.  The API's row, column and capacity values are of
   // type Index; the device arrays (offsets, column indexes) are 32-bit, as the
   // kernels read them: a matrix whose rows, columns or storage do not fit
   // int32 raises std::overflow_error, as the reference's Ellpack does for
   // its own int32 overflow (Ellpack.hpp:91-95).
   static_assert( std::is_same_v< Index, int > || std::is_same_v< Index, long >
                     || std::is_same_v< Index, long long >,
                  "SparseMatrix on Devices::MI355X supports Index = int or long" );
   using DeviceIndexVector = Containers::Vector< int32_t, Devices::MI355X, int32_t >;
   // host index values -> the 32-bit device array (std::overflow_error if one does not fit)
   static DeviceIndexVector toDevice32( const std::vector< Index >& v, const char* what )
   {
      DeviceIndexVector d;
      if constexpr( std::is_same_v< Index, int32_t > )
         d = v;
      else {
         std::vector< int32_t > w( v.size() );
         for( std::size_t i = 0; i < v.size(); i++ ) {
            if( v[ i ] < INT32_MIN || v[ i ] > INT32_MAX )
               throw std::overflow_error( std::string( what ) + ": the value " + std::to_string( v[ i ] )
                                          + " does not fit the 32-bit device indexes" );
            w[ i ] = (int32_t) v[ i ];
         }
         d = w;
      }
      return d;
   }
   static void checkDimension( Index n, const char* what )
   {
      if( n > (Index) INT32_MAX )
         throw std::overflow_error( std::string( what ) + " " + std::to_string( n )
                                    + " does not fit the 32-bit device indexes" );
   }
   // the segments type with this device (only named, never instantiated: its
   // template arguments -- e.g. the default ElementsOrganization, ColumnMajorOrder
   // on every device but Host (ElementsOrganization.h:21-32) -- select the format)
   using SegTag = MI355XDetail::SegmentsFormat< Segments< Devices::MI355X, Index, IndexAllocator > >;

public:
   using RealType = Real;
   using DeviceType = Devices::MI355X;
   using IndexType = Index;
   using ComputeRealType = ComputeReal;
   using RealAllocatorType = RealAllocator;
   using IndexAllocatorType = IndexAllocator;
   using ValuesVectorType = Containers::Vector< Real, Devices::MI355X, Index, RealAllocator >;
   using ColumnIndexesVectorType = Containers::Vector< Index, Devices::MI355X, Index, IndexAllocator >;
   using RowCapacitiesType = Containers::Vector< Index, Devices::MI355X, Index, IndexAllocator >;
   template< typename _Real = Real, typename _Device = Devices::MI355X, typename _Index = Index,
             typename _MatrixType = MatrixType,
             template< typename, typename, typename > class _Segments = Segments,
             typename _ComputeReal = ComputeReal,
             typename _RealAllocator = typename Allocators::Default< _Device >::template Allocator< _Real >,
             typename _IndexAllocator = typename Allocators::Default< _Device >::template Allocator< _Index > >
   using Self = SparseMatrix< _Real, _Device, _Index, _MatrixType, _Segments, _ComputeReal, _RealAllocator,
                              _IndexAllocator >;

   SparseMatrix() = default;
   SparseMatrix( Index rows, Index columns ) { setDimensions( rows, columns ); }
   //! SparseMatrix( rows, columns, { { row, column, value }, ... } ) (SparseMatrix.hpp:336-367)
   SparseMatrix( Index rows, Index columns, const std::initializer_list< std::tuple< Index, Index, Real > >& data )
   {
      setDimensions( rows, columns );
      std::map< std::pair< Index, Index >, Real > m;
      for( const auto& [ r, c, v ] : data )
         m[ { r, c } ] = v;
      setElements( m );
   }
   //! SparseMatrix( rows, columns, std::map ) (SparseMatrix.hpp:369-378)
   template< typename MapIndex, typename MapValue >
   SparseMatrix( Index rows, Index columns, const std::map< std::pair< MapIndex, MapIndex >, MapValue >& map )
   {
      setDimensions( rows, columns );
      setElements( map );
   }
   //! SparseMatrix( rowCapacities, columns ): empty rows of the given capacities
   SparseMatrix( const std::initializer_list< Index >& rowCapacities, Index columns )
   {
      setDimensions( (Index) rowCapacities.size(), columns );
      setRowCapacities( std::vector< Index >( rowCapacities ) );
   }
   SparseMatrix( const SparseMatrix& o ) { *this = o; }
   SparseMatrix& operator=( const SparseMatrix& o )
   {
      if( this == &o )
         return *this;
      if( o.segments_ )
         copyFrom( o );
      else
         setDimensions( o.rows_, o.columns_ );
      return *this;
   }
   //! cross-device assignment from the reference's host matrix
   //! (SparseMatrix.hpp:545-575): every row's stored (column, value) pairs in
   //! slot order, padding skipped
   template< typename R2, typename M2, template< typename, typename, typename > class S2, typename C2, typename RA2,
             typename IA2 >
   SparseMatrix& operator=( const SparseMatrix< R2, Devices::Host, Index, M2, S2, C2, RA2, IA2 >& host )
   {
      static_assert( M2::isSymmetric() == MatrixType::isSymmetric(),
                     "a general and a symmetric matrix cannot be assigned to each other" );
      const Index rows = host.getRows();
      std::vector< Index > off( rows + 1, 0 ), cols;
      std::vector< Real > vals;
      for( Index r = 0; r < rows; r++ ) {
         const auto row = host.getRow( r );
         for( Index k = 0; k < row.getSize(); k++ )
            if( row.getColumnIndex( k ) >= 0 ) {
               cols.push_back( row.getColumnIndex( k ) );
               vals.push_back( (Real) row.getValue( k ) );
            }
         off[ r + 1 ] = (Index) cols.size();
      }
      rows_ = rows;
      columns_ = host.getColumns();
      setFromCSR( off, cols, vals );
      return *this;
   }
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
      const tnlmi_matrix src = B.abi();
      release();
      rows_ = B.getRows();
      columns_ = B.getColumns();
      DeviceIndexVector len( std::max< int32_t >( 1, (int32_t) rows_ ) );
      MI355XDetail::check( tnlmi_matrix_row_lengths( &src, len.getData(), nullptr ) );
      MI355XDetail::check( tnlmi_segments_create( SegTag::format, SegTag::organization, SegTag::param,
                                                    len.getData(), rows_, nullptr, &segments_ ) );
      allocateStorage();
      tnlmi_matrix dst = abi();
      MI355XDetail::check( tnlmi_matrix_copy( &src, &dst, len.getData(), nullptr ) );
      mirror();
   }

   void setDimensions( Index rows, Index columns )
   {
      if( rows < 0 || columns < 0 )
         throw std::invalid_argument( "setDimensions: negative dimensions" );
      checkDimension( rows, "setDimensions: rows" );
      checkDimension( columns, "setDimensions: columns" );
      rows_ = rows;
      columns_ = columns;
      setRowCapacities( std::vector< Index >( rows, 0 ) );
   }
   //! setDimensions and row capacities of another matrix (SparseMatrix.hpp:238-255)
   template< typename Matrix_ >
   void setLike( const Matrix_& m )
   {
      rows_ = m.getRows();
      columns_ = m.getColumns();
      RowCapacitiesType caps;
      m.getRowCapacities( caps );
      setRowCapacities( caps );
   }
   Index getRows() const { return rows_; }
   Index getColumns() const { return columns_; }

   //! setRowCapacities (SparseMatrix.hpp:301-327): builds the format's
   //! segments; values = 0, column indexes = -1 (padding)
   void setRowCapacities( const std::vector< Index >& capacities )
   {
      if( (Index) capacities.size() != rows_ )
         throw std::invalid_argument( "setRowCapacities: the number of row capacities differs from the rows" );
      std::vector< Index > off( rows_ + 1, 0 );
      for( Index r = 0; r < rows_; r++ )
         off[ r + 1 ] = off[ r ] + capacities[ r ];
      std::vector< Index > cols( off[ rows_ ], -1 );
      std::vector< Real > vals( off[ rows_ ], Real( 0 ) );
      assemble( capacities, off, cols, vals );
   }
   template< typename A >
   void setRowCapacities( const Containers::Vector< Index, Devices::MI355X, Index, A >& capacities )
   {
      setRowCapacities( capacities.toHost() );
   }
   template< typename A >
   void setRowCapacities( const Containers::Vector< Index, Devices::Host, Index, A >& capacities )
   {
      setRowCapacities( std::vector< Index >( capacities.getData(), capacities.getData() + capacities.getSize() ) );
   }

   //! setElements( map ) (SparseMatrix.hpp:380-470, Complete encoding):
   //! capacities = entries per row; a symmetric matrix keeps the entries on
   //! and below the diagonal and checks that the others mirror them
   template< typename MapIndex, typename MapValue >
   void setElements( const std::map< std::pair< MapIndex, MapIndex >, MapValue >& m )
   {
      std::vector< Index > caps( rows_, 0 );
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
      std::vector< Index > off( rows_ + 1, 0 );
      for( Index r = 0; r < rows_; r++ )
         off[ r + 1 ] = off[ r ] + caps[ r ];
      std::vector< Index > cols( off[ rows_ ] ), pos( off.begin(), off.end() - 1 );
      std::vector< Real > vals( off[ rows_ ] );
      for( const auto& [ rc, v ] : m ) {  // std::map is ordered: columns ascend within a row
         if( MatrixType::isSymmetric() && rc.first < rc.second )
            continue;
         cols[ pos[ rc.first ] ] = (Index) rc.second;
         vals[ pos[ rc.first ]++ ] = (Real) v;
      }
      assemble( caps, off, cols, vals );
   }

   //! Assembly from CSR-ordered host arrays with row capacities taken from
   //! `offsets` (an extension of this device; the reference's device build
   //! fills a host matrix and copies it, SparseMatrix.hpp:360-367).
   void setFromCSR( const std::vector< Index >& offsets, const std::vector< Index >& columns,
                    const std::vector< Real >& values )
   {
      std::vector< Index > caps( rows_ );
      for( Index r = 0; r < rows_; r++ )
         caps[ r ] = offsets[ r + 1 ] - offsets[ r ];
      assemble( caps, offsets, columns, values );
   }

   //! vectorProduct (SparseMatrixBase.hpp:267-391): out = a*A*in + b*out on [begin, end)
   template< typename InVector, typename OutVector >
   void vectorProduct( const InVector& in, OutVector& out, ComputeReal matrixMultiplicator = 1.0,
                       ComputeReal outVectorMultiplicator = 0.0, Index begin = 0, Index end = 0,
                       const Algorithms::Segments::LaunchConfiguration& launchConfig = {} ) const
   {
      if( getColumns() != in.getSize() )
         throw std::invalid_argument( "vectorProduct: size of the input vector does not match the number of matrix columns" );
      if( getRows() != out.getSize() )
         throw std::invalid_argument( "vectorProduct: size of the output vector does not match the number of matrix rows" );
      using OutReal = std::remove_const_t< typename OutVector::RealType >;
      const tnlmi_matrix m = abi();
      const tnlmi_launch cfg = MI355XDetail::launch( launchConfig );
      MI355XDetail::check( tnlmi_spmv( &m, MI355XDetail::dtype< OutReal >(), in.getData(), out.getData(),
                                         MI355XDetail::dtype< ComputeReal >(), matrixMultiplicator,
                                         outVectorMultiplicator, begin, end, &cfg, nullptr ) );
      MI355XDetail::hip( hipDeviceSynchronize() );
   }
   template< typename InVector, typename OutVector >
   void vectorProduct( const InVector& in, OutVector& out, const Algorithms::Segments::LaunchConfiguration& cfg ) const
   {
      vectorProduct( in, out, 1.0, 0.0, 0, 0, cfg );
   }

   //! transposedVectorProduct (SparseMatrixBase.hpp:541-588): out = a*A^T*in + b*out on columns [begin, end)
   template< typename InVector, typename OutVector >
   void transposedVectorProduct( const InVector& in, OutVector& out, ComputeReal matrixMultiplicator = 1.0,
                                 ComputeReal outVectorMultiplicator = 0.0, Index begin = 0, Index end = 0 ) const
   {
      if( getRows() != in.getSize() )
         throw std::invalid_argument(
            "transposedVectorProduct: size of the input vector does not match the number of matrix rows" );
      if( getColumns() != out.getSize() )
         throw std::invalid_argument(
            "transposedVectorProduct: size of the output vector does not match the number of matrix columns" );
      using OutReal = std::remove_const_t< typename OutVector::RealType >;
      const tnlmi_matrix m = abi();
      MI355XDetail::check( tnlmi_spmv_transposed( &m, MI355XDetail::dtype< OutReal >(), in.getData(), out.getData(),
                                                    matrixMultiplicator, outVectorMultiplicator, begin, end,
                                                    nullptr ) );
      MI355XDetail::hip( hipDeviceSynchronize() );
   }

   //! getElement( row, column ) (SparseMatrixBase.hpp:231-265), 0 if absent;
   //! a symmetric matrix reads (column, row) for row < column
   Real getElement( Index row, Index column ) const
   {
      const tnlmi_matrix m = abi();
      double v = 0;
      MI355XDetail::check( tnlmi_matrix_get_element( &m, row, column, &v, nullptr ) );
      return (Real) v;
   }
   //! setElement( row, column, value ) (SparseMatrixBase.hpp:152-161)
   void setElement( Index row, Index column, const Real& value ) { addElement( row, column, value, Real( 0 ) ); }
   //! addElement( row, column, value, thisElementMultiplicator ) (SparseMatrixBase.hpp:166-231):
   //! an existing element becomes multiplicator * old + value; a new one is
   //! inserted in column order into the row's capacity (std::logic_error
   //! "The capacity of the sparse matrix row number r was exceeded." if full)
   void addElement( Index row, Index column, const Real& value, const Real& thisElementMultiplicator = 1.0 )
   {
      tnlmi_matrix m = abi();
      MI355XDetail::check( tnlmi_matrix_add_element( &m, row, column, value, thisElementMultiplicator, nullptr ) );
   }
   //! getRowCapacity( row ): the row's segment size (Ellpack: the width)
   Index getRowCapacity( Index row ) const { return hoff_[ row + 1 ] - hoff_[ row ]; }
   //! getRowCapacities (SparseMatrixBase.hpp:75-92): stored (non-padding) slots per row
   template< typename V >
   void getRowCapacities( V& out ) const
   {
      rowCounts( 1, out );
   }
   //! getCompressedRowLengths (SparseMatrixBase.hpp:55-73): stored nonzero values per row
   template< typename V >
   void getCompressedRowLengths( V& out ) const
   {
      rowCounts( 2, out );
   }
   //! getNonzeroElementsCount (SparseMatrixBase.hpp:102-132): a symmetric matrix
   //! counts its stored off-diagonal entries twice
   Index getNonzeroElementsCount() const
   {
      RowCapacitiesType c;
      rowCounts( MatrixType::isSymmetric() ? 3 : 2, c );
      int64_t total = 0;
      for( Index v : c.toHost() )
         total += v;
      return (Index) total;
   }
   Index getAllocatedElementsCount() const
   {
      return segments_ ? (Index) tnlmi_segments_storage_size( segments_ ) : 0;
   }
   //! the values and column indexes on the device (an extension of this
   //! device: writing through them changes the matrix)
   Real* getValuesData() const { return values_; }
   int32_t* getColumnIndexesData() const { return columnIndexes_; }

   //! getSerializationType (SparseMatrixBase.hpp:46-53)
   static std::string getSerializationType()
   {
      tnlmi_segments* empty = nullptr;
      MI355XDetail::check( tnlmi_segments_create( SegTag::format, SegTag::organization, SegTag::param, nullptr, 0,
                                                    nullptr, &empty ) );
      std::unique_ptr< tnlmi_segments, void ( * )( tnlmi_segments* ) > own( empty, tnlmi_segments_destroy );
      const tnlmi_matrix m{ 0, 0, empty, MI355XDetail::dtype< Real >(), nullptr, nullptr, matrixType() };
      int64_t n = 0;
      MI355XDetail::check( tnlmi_matrix_serialization_type( &m, nullptr, 0, &n ) );
      std::string t( (std::size_t) n + 1, '\0' );
      MI355XDetail::check( tnlmi_matrix_serialization_type( &m, t.data(), n + 1, &n ) );
      t.resize( (std::size_t) n );
      return t;
   }

   //! save( fileName ) (SparseMatrix.hpp:640-646): TNL's binary format
   void save( const std::string& fileName ) const
   {
      const tnlmi_matrix m = abi();
      MI355XDetail::check( tnlmi_matrix_save( &m, fileName.c_str(), nullptr ) );
   }

   //! load( fileName ) (SparseMatrix.hpp:656-694): a file of this matrix type
   //! (other types: "object type does not match"), its segments arrays as saved
   void load( const std::string& fileName )
   {
      tnlmi_matrix_file* f = nullptr;
      MI355XDetail::check( tnlmi_matrix_file_read( fileName.c_str(), &f ) );
      std::unique_ptr< tnlmi_matrix_file, void ( * )( tnlmi_matrix_file* ) > own( f, tnlmi_matrix_file_destroy );
      const std::string found = tnlmi_matrix_file_type( f ), expected = getSerializationType();
      if( found != expected )
         throw std::runtime_error( "Failed to deserialize an object from the file '" + fileName
                                   + "': object type does not match (expected " + expected + ", found " + found + ")." );
      int64_t rows = 0, columns = 0, storage = 0;
      MI355XDetail::check(
         tnlmi_matrix_file_info( f, &rows, &columns, nullptr, nullptr, nullptr, nullptr, &storage, nullptr ) );
      release();
      rows_ = (Index) rows;
      columns_ = (Index) columns;
      MI355XDetail::check( tnlmi_matrix_file_segments( f, nullptr, &segments_ ) );
      allocateStorage();
      MI355XDetail::check( tnlmi_matrix_file_upload( f, values_, columnIndexes_, nullptr ) );
      mirror();
   }

   //! The C ABI description of this matrix.
   tnlmi_matrix abi() const
   {
      return { rows_, columns_, segments_, MI355XDetail::dtype< Real >(), values_, columnIndexes_, matrixType() };
   }

#if defined( __HIPCC__ ) || defined( __HIP__ )
   // ---- rows and elements on the device (code compiled by hipcc) ----
   using ViewType = MI355XDetail::DeviceMatrixView< Real, Index >;
   using RowView = typename ViewType::RowView;
   using ConstRowView = typename ViewType::ConstRowView;
   //! getView (SparseMatrix.h): the device view, for getRow inside kernels
   ViewType getView() { return ViewType( segmentsView(), rows_, columns_, values_, columnIndexes_ ); }
   const ViewType getConstView() const
   {
      return ViewType( segmentsView(), rows_, columns_, values_, columnIndexes_ );
   }
   //! forRows( begin, end, f ): f( RowView& ) for every row of [begin, end)
   //! on the device (SparseMatrixBase.hpp:921-935)
   template< typename Function >
   void forRows( Index begin, Index end, Function&& function )
   {
      launchRows< false >( begin, end, function );
      modified();
   }
   template< typename Function >
   void forRows( Index begin, Index end, Function&& function ) const
   {
      launchRows< true >( begin, end, function );
   }
   template< typename Function >
   void forAllRows( Function&& function )
   {
      forRows( (Index) 0, rows_, function );
   }
   template< typename Function >
   void forAllRows( Function&& function ) const
   {
      forRows( (Index) 0, rows_, function );
   }
   //! forElements( begin, end, f ): f( rowIdx, localIdx, columnIdx, value )
   //! for every slot of the rows [begin, end), column and value by reference
   //! (SparseMatrixBase.hpp:746-766); free slots have column -1
   template< typename Function >
   void forElements( Index begin, Index end, Function&& function )
   {
      launchElements< false >( begin, end, function );
      modified();
   }
   template< typename Function >
   void forElements( Index begin, Index end, Function&& function ) const
   {
      launchElements< true >( begin, end, function );
   }
   template< typename Function >
   void forAllElements( Function&& function )
   {
      forElements( (Index) 0, rows_, function );
   }
   template< typename Function >
   void forAllElements( Function&& function ) const
   {
      forElements( (Index) 0, rows_, function );
   }

private:
   tnlmi_segments_view segmentsView() const
   {
      tnlmi_segments_view v{};
      if( segments_ )
         MI355XDetail::check( tnlmi_segments_get_view( segments_, &v ) );
      return v;
   }
   void checkRange( Index begin, Index end ) const
   {
      if( begin < 0 || end > rows_ || begin > end )
         throw std::invalid_argument( "row range [" + std::to_string( begin ) + ", " + std::to_string( end )
                                      + ") outside the matrix's " + std::to_string( rows_ ) + " rows" );
   }
   template< bool Const, typename Function >
   void launchRows( Index begin, Index end, Function& f ) const
   {
      checkRange( begin, end );
      if( begin == end || ! segments_ )
         return;
      hipLaunchKernelGGL( ( MI355XDetail::for_rows_kernel< ViewType, Const, std::decay_t< Function > > ),
                          MI355XDetail::rows_grid( end - begin ), dim3( 256 ), 0, 0, getConstView(),
                          (int64_t) begin, (int64_t) end, f );
      MI355XDetail::hip( hipGetLastError() );
      MI355XDetail::hip( hipDeviceSynchronize() );
   }
   template< bool Const, typename Function >
   void launchElements( Index begin, Index end, Function& f ) const
   {
      checkRange( begin, end );
      if( begin == end || ! segments_ )
         return;
      hipLaunchKernelGGL( ( MI355XDetail::for_elements_kernel< ViewType, Const, std::decay_t< Function > > ),
                          MI355XDetail::rows_grid( end - begin ), dim3( 256 ), 0, 0, getConstView(),
                          (int64_t) begin, (int64_t) end, f );
      MI355XDetail::hip( hipGetLastError() );
      MI355XDetail::hip( hipDeviceSynchronize() );
   }
   // the values or columns were written on the device: derived plans drop
   void modified()
   {
      const tnlmi_matrix m = abi();
      MI355XDetail::check( tnlmi_matrix_modified( &m ) );
   }

public:
#endif
   static constexpr bool isSymmetric() { return MatrixType::isSymmetric(); }
   static constexpr int32_t matrixType() { return MatrixType::isSymmetric() ? TNLMI_SYMMETRIC : TNLMI_GENERAL; }

private:
   template< typename V >
   void rowCounts( int kind, V& out ) const
   {
      DeviceIndexVector c( std::max< int32_t >( 1, (int32_t) rows_ ) );
      const tnlmi_matrix m = abi();
      MI355XDetail::check( tnlmi_matrix_row_counts( &m, kind, c.getData(), nullptr ) );
      using OutIndex = std::remove_const_t< typename V::RealType >;
      if constexpr( std::is_same_v< typename V::DeviceType, Devices::MI355X >
                    && std::is_same_v< OutIndex, int32_t > ) {
         out.setSize( rows_ );
         if( rows_ > 0 )
            MI355XDetail::hip( hipMemcpy( out.getData(), c.getData(), sizeof( int32_t ) * rows_,
                                          hipMemcpyDeviceToDevice ) );
      }
      else {
         const auto h = c.toHost();
         std::vector< OutIndex > w( (std::size_t) rows_ );
         for( Index r = 0; r < rows_; r++ )
            w[ r ] = (OutIndex) h[ r ];
         if constexpr( std::is_same_v< typename V::DeviceType, Devices::MI355X > )
            out = w;
         else {
            out.setSize( rows_ );
            for( Index r = 0; r < rows_; r++ )
               out[ r ] = w[ r ];
         }
      }
   }
   // the rows' segment sizes on the host (getRowCapacity)
   void mirror()
   {
      RowCapacitiesType c;
      rowCounts( 0, c );
      const std::vector< Index > sizes = c.toHost();
      hoff_.assign( rows_ + 1, 0 );
      for( Index r = 0; r < rows_; r++ )
         hoff_[ r + 1 ] = hoff_[ r ] + sizes[ r ];
   }
   void allocateStorage()
   {
      const std::size_t storage = (std::size_t) std::max< int64_t >( 1, tnlmi_segments_storage_size( segments_ ) );
      values_ = RealAllocator().allocate( storage );
      columnIndexes_ = Allocators::MI355X< int32_t >().allocate( storage );  // 32-bit on the device
      storage_ = storage;
   }
   void release()
   {
      if( segments_ )
         tnlmi_segments_destroy( segments_ );
      if( values_ )
         RealAllocator().deallocate( values_, storage_ );
      if( columnIndexes_ )
         Allocators::MI355X< int32_t >().deallocate( columnIndexes_, storage_ );
      segments_ = nullptr;
      values_ = nullptr;
      columnIndexes_ = nullptr;
      storage_ = 0;
   }
   void assemble( const std::vector< Index >& caps, const std::vector< Index >& off, const std::vector< Index >& cols,
                  const std::vector< Real >& vals )
   {
      release();
      DeviceIndexVector d_caps = toDevice32( caps, "row capacity" );
      DeviceIndexVector d_off = toDevice32( off, "storage offset" );
      DeviceIndexVector d_cols = toDevice32( cols, "column index" );
      ValuesVectorType d_vals;
      d_vals = vals;
      MI355XDetail::check( tnlmi_segments_create( SegTag::format, SegTag::organization, SegTag::param,
                                                    d_caps.getData(), rows_, nullptr, &segments_ ) );
      allocateStorage();
      MI355XDetail::check( tnlmi_matrix_fill( segments_, d_off.getData(), d_cols.getData(), d_vals.getData(),
                                                MI355XDetail::dtype< Real >(), values_,
                                                MI355XDetail::dtype< Real >(), columnIndexes_, nullptr ) );
      MI355XDetail::hip( hipDeviceSynchronize() );
      mirror();
   }

   Index rows_ = 0, columns_ = 0;
   tnlmi_segments* segments_ = nullptr;
   Real* values_ = nullptr;
   int32_t* columnIndexes_ = nullptr;
   std::size_t storage_ = 0;
   std::vector< Index > hoff_{ 0 };  // prefix sums of the row capacities
};

//! copySparseMatrix (Matrices/SparseOperations.hpp:716-721) between two
//! MI355X matrices of any formats: a more specialised overload than the
//! reference's generic template, so the call of the reference picks it.
template< typename R1, typename I, typename M1, template< typename, typename, typename > class S1, typename C1,
          typename RA1, typename IA1, typename R2, typename M2, template< typename, typename, typename > class S2,
          typename C2, typename RA2, typename IA2 >
void copySparseMatrix( SparseMatrix< R1, Devices::MI355X, I, M1, S1, C1, RA1, IA1 >& A,
                       const SparseMatrix< R2, Devices::MI355X, I, M2, S2, C2, RA2, IA2 >& B )
{
   A.copyFrom( B );
}

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
   using RealType = typename Matrix::RealType;
   using DeviceType = Devices::MI355X;
   using IndexType = typename Matrix::IndexType;

   static void readMtx( const std::string& fileName, Matrix& matrix, bool verbose = false )
   {
      tnlmi_host_csr* h = nullptr;
      MI355XDetail::check( tnlmi_mtx_read( fileName.c_str(), &h ) );
      load( h, matrix, verbose );
   }
   static void readMtx( std::istream& str, Matrix& matrix, bool verbose = false )
   {
      const std::string text( ( std::istreambuf_iterator< char >( str ) ), std::istreambuf_iterator< char >() );
      tnlmi_host_csr* h = nullptr;
      MI355XDetail::check( tnlmi_mtx_parse( text.data(), (int64_t) text.size(), &h ) );
      load( h, matrix, verbose );
   }

private:
   static void load( tnlmi_host_csr* h, Matrix& matrix, bool verbose )
   {
      std::unique_ptr< tnlmi_host_csr, void ( * )( tnlmi_host_csr* ) > own( h, tnlmi_host_csr_destroy );
      int64_t rows = 0, columns = 0, nnz = 0;
      const int32_t *off = nullptr, *cols = nullptr;
      const double* vals = nullptr;
      MI355XDetail::check( tnlmi_host_csr_get( h, &rows, &columns, &nnz, &off, &cols, &vals ) );
      if( verbose )
         std::cout << "Matrix dimensions are " << rows << " x " << columns << '\n';
      matrix.setDimensions( (IndexType) rows, (IndexType) columns );
      matrix.setFromCSR( std::vector< IndexType >( off, off + rows + 1 ), std::vector< IndexType >( cols, cols + nnz ),
                         std::vector< RealType >( vals, vals + nnz ) );
   }
};

}  // namespace Matrices

// ------------------------------------------------------------------ solvers
namespace MI355XDetail {
//! What a solver of this device needs of a preconditioner: its handle in the
//! library (Diagonal and ILU0 of this device have one).
struct DevicePreconditioner
{
   virtual ~DevicePreconditioner()
   {
      if( handle_ )
         tnlmi_preconditioner_destroy( handle_ );
   }
   const tnlmi_preconditioner* handle() const { return handle_; }

protected:
   template< typename MatrixPointer >
   void build( tnlmi_precond_kind kind, const MatrixPointer& matrix )
   {
      if( handle_ )
         tnlmi_preconditioner_destroy( handle_ );
      handle_ = nullptr;
      const tnlmi_matrix m = matrix->abi();
      check( tnlmi_preconditioner_create( kind, &m, nullptr, &handle_ ) );
   }
   void apply( const void* b, void* x ) const
   {
      if( ! handle_ )
         throw std::logic_error( "Preconditioner::solve: update() was not called" );
      check( tnlmi_preconditioner_solve( handle_, (const double*) b, (double*) x, nullptr ) );
      hip( hipDeviceSynchronize() );
   }
   tnlmi_preconditioner* handle_ = nullptr;
};

//! The shared part of this device's solvers: the reference's LinearSolver
//! (setup, setMatrix, setPreconditioner and the IterativeSolver settings),
//! the solver state handed to the library and the result read back.
template< typename Matrix >
class DeviceSolver : public Solvers::Linear::LinearSolver< Matrix >
{
protected:
   tnlmi_solver_state state() const
   {
      return { (int64_t) this->getMaxIterations(), (int64_t) this->getMinIterations(),
               (double) this->getConvergenceResidue(), (double) this->getDivergenceResidue(), 0, 0.0, 0 };
   }
   bool finish( const tnlmi_solver_state& st )
   {
      this->resetIterations();
      this->currentIteration = (typename Matrix::IndexType) st.iterations;
      this->setResidue( st.residue );
      return st.converged != 0;
   }
   const tnlmi_preconditioner* pre() const
   {
      if( ! this->preconditioner )
         return nullptr;
      const auto* p = dynamic_cast< const DevicePreconditioner* >( this->preconditioner.get() );
      if( ! p )
         throw Exceptions::NotImplementedError(
            "Devices::MI355X solvers take the Diagonal or ILU0 preconditioner of Devices::MI355X" );
      return p->handle();
   }
   tnlmi_matrix mat( const char* who ) const
   {
      if( ! this->matrix )
         throw std::logic_error( std::string( "No matrix was set in " ) + who + ". Call setMatrix() before solve()." );
      static_assert( std::is_same_v< typename Matrix::RealType, double >,
                     "the solvers of Devices::MI355X take FP64 vectors (a FP32 matrix is accepted through its "
                     "product's FP64 accumulation)" );
      return this->matrix->abi();
   }
};
}  // namespace MI355XDetail

#define TNLMI_SPARSE_MATRIX_PARAMS                                                                                    \
   typename Real_, typename Index_, typename MatrixType_, template< typename, typename, typename > class Segments_,   \
      typename ComputeReal_, typename RealAllocator_, typename IndexAllocator_
#define TNLMI_SPARSE_MATRIX                                                                                           \
   Matrices::SparseMatrix< Real_, Devices::MI355X, Index_, MatrixType_, Segments_, ComputeReal_, RealAllocator_,      \
                           IndexAllocator_ >

namespace Solvers::Linear {

namespace Preconditioners {

//! Preconditioners::Diagonal (Diagonal.hpp:12-43): d_i = A(i, i), x = b / d
template< TNLMI_SPARSE_MATRIX_PARAMS >
class Diagonal< TNLMI_SPARSE_MATRIX > : public Preconditioner< TNLMI_SPARSE_MATRIX >,
                                        public MI355XDetail::DevicePreconditioner
{
   using Base = Preconditioner< TNLMI_SPARSE_MATRIX >;

public:
   using typename Base::ConstVectorViewType;
   using typename Base::MatrixPointer;
   using typename Base::VectorViewType;
   void update( const MatrixPointer& matrix ) override { this->build( TNLMI_PRECOND_DIAGONAL, matrix ); }
   void solve( ConstVectorViewType b, VectorViewType x ) const override { this->apply( b.getData(), x.getData() ); }
};

//! Preconditioners::ILU0 (ILU0.hpp:12-139) through the reference's per-device
//! ILU0_impl hook (ILU0.h:26-28); the reference's GPU version throws
//! NotImplementedError, this device factorises on the host (the reference's
//! IKJ order) and solves on the GPU
template< typename Matrix, typename Real, typename Index >
class ILU0_impl< Matrix, Real, Devices::MI355X, Index > : public Preconditioner< Matrix >,
                                                          public MI355XDetail::DevicePreconditioner
{
   using Base = Preconditioner< Matrix >;

public:
   using typename Base::ConstVectorViewType;
   using typename Base::MatrixPointer;
   using typename Base::VectorViewType;
   void update( const MatrixPointer& matrix ) override { this->build( TNLMI_PRECOND_ILU0, matrix ); }
   void solve( ConstVectorViewType b, VectorViewType x ) const override { this->apply( b.getData(), x.getData() ); }
};

}  // namespace Preconditioners

//! CG (CG.hpp:10-108)
template< TNLMI_SPARSE_MATRIX_PARAMS >
class CG< TNLMI_SPARSE_MATRIX > : public MI355XDetail::DeviceSolver< TNLMI_SPARSE_MATRIX >
{
   using Base = MI355XDetail::DeviceSolver< TNLMI_SPARSE_MATRIX >;

public:
   using typename Base::ConstVectorViewType;
   using typename Base::VectorViewType;
   bool solve( ConstVectorViewType b, VectorViewType x ) override
   {
      const tnlmi_matrix m = this->mat( "CG" );
      tnlmi_solver_state st = this->state();
      MI355XDetail::check( tnlmi_cg_solve( &m, nullptr, this->pre(), b.getData(), x.getData(), &st, nullptr ) );
      return this->finish( st );
   }
};

//! GMRES (GMRES.hpp:18-180): variant and restarting parameters via setup()
template< TNLMI_SPARSE_MATRIX_PARAMS >
class GMRES< TNLMI_SPARSE_MATRIX > : public MI355XDetail::DeviceSolver< TNLMI_SPARSE_MATRIX >
{
   using Base = MI355XDetail::DeviceSolver< TNLMI_SPARSE_MATRIX >;

public:
   using typename Base::ConstVectorViewType;
   using typename Base::VectorViewType;
   bool setup( const Config::ParameterContainer& p, const String& prefix = "" ) override
   {
      if( p.checkParameter( prefix + "gmres-variant" ) ) {
         const String v = p.getParameter< String >( prefix + "gmres-variant" );
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
      return Base::setup( p, prefix );
   }
   bool solve( ConstVectorViewType b, VectorViewType x ) override
   {
      const tnlmi_matrix m = this->mat( "GMRES" );
      // GMRES.hpp:78-89
      if( rmin_ <= 0 || rmax_ <= 0 || rmin_ > rmax_ ) {
         std::cerr << "Wrong value for the GMRES restarting parameters: r_min = " << rmin_ << ", r_max = " << rmax_
                   << '\n';
         return false;
      }
      if( dmin_ < 0 || dmax_ < 0 || dmin_ > dmax_ ) {
         std::cerr << "Wrong value for the GMRES restarting adjustment parameters: d_min = " << dmin_
                   << ", d_max = " << dmax_ << '\n';
         return false;
      }
      tnlmi_solver_state st = this->state();
      MI355XDetail::check( tnlmi_gmres_solve( &m, nullptr, this->pre(), variant_, rmin_, rmax_, dmin_, dmax_,
                                                b.getData(), x.getData(), &st, nullptr ) );
      return this->finish( st );
   }

private:
   tnlmi_gmres_variant variant_ = TNLMI_GMRES_CWY;  // class default, GMRES.h:212
   int rmin_ = 10, rmax_ = 10, dmin_ = 3, dmax_ = 3;
};

//! BICGStab (BICGStab.hpp:12-127): bicgstab-exact-residue via setup()
template< TNLMI_SPARSE_MATRIX_PARAMS >
class BICGStab< TNLMI_SPARSE_MATRIX > : public MI355XDetail::DeviceSolver< TNLMI_SPARSE_MATRIX >
{
   using Base = MI355XDetail::DeviceSolver< TNLMI_SPARSE_MATRIX >;

public:
   using typename Base::ConstVectorViewType;
   using typename Base::VectorViewType;
   bool setup( const Config::ParameterContainer& p, const String& prefix = "" ) override
   {
      if( p.checkParameter( prefix + "bicgstab-exact-residue" ) )
         exact_residue_ = p.getParameter< bool >( prefix + "bicgstab-exact-residue" );
      return Base::setup( p, prefix );
   }
   bool solve( ConstVectorViewType b, VectorViewType x ) override
   {
      const tnlmi_matrix m = this->mat( "BICGStab" );
      tnlmi_solver_state st = this->state();
      MI355XDetail::check( tnlmi_bicgstab_solve( &m, nullptr, this->pre(), exact_residue_ ? 1 : 0, b.getData(),
                                                   x.getData(), &st, nullptr ) );
      return this->finish( st );
   }

private:
   bool exact_residue_ = false;  // configSetup default, BICGStab.hpp:17-20
};

//! TFQMR (TFQMR.hpp:12-121)
template< TNLMI_SPARSE_MATRIX_PARAMS >
class TFQMR< TNLMI_SPARSE_MATRIX > : public MI355XDetail::DeviceSolver< TNLMI_SPARSE_MATRIX >
{
   using Base = MI355XDetail::DeviceSolver< TNLMI_SPARSE_MATRIX >;

public:
   using typename Base::ConstVectorViewType;
   using typename Base::VectorViewType;
   bool solve( ConstVectorViewType b, VectorViewType x ) override
   {
      const tnlmi_matrix m = this->mat( "TFQMR" );
      tnlmi_solver_state st = this->state();
      MI355XDetail::check(
         tnlmi_tfqmr_solve( &m, nullptr, this->pre(), b.getData(), x.getData(), &st, nullptr ) );
      return this->finish( st );
   }
};

//! BICGStabL (BICGStabL.hpp:12-257): bicgstab-ell and bicgstab-exact-residue via setup()
template< TNLMI_SPARSE_MATRIX_PARAMS >
class BICGStabL< TNLMI_SPARSE_MATRIX > : public MI355XDetail::DeviceSolver< TNLMI_SPARSE_MATRIX >
{
   using Base = MI355XDetail::DeviceSolver< TNLMI_SPARSE_MATRIX >;

public:
   using typename Base::ConstVectorViewType;
   using typename Base::VectorViewType;
   bool setup( const Config::ParameterContainer& p, const String& prefix = "" ) override
   {
      if( p.checkParameter( prefix + "bicgstab-ell" ) )
         ell_ = p.getParameter< int >( prefix + "bicgstab-ell" );
      if( p.checkParameter( prefix + "bicgstab-exact-residue" ) )
         exact_residue_ = p.getParameter< bool >( prefix + "bicgstab-exact-residue" );
      return Base::setup( p, prefix );
   }
   bool solve( ConstVectorViewType b, VectorViewType x ) override
   {
      const tnlmi_matrix m = this->mat( "BICGStabL" );
      tnlmi_solver_state st = this->state();
      MI355XDetail::check( tnlmi_bicgstabl_solve( &m, nullptr, this->pre(), ell_, exact_residue_ ? 1 : 0,
                                                    b.getData(), x.getData(), &st, nullptr ) );
      return this->finish( st );
   }

private:
   int ell_ = 1;                 // configSetup default, BICGStabL.hpp:17
   bool exact_residue_ = false;  // BICGStabL.hpp:18-21
};

//! IDRs (IDRs.hpp:12-386): idr-s and idr-residual-smoothing via setup() or the setters
template< TNLMI_SPARSE_MATRIX_PARAMS >
class IDRs< TNLMI_SPARSE_MATRIX > : public MI355XDetail::DeviceSolver< TNLMI_SPARSE_MATRIX >
{
   using Base = MI355XDetail::DeviceSolver< TNLMI_SPARSE_MATRIX >;

public:
   using typename Base::ConstVectorViewType;
   using typename Base::VectorViewType;
   bool setup( const Config::ParameterContainer& p, const String& prefix = "" ) override
   {
      if( p.checkParameter( prefix + "idr-s" ) )
         s_ = p.getParameter< int >( prefix + "idr-s" );
      if( p.checkParameter( prefix + "idr-residual-smoothing" ) )
         smoothing_ = p.getParameter< bool >( prefix + "idr-residual-smoothing" );
      return Base::setup( p, prefix );
   }
   void setShadowSpaceDimension( int s ) { s_ = s; }
   void setResidualSmoothing( bool smoothing ) { smoothing_ = smoothing; }
   bool solve( ConstVectorViewType b, VectorViewType x ) override
   {
      const tnlmi_matrix m = this->mat( "IDRs" );
      tnlmi_solver_state st = this->state();
      MI355XDetail::check( tnlmi_idrs_solve( &m, nullptr, this->pre(), s_, smoothing_ ? 1 : 0, b.getData(),
                                               x.getData(), &st, nullptr ) );
      return this->finish( st );
   }

private:
   int s_ = 4;               // IDRs.h default shadow space dimension
   bool smoothing_ = false;  // IDRs.h default
};

//! Jacobi (Jacobi.hpp:65-125): jacobi-omega and residue-period via setup() or the setters
template< TNLMI_SPARSE_MATRIX_PARAMS >
class Jacobi< TNLMI_SPARSE_MATRIX > : public MI355XDetail::DeviceSolver< TNLMI_SPARSE_MATRIX >
{
   using Base = MI355XDetail::DeviceSolver< TNLMI_SPARSE_MATRIX >;

public:
   using typename Base::ConstVectorViewType;
   using typename Base::VectorViewType;
   using RealType = typename Base::RealType;
   using IndexType = typename Base::IndexType;
   bool setup( const Config::ParameterContainer& p, const String& prefix = "" ) override
   {
      if( p.checkParameter( prefix + "jacobi-omega" ) )
         setOmega( p.getParameter< double >( prefix + "jacobi-omega" ) );
      if( omega_ <= 0.0 || omega_ > 2.0 )
         std::cerr << "Warning: The Jacobi method parameter omega is out of interval (0,2). The value is " << omega_
                   << " the method will not converge.\n";
      if( p.checkParameter( prefix + "residue-period" ) )
         setResiduePeriod( p.getParameter< int >( prefix + "residue-period" ) );
      return Base::setup( p, prefix );
   }
   void setOmega( RealType omega ) { omega_ = omega; }
   RealType getOmega() const { return omega_; }
   void setResiduePeriod( IndexType period ) { period_ = period; }
   IndexType getResiduePerid() const { return period_; }
   bool solve( ConstVectorViewType b, VectorViewType x ) override
   {
      const tnlmi_matrix m = this->mat( "Jacobi" );
      tnlmi_solver_state st = this->state();
      MI355XDetail::check(
         tnlmi_jacobi_solve( &m, nullptr, omega_, period_, b.getData(), x.getData(), &st, nullptr ) );
      return this->finish( st );
   }

private:
   RealType omega_ = 1.0;  // Jacobi.h:121-123
   IndexType period_ = 4;
};

}  // namespace Solvers::Linear

#undef TNLMI_SPARSE_MATRIX
#undef TNLMI_SPARSE_MATRIX_PARAMS

}  // namespace TNL
