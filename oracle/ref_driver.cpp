/*
 * ref_driver.cpp -- a C-ABI wrapper around TNL's own Devices::Host code path,
 * compiled against the reference headers (TNL is header-only) by
 * oracle/Makefile into oracle/_ref/libtnlref.so.
 *
 * TEST INFRASTRUCTURE ONLY.  It pins the CPU restatement in oracle.cpp and
 * serves as the measured CPU baseline ("kind": "reference").  It contains no
 * reference source: it only instantiates the reference's templates.
 */
#include <omp.h>

#include <TNL/Matrices/SparseMatrix.h>
#include <TNL/Algorithms/Segments/CSR.h>
#include <TNL/Algorithms/Segments/AdaptiveCSR.h>
#include <TNL/Algorithms/Segments/Ellpack.h>
#include <TNL/Algorithms/Segments/SlicedEllpack.h>
#include <TNL/Algorithms/Segments/BiEllpack.h>
#include <TNL/Algorithms/Segments/ChunkedEllpack.h>
#include <TNL/Algorithms/Segments/SortedSegments.h>
#include <TNL/Solvers/Linear/CG.h>
#include <TNL/Solvers/Linear/GMRES.h>
#include <TNL/Solvers/Linear/BICGStab.h>
#include <TNL/Solvers/Linear/TFQMR.h>
#include <TNL/Solvers/Linear/BICGStabL.h>
#include <TNL/Solvers/Linear/IDRs.h>
#include <TNL/Solvers/Linear/Jacobi.h>
#include <TNL/Solvers/Linear/Preconditioners/Diagonal.h>
#include <TNL/Solvers/Linear/Preconditioners/ILU0.h>
#include <TNL/Solvers/Linear/Utils/LinearResidueGetter.h>
#include <TNL/Matrices/MatrixReader.h>
#include <TNL/Matrices/SparseOperations.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

using namespace TNL;
namespace Seg = TNL::Algorithms::Segments;
using Host = Devices::Host;
using IVec = Containers::Vector< int, Host, int >;
using DVec = Containers::Vector< double, Host, int >;

template< class D, class I, class A >
using EllRM = Seg::Ellpack< D, I, A, Seg::RowMajorOrder, 32 >;
template< class D, class I, class A >
using EllCM = Seg::Ellpack< D, I, A, Seg::ColumnMajorOrder, 32 >;
template< class D, class I, class A >
using Sl32RM = Seg::SlicedEllpack< D, I, A, Seg::RowMajorOrder, 32 >;
template< class D, class I, class A >
using Sl32CM = Seg::SlicedEllpack< D, I, A, Seg::ColumnMajorOrder, 32 >;
template< class D, class I, class A >
using Sl64RM = Seg::SlicedEllpack< D, I, A, Seg::RowMajorOrder, 64 >;
template< class D, class I, class A >
using Sl64CM = Seg::SlicedEllpack< D, I, A, Seg::ColumnMajorOrder, 64 >;
template< class D, class I, class A >
using BiRM = Seg::BiEllpack< D, I, A, Seg::RowMajorOrder, 32 >;
template< class D, class I, class A >
using BiCM = Seg::BiEllpack< D, I, A, Seg::ColumnMajorOrder, 32 >;
template< class D, class I, class A >
using ChRM = Seg::ChunkedEllpack< D, I, A, Seg::RowMajorOrder >;
template< class D, class I, class A >
using ChCM = Seg::ChunkedEllpack< D, I, A, Seg::ColumnMajorOrder >;
template< class D, class I, class A >
using SortedCSR = Seg::SortedSegments< Seg::CSR< D, I, A > >;

template< template< class, class, class > class S >
using Mat = Matrices::SparseMatrix< double, Host, int, Matrices::GeneralMatrix, S >;
using CsrMat = Mat< Seg::CSR >;

namespace {

IVec to_ivec( const int32_t* p, int64_t n )
{
   IVec v( n );
   for( int64_t i = 0; i < n; i++ )
      v[ i ] = p[ i ];
   return v;
}

// Builds an n x ncols matrix of segments type S from CSR capacity arrays
// (row r owns cols[row_ptr[r] .. row_ptr[r+1]), column -1 = unused slot).
template< class M >
void build( M& m, int64_t n, int64_t ncols, const int32_t* row_ptr, const int32_t* cols, const double* vals )
{
   m.setDimensions( n, ncols );
   IVec caps( n );
   for( int64_t r = 0; r < n; r++ )
      caps[ r ] = row_ptr[ r + 1 ] - row_ptr[ r ];
   m.setRowCapacities( caps );
   for( int64_t r = 0; r < n; r++ ) {
      auto row = m.getRow( r );
      int k = 0;
      for( int32_t g = row_ptr[ r ]; g < row_ptr[ r + 1 ]; g++ )
         if( cols[ g ] >= 0 )
            row.setElement( k++, cols[ g ], vals[ g ] );
   }
}

template< class M >
int64_t fill_impl( int64_t n, int64_t ncols, const int32_t* row_ptr, const int32_t* cols, const double* vals,
                   int32_t* out_cols, double* out_vals, int64_t cap )
{
   M m;
   build( m, n, ncols, row_ptr, cols, vals );
   const int64_t storage = m.getValues().getSize();
   if( storage <= cap )
      for( int64_t i = 0; i < storage; i++ ) {
         out_cols[ i ] = m.getColumnIndexes()[ i ];
         out_vals[ i ] = m.getValues()[ i ];
      }
   return storage;
}

// message of the last exception a driver call caught
static std::string g_err;

// setRowCapacities from the CSR arrays (columns -1: empty rows of that
// capacity), then addElement( row, column, value, multiplicator ) per op until
// one throws (its index in *failed, its message in g_err; -1 if none); the
// storage arrays as fill_impl exports them
template< class M >
int64_t elements_impl( int64_t n, int64_t ncols, const int32_t* row_ptr, const int32_t* cols, const double* vals,
                       const int64_t* op_pos, const double* op_val, int64_t nops, int64_t* failed, int32_t* out_cols,
                       double* out_vals, int64_t cap )
{
   M m;
   build( m, n, ncols, row_ptr, cols, vals );
   *failed = -1;
   for( int64_t i = 0; i < nops; i++ ) {
      try {
         m.addElement( (int) op_pos[ 2 * i ], (int) op_pos[ 2 * i + 1 ], op_val[ 2 * i ], op_val[ 2 * i + 1 ] );
      }
      catch( const std::logic_error& e ) {
         g_err = e.what();
         *failed = i;
         break;
      }
   }
   const int64_t storage = m.getValues().getSize();
   if( storage <= cap )
      for( int64_t i = 0; i < storage; i++ ) {
         out_cols[ i ] = m.getColumnIndexes()[ i ];
         out_vals[ i ] = m.getValues()[ i ];
      }
   return storage;
}

template< class M >
void spmv_impl( int64_t n, int64_t ncols, const int32_t* row_ptr, const int32_t* cols, const double* vals,
                const double* x, double* y, double alpha, double beta, int64_t begin, int64_t end )
{
   M m;
   build( m, n, ncols, row_ptr, cols, vals );
   DVec xv( ncols ), yv( n );
   for( int64_t i = 0; i < ncols; i++ )
      xv[ i ] = x[ i ];
   for( int64_t i = 0; i < n; i++ )
      yv[ i ] = y[ i ];
   m.vectorProduct( xv, yv, alpha, beta, begin, end );
   for( int64_t i = 0; i < n; i++ )
      y[ i ] = yv[ i ];
}

// format: 0 CSR, 1 AdaptiveCSR, 2 Ellpack, 3 SlicedEllpack32, 4 SlicedEllpack64,
// 5 BiEllpack, 6 ChunkedEllpack, 7 SortedCSR; col_major selects the
// column-major organisation (the reference's GPU default).
template< template< template< class, class, class > class > class F, class... Args >
int64_t dispatch( int format, int col_major, Args... args )
{
   switch( format ) {
      case 0:
         return F< Seg::CSR >::run( args... );
      case 1:
         return F< Seg::AdaptiveCSR >::run( args... );
      case 2:
         return col_major ? F< EllCM >::run( args... ) : F< EllRM >::run( args... );
      case 3:
         return col_major ? F< Sl32CM >::run( args... ) : F< Sl32RM >::run( args... );
      case 4:
         return col_major ? F< Sl64CM >::run( args... ) : F< Sl64RM >::run( args... );
      case 5:
         return col_major ? F< BiCM >::run( args... ) : F< BiRM >::run( args... );
      case 6:
         return col_major ? F< ChCM >::run( args... ) : F< ChRM >::run( args... );
      case 7:
         return F< SortedCSR >::run( args... );
   }
   return -1;
}

template< template< class, class, class > class S >
struct FillOp
{
   static int64_t run( int64_t n, int64_t ncols, const int32_t* rp, const int32_t* c, const double* v, int32_t* oc,
                       double* ov, int64_t cap )
   {
      return fill_impl< Mat< S > >( n, ncols, rp, c, v, oc, ov, cap );
   }
};
template< template< class, class, class > class S >
struct SpmvOp
{
   static int64_t run( int64_t n, int64_t ncols, const int32_t* rp, const int32_t* c, const double* v, const double* x,
                       double* y, double a, double b, int64_t be, int64_t en )
   {
      spmv_impl< Mat< S > >( n, ncols, rp, c, v, x, y, a, b, be, en );
      return 0;
   }
};

// copySparseMatrix( dst, src ) (SparseOperations.hpp:600-690) from a source of
// format S built from the CSR arrays into a target of the dispatched format;
// the target's storage arrays as fill_impl exports them
template< template< class, class, class > class S >
struct CopyOp
{
   template< template< class, class, class > class T >
   struct Into
   {
      static int64_t run( const Mat< S >* src, int32_t* oc, double* ov, int64_t cap )
      {
         Mat< T > dst;
         Matrices::copySparseMatrix( dst, *src );
         const int64_t storage = dst.getValues().getSize();
         if( storage <= cap )
            for( int64_t i = 0; i < storage; i++ ) {
               oc[ i ] = dst.getColumnIndexes()[ i ];
               ov[ i ] = dst.getValues()[ i ];
            }
         return storage;
      }
   };
   static int64_t run( int64_t n, int64_t ncols, const int32_t* rp, const int32_t* c, const double* v, int dst_format,
                       int dst_col_major, int32_t* oc, double* ov, int64_t cap )
   {
      Mat< S > src;
      build( src, n, ncols, rp, c, v );
      return dispatch< CopyOp< S >::template Into >( dst_format, dst_col_major, &src, oc, ov, cap );
   }
};

template< template< class, class, class > class S >
using SymMat = Matrices::SparseMatrix< double, Host, int, Matrices::SymmetricMatrix, S >;

// a SymmetricMatrix built from lower-triangle CSR arrays, vectorProduct
template< template< class, class, class > class S >
struct SymSpmvOp
{
   static int64_t run( int64_t n, int64_t ncols, const int32_t* rp, const int32_t* c, const double* v, const double* x,
                       double* y, double a, double b, int64_t be, int64_t en )
   {
      spmv_impl< SymMat< S > >( n, ncols, rp, c, v, x, y, a, b, be, en );
      return 0;
   }
};

// transposedVectorProduct of a general matrix
template< template< class, class, class > class S >
struct TransSpmvOp
{
   static int64_t run( int64_t n, int64_t ncols, const int32_t* rp, const int32_t* c, const double* v, const double* x,
                       double* y, double a, double b, int64_t be, int64_t en )
   {
      Mat< S > m;
      build( m, n, ncols, rp, c, v );
      DVec xv( n ), yv( ncols );
      for( int64_t i = 0; i < n; i++ )
         xv[ i ] = x[ i ];
      for( int64_t i = 0; i < ncols; i++ )
         yv[ i ] = y[ i ];
      m.transposedVectorProduct( xv, yv, a, b, be, en );
      for( int64_t i = 0; i < ncols; i++ )
         y[ i ] = yv[ i ];
      return 0;
   }
};

template< template< class, class, class > class S >
struct ElementsOp
{
   static int64_t run( int64_t n, int64_t ncols, const int32_t* rp, const int32_t* c, const double* v,
                       const int64_t* op, const double* ov, int64_t nops, int64_t* failed, int32_t* oc, double* os,
                       int64_t cap )
   {
      return elements_impl< Mat< S > >( n, ncols, rp, c, v, op, ov, nops, failed, oc, os, cap );
   }
};
template< template< class, class, class > class S >
struct SymElementsOp
{
   static int64_t run( int64_t n, int64_t ncols, const int32_t* rp, const int32_t* c, const double* v,
                       const int64_t* op, const double* ov, int64_t nops, int64_t* failed, int32_t* oc, double* os,
                       int64_t cap )
   {
      return elements_impl< SymMat< S > >( n, ncols, rp, c, v, op, ov, nops, failed, oc, os, cap );
   }
};

// getRowCapacity( r ) per row, getRowCapacities, getCompressedRowLengths and
// getNonzeroElementsCount of the matrix build() makes (SparseMatrixBase.hpp:55-132)
template< class M >
int64_t counts_impl( int64_t n, int64_t ncols, const int32_t* rp, const int32_t* c, const double* v, int32_t* cap,
                     int32_t* stored, int32_t* nonzero )
{
   M m;
   build( m, n, ncols, rp, c, v );
   IVec a, b;
   m.getRowCapacities( a );
   m.getCompressedRowLengths( b );
   for( int64_t r = 0; r < n; r++ ) {
      cap[ r ] = m.getRowCapacity( r );
      stored[ r ] = a[ r ];
      nonzero[ r ] = b[ r ];
   }
   return m.getNonzeroElementsCount();
}
template< template< class, class, class > class S >
struct CountsOp
{
   static int64_t run( int64_t n, int64_t ncols, const int32_t* rp, const int32_t* c, const double* v, int32_t* a,
                       int32_t* b, int32_t* d )
   {
      return counts_impl< Mat< S > >( n, ncols, rp, c, v, a, b, d );
   }
};
template< template< class, class, class > class S >
struct SymCountsOp
{
   static int64_t run( int64_t n, int64_t ncols, const int32_t* rp, const int32_t* c, const double* v, int32_t* a,
                       int32_t* b, int32_t* d )
   {
      return counts_impl< SymMat< S > >( n, ncols, rp, c, v, a, b, d );
   }
};

template< template< class, class, class > class S >
struct SymSaveOp
{
   static int64_t run( int64_t n, int64_t ncols, const int32_t* rp, const int32_t* c, const double* v,
                       const char* path )
   {
      SymMat< S > m;
      build( m, n, ncols, rp, c, v );
      m.save( path );
      return 0;
   }
};

template< template< class, class, class > class S >
struct SaveOp
{
   static int64_t run( int64_t n, int64_t ncols, const int32_t* rp, const int32_t* c, const double* v,
                       const char* path )
   {
      Mat< S > m;
      build( m, n, ncols, rp, c, v );
      m.save( path );
      return 0;
   }
};
template< template< class, class, class > class S >
struct LoadOp
{
   static int64_t run( const char* path, int64_t* dims, int32_t* oc, double* ov, int64_t cap )
   {
      Mat< S > m;
      m.load( path );
      dims[ 0 ] = m.getRows();
      dims[ 1 ] = m.getColumns();
      const int64_t storage = m.getValues().getSize();
      if( storage <= cap )
         for( int64_t i = 0; i < storage; i++ ) {
            oc[ i ] = m.getColumnIndexes()[ i ];
            ov[ i ] = m.getValues()[ i ];
         }
      return storage;
   }
};

struct State
{
   int64_t max_iterations, min_iterations;
   double convergence_residue, divergence_residue;
   int64_t iterations;
   double residue;
   int converged;
};

template< class Solver >
void run_solver( Solver& solver, const std::shared_ptr< CsrMat >& A, int precond, const double* b, double* x,
                 int64_t n, State* st )
{
   using namespace Solvers::Linear::Preconditioners;
   solver.setMatrix( A );
   solver.setMaxIterations( st->max_iterations );
   solver.setMinIterations( st->min_iterations );
   solver.setConvergenceResidue( st->convergence_residue );
   solver.setDivergenceResidue( st->divergence_residue );
   std::shared_ptr< Diagonal< CsrMat > > diag;
   std::shared_ptr< ILU0< CsrMat > > ilu;
   if( precond == 1 ) {
      diag = std::make_shared< Diagonal< CsrMat > >();
      diag->update( A );
      solver.setPreconditioner( diag );
   }
   else if( precond == 2 ) {
      ilu = std::make_shared< ILU0< CsrMat > >();
      ilu->update( A );
      solver.setPreconditioner( ilu );
   }
   DVec bv( n ), xv( n );
   for( int64_t i = 0; i < n; i++ ) {
      bv[ i ] = b[ i ];
      xv[ i ] = x[ i ];
   }
   st->converged = solver.solve( bv.getConstView(), xv.getView() );
   st->iterations = solver.getIterations();
   st->residue = solver.getResidue();
   for( int64_t i = 0; i < n; i++ )
      x[ i ] = xv[ i ];
}

std::shared_ptr< CsrMat > make_csr( int64_t n, const int32_t* rp, const int32_t* c, const double* v )
{
   auto A = std::make_shared< CsrMat >();
   build( *A, n, n, rp, c, v );
   return A;
}

}  // namespace

extern "C" {

void ref_csr_offsets( const int32_t* sizes, int64_t n, int32_t* offsets )
{
   Seg::CSR< Host, int > s;
   s.setSegmentsSizes( to_ivec( sizes, n ) );
   for( int64_t i = 0; i <= n; i++ )
      offsets[ i ] = s.getOffsets()[ i ];
}

int ref_ellpack_dims( const int32_t* sizes, int64_t n, int col_major, int32_t* w, int32_t* aligned )
{
   try {
      if( col_major ) {
         EllCM< Host, int, Allocators::Default< Host >::Allocator< int > > s;
         s.setSegmentsSizes( to_ivec( sizes, n ) );
         *w = s.getSegmentSize( 0 );
         *aligned = s.getAlignedSize();
      }
      else {
         EllRM< Host, int, Allocators::Default< Host >::Allocator< int > > s;
         s.setSegmentsSizes( to_ivec( sizes, n ) );
         *w = s.getSegmentSize( 0 );
         *aligned = s.getAlignedSize();
      }
   }
   catch( const std::overflow_error& ) {
      return 3;
   }
   return 0;
}

void ref_sliced_ellpack( const int32_t* sizes, int64_t n, int slice_size, int32_t* offsets, int32_t* seg_sizes )
{
   auto copy_out = [ & ]( auto& s ) {
      for( int64_t i = 0; i < s.getSliceOffsetsView().getSize(); i++ )
         offsets[ i ] = s.getSliceOffsetsView()[ i ];
      for( int64_t i = 0; i < s.getSliceSegmentSizesView().getSize(); i++ )
         seg_sizes[ i ] = s.getSliceSegmentSizesView()[ i ];
   };
   if( slice_size == 32 ) {
      Sl32CM< Host, int, Allocators::Default< Host >::Allocator< int > > s;
      s.setSegmentsSizes( to_ivec( sizes, n ) );
      copy_out( s );
   }
   else {
      Sl64CM< Host, int, Allocators::Default< Host >::Allocator< int > > s;
      s.setSegmentsSizes( to_ivec( sizes, n ) );
      copy_out( s );
   }
}

void ref_biellpack( const int32_t* sizes, int64_t n, int32_t* perm, int32_t* group_ptrs )
{
   BiCM< Host, int, Allocators::Default< Host >::Allocator< int > > s;
   s.setSegmentsSizes( to_ivec( sizes, n ) );
   for( int64_t i = 0; i < s.getSegmentsPermutationView().getSize(); i++ )
      perm[ i ] = s.getSegmentsPermutationView()[ i ];
   for( int64_t i = 0; i < s.getGroupPointersView().getSize(); i++ )
      group_ptrs[ i ] = s.getGroupPointersView()[ i ];
}

int64_t ref_chunked_ellpack( const int32_t* sizes, int64_t n, int32_t* slices, int32_t* seg_to_chunk,
                             int32_t* seg_to_slice, int32_t* seg_ptrs, int32_t* chunks_to_seg, int64_t chunks_cap,
                             int32_t* storage )
{
   ChCM< Host, int, Allocators::Default< Host >::Allocator< int > > s;
   s.setSegmentsSizes( to_ivec( sizes, n ) );
   const int64_t nslices = s.getSegmentToSliceMappingView().getSize() ? s.getSegmentToSliceMappingView()[ n - 1 ] + 1 : 0;
   for( int64_t i = 0; i < nslices; i++ ) {
      const auto& si = s.getSlicesView()[ i ];
      slices[ 4 * i + 0 ] = si.size;
      slices[ 4 * i + 1 ] = si.chunkSize;
      slices[ 4 * i + 2 ] = si.firstSegment;
      slices[ 4 * i + 3 ] = si.pointer;
   }
   for( int64_t i = 0; i < n; i++ ) {
      seg_to_chunk[ i ] = s.getSegmentToChunkMappingView()[ i ];
      seg_to_slice[ i ] = s.getSegmentToSliceMappingView()[ i ];
   }
   for( int64_t i = 0; i <= n; i++ )
      seg_ptrs[ i ] = s.getSegmentPointersView()[ i ];
   const int64_t nc = s.getChunksToSegmentsMappingView().getSize();
   if( nc > chunks_cap )
      return -1;
   // only the entries the reference writes (one per allocated chunk) are defined
   int64_t defined = 0;
   for( int64_t i = 0; i < n; i++ )
      defined += s.getSegmentToChunkMappingView()[ i ]
               - ( i == s.getSlicesView()[ seg_to_slice[ i ] ].firstSegment ? 0 : s.getSegmentToChunkMappingView()[ i - 1 ] );
   for( int64_t i = 0; i < defined; i++ )
      chunks_to_seg[ i ] = s.getChunksToSegmentsMappingView()[ i ];
   *storage = s.getStorageSize();
   return nslices;
}

int64_t ref_adaptive_csr_blocks( const int32_t* sizes, int64_t n, int value_size, int32_t* blocks, int64_t cap )
{
   Seg::AdaptiveCSR< Host, int > s;
   try {
      s.setSegmentsSizes( to_ivec( sizes, n ) );
   }
   catch( const std::exception& ) {
      // a host-only build cannot produce LONG descriptors: their padding
      // queries the GPU's warp size (Backend::getDevice() throws
      // BackendSupportMissing without a GPU backend)
      return -2;
   }
   int log = value_size == 1 ? 0 : value_size == 2 ? 1 : value_size <= 4 ? 2 : value_size <= 8 ? 3 : value_size <= 16 ? 4 : 5;
   const auto& b = s.getBlocks()[ log ];
   if( b.getSize() > cap )
      return -1;
   for( int64_t i = 0; i < b.getSize(); i++ ) {
      const auto& d = b[ i ];
      const int type = (int) d.getType();
      int32_t* o = blocks + 6 * i;
      o[ 0 ] = type;
      o[ 1 ] = d.getFirstSegment();
      // fields left uninitialised by the reference's constructors are reported as 0
      const bool sized = type != 0 && i != b.getSize() - 1;
      o[ 2 ] = sized ? d.getSize() : 0;
      o[ 3 ] = sized ? d.getSegmentsInBlock() : 0;
      o[ 4 ] = d.getWarpIdx();
      o[ 5 ] = type == 0 ? d.getWarpsCount() : 0;
   }
   return b.getSize();
}

void ref_sorted_segments( const int32_t* sizes, int64_t n, int32_t* embedded_offsets, int32_t* inverse_perm,
                          int32_t* perm )
{
   SortedCSR< Host, int, Allocators::Default< Host >::Allocator< int > > s;
   s.setSegmentsSizes( to_ivec( sizes, n ) );
   for( int64_t i = 0; i < n; i++ ) {
      inverse_perm[ i ] = s.getInverseSegmentsPermutationView()[ i ];
      perm[ i ] = s.getSegmentsPermutationView()[ i ];
   }
   for( int64_t i = 0; i <= n; i++ )
      embedded_offsets[ i ] = s.getEmbeddedSegments().getOffsets()[ i ];
}

int64_t ref_fill( int format, int col_major, int64_t n, int64_t ncols, const int32_t* row_ptr, const int32_t* cols,
                  const double* vals, int32_t* out_cols, double* out_vals, int64_t cap )
{
   return dispatch< FillOp >( format, col_major, n, ncols, row_ptr, cols, vals, out_cols, out_vals, cap );
}

// copySparseMatrix from a matrix of format src_format (built from the CSR
// arrays) into one of format dst_format: the target's storage size; its
// column indexes and values when they fit cap
int64_t ref_copy( int src_format, int src_col_major, int dst_format, int dst_col_major, int64_t n, int64_t ncols,
                  const int32_t* row_ptr, const int32_t* cols, const double* vals, int32_t* out_cols,
                  double* out_vals, int64_t cap )
{
   return dispatch< CopyOp >( src_format, src_col_major, n, ncols, row_ptr, cols, vals, dst_format, dst_col_major,
                              out_cols, out_vals, cap );
}

int64_t ref_spmv( int format, int col_major, int64_t n, int64_t ncols, const int32_t* row_ptr, const int32_t* cols,
                  const double* vals, const double* x, double* y, double alpha, double beta, int64_t begin,
                  int64_t end )
{
   return dispatch< SpmvOp >( format, col_major, n, ncols, row_ptr, cols, vals, x, y, alpha, beta, begin, end );
}

void ref_spmv_csr_f32( int64_t n, int64_t ncols, const int32_t* row_ptr, const int32_t* cols, const float* vals,
                       const float* x, float* y )
{
   Matrices::SparseMatrix< float, Host, int > m;
   m.setDimensions( n, ncols );
   IVec caps( n );
   for( int64_t r = 0; r < n; r++ )
      caps[ r ] = row_ptr[ r + 1 ] - row_ptr[ r ];
   m.setRowCapacities( caps );
   for( int64_t g = 0; g < row_ptr[ n ]; g++ ) {
      m.getColumnIndexes()[ g ] = cols[ g ];
      m.getValues()[ g ] = vals[ g ];
   }
   Containers::Vector< float, Host, int > xv( ncols ), yv( n );
   for( int64_t i = 0; i < ncols; i++ )
      xv[ i ] = x[ i ];
   m.vectorProduct( xv, yv );
   for( int64_t i = 0; i < n; i++ )
      y[ i ] = yv[ i ];
}

/* TNL Host CSR vectorProduct (OpenMP, schedule(dynamic,100)) on a CSR matrix
 * whose arrays are copied in once; reps timed products; returns the median
 * seconds per product and leaves the last product in y. */
double ref_spmv_csr_time( int64_t n, int64_t ncols, const int32_t* row_ptr, const int32_t* cols, const double* vals,
                          const double* x, double* y, int reps, int threads )
{
   Host::setMaxThreadsCount( threads );
   CsrMat m;
   m.setDimensions( n, ncols );
   IVec caps( n );
   for( int64_t r = 0; r < n; r++ )
      caps[ r ] = row_ptr[ r + 1 ] - row_ptr[ r ];
   m.setRowCapacities( caps );
   const int64_t nnz = row_ptr[ n ];
   std::copy( cols, cols + nnz, m.getColumnIndexes().getData() );
   std::copy( vals, vals + nnz, m.getValues().getData() );
   DVec xv( ncols ), yv( n );
   std::copy( x, x + ncols, xv.getData() );
   m.vectorProduct( xv, yv );  // warm-up
   std::vector< double > t;
   for( int i = 0; i < reps; i++ ) {
      const auto t0 = std::chrono::steady_clock::now();
      m.vectorProduct( xv, yv );
      const auto t1 = std::chrono::steady_clock::now();
      t.push_back( std::chrono::duration< double >( t1 - t0 ).count() );
   }
   std::copy( yv.getData(), yv.getData() + n, y );
   std::sort( t.begin(), t.end() );
   return t.empty() ? 0.0 : t[ t.size() / 2 ];
}

// the SymmetricMatrix product (CSR storage of the lower triangle) timed as ref_spmv_csr_time
double ref_spmv_symmetric_time( int64_t n, const int32_t* row_ptr, const int32_t* cols, const double* vals,
                                const double* x, double* y, int reps, int threads )
{
   Host::setMaxThreadsCount( threads );
   SymMat< Seg::CSR > m;
   m.setDimensions( n, n );
   IVec caps( n );
   for( int64_t r = 0; r < n; r++ )
      caps[ r ] = row_ptr[ r + 1 ] - row_ptr[ r ];
   m.setRowCapacities( caps );
   const int64_t nnz = row_ptr[ n ];
   std::copy( cols, cols + nnz, m.getColumnIndexes().getData() );
   std::copy( vals, vals + nnz, m.getValues().getData() );
   DVec xv( n ), yv( n );
   std::copy( x, x + n, xv.getData() );
   m.vectorProduct( xv, yv );  // warm-up
   std::vector< double > t;
   for( int i = 0; i < reps; i++ ) {
      const auto t0 = std::chrono::steady_clock::now();
      m.vectorProduct( xv, yv );
      const auto t1 = std::chrono::steady_clock::now();
      t.push_back( std::chrono::duration< double >( t1 - t0 ).count() );
   }
   std::copy( yv.getData(), yv.getData() + n, y );
   std::sort( t.begin(), t.end() );
   return t.empty() ? 0.0 : t[ t.size() / 2 ];
}

void ref_cg( int64_t n, const int32_t* rp, const int32_t* c, const double* v, const double* b, double* x,
             int precond, State* st )
{
   auto A = make_csr( n, rp, c, v );
   Solvers::Linear::CG< CsrMat > s;
   run_solver( s, A, precond, b, x, n, st );
}

void ref_gmres( int64_t n, const int32_t* rp, const int32_t* c, const double* v, const double* b, double* x,
                int precond, int variant, int rmin, int rmax, int dmin, int dmax, State* st )
{
   auto A = make_csr( n, rp, c, v );
   Solvers::Linear::GMRES< CsrMat > s;
   Config::ConfigDescription cd;
   Solvers::Linear::GMRES< CsrMat >::configSetup( cd );
   Config::ParameterContainer pc;
   const char* names[] = { "CGS", "CGSR", "MGS", "MGSR", "CWY" };
   pc.addParameter< String >( "gmres-variant", names[ variant ] );
   pc.addParameter< int >( "gmres-restarting-min", rmin );
   pc.addParameter< int >( "gmres-restarting-max", rmax );
   pc.addParameter< int >( "gmres-restarting-step-min", dmin );
   pc.addParameter< int >( "gmres-restarting-step-max", dmax );
   s.setup( pc );
   run_solver( s, A, precond, b, x, n, st );
}

void ref_bicgstab( int64_t n, const int32_t* rp, const int32_t* c, const double* v, const double* b, double* x,
                   int precond, int exact_residue, State* st )
{
   auto A = make_csr( n, rp, c, v );
   Solvers::Linear::BICGStab< CsrMat > s;
   Config::ConfigDescription cd;
   Solvers::Linear::BICGStab< CsrMat >::configSetup( cd );
   Config::ParameterContainer pc;
   pc.addParameter< bool >( "bicgstab-exact-residue", exact_residue != 0 );
   s.setup( pc );
   run_solver( s, A, precond, b, x, n, st );
}

void ref_tfqmr( int64_t n, const int32_t* rp, const int32_t* c, const double* v, const double* b, double* x,
                int precond, State* st )
{
   auto A = make_csr( n, rp, c, v );
   Solvers::Linear::TFQMR< CsrMat > s;
   run_solver( s, A, precond, b, x, n, st );
}

void ref_bicgstabl( int64_t n, const int32_t* rp, const int32_t* c, const double* v, const double* b, double* x,
                    int precond, int ell, int exact_residue, State* st )
{
   auto A = make_csr( n, rp, c, v );
   Solvers::Linear::BICGStabL< CsrMat > s;
   Config::ConfigDescription cd;
   Solvers::Linear::BICGStabL< CsrMat >::configSetup( cd );
   Config::ParameterContainer pc;
   pc.addParameter< int >( "bicgstab-ell", ell );
   pc.addParameter< bool >( "bicgstab-exact-residue", exact_residue != 0 );
   s.setup( pc );
   run_solver( s, A, precond, b, x, n, st );
}

void ref_idrs( int64_t n, const int32_t* rp, const int32_t* c, const double* v, const double* b, double* x,
               int precond, int s_dim, int smoothing, State* st )
{
   auto A = make_csr( n, rp, c, v );
   Solvers::Linear::IDRs< CsrMat > s;
   s.setShadowSpaceDimension( s_dim );
   s.setResidualSmoothing( smoothing != 0 );
   run_solver( s, A, precond, b, x, n, st );
}

void ref_jacobi( int64_t n, const int32_t* rp, const int32_t* c, const double* v, const double* b, double* x,
                 double omega, int64_t period, State* st )
{
   auto A = make_csr( n, rp, c, v );
   Solvers::Linear::Jacobi< CsrMat > s;
   s.setOmega( omega );
   s.setResiduePeriod( period );
   run_solver( s, A, 0, b, x, n, st );
}

void ref_ilu0_apply( int64_t n, const int32_t* rp, const int32_t* c, const double* v, const double* b, double* x )
{
   auto A = make_csr( n, rp, c, v );
   Solvers::Linear::Preconditioners::ILU0< CsrMat > p;
   p.update( A );
   DVec bv( n ), xv( n );
   std::copy( b, b + n, bv.getData() );
   p.solve( bv.getConstView(), xv.getView() );
   std::copy( xv.getData(), xv.getData() + n, x );
}

// MatrixReader< CSR Host matrix >::readMtx; the result is kept for
// ref_read_mtx_arrays.  Returns 0, or -1 with the exception's message in
// ref_last_error().
static std::vector< int > g_off, g_col;
static std::vector< double > g_val;

int ref_read_mtx( const char* path, int64_t* rows, int64_t* cols, int64_t* nnz )
{
   try {
      CsrMat A;
      Matrices::MatrixReader< CsrMat >::readMtx( std::string( path ), A );
      const auto off = A.getSegments().getOffsets();
      g_off.assign( off.getData(), off.getData() + off.getSize() );
      const auto ci = A.getColumnIndexes();
      const auto va = A.getValues();
      g_col.assign( ci.getData(), ci.getData() + ci.getSize() );
      g_val.assign( va.getData(), va.getData() + va.getSize() );
      *rows = A.getRows();
      *cols = A.getColumns();
      *nnz = (int64_t) g_col.size();
      return 0;
   }
   catch( const std::exception& e ) {
      g_err = e.what();
      return -1;
   }
}

void ref_read_mtx_arrays( int32_t* off, int32_t* col, double* val )
{
   std::copy( g_off.begin(), g_off.end(), off );
   std::copy( g_col.begin(), g_col.end(), col );
   std::copy( g_val.begin(), g_val.end(), val );
}

const char* ref_last_error()
{
   return g_err.c_str();
}

// SparseMatrix::save / load (SparseMatrix.hpp:640-700) of a matrix of the given
// format (numbering of ref_fill); -1 with the message in ref_last_error() when
// TNL throws.  ref_load returns the storage size (arrays copied when <= cap).
int64_t ref_save( int format, int col_major, int64_t n, int64_t ncols, const int32_t* row_ptr, const int32_t* cols,
                  const double* vals, const char* path )
{
   try {
      return dispatch< SaveOp >( format, col_major, n, ncols, row_ptr, cols, vals, path );
   }
   catch( const std::exception& e ) {
      g_err = e.what();
      return -1;
   }
}

int64_t ref_load( int format, int col_major, const char* path, int64_t* dims, int32_t* out_cols, double* out_vals,
                  int64_t cap )
{
   try {
      return dispatch< LoadOp >( format, col_major, path, dims, out_cols, out_vals, cap );
   }
   catch( const std::exception& e ) {
      g_err = e.what();
      return -1;
   }
}

// SparseMatrixBase.hpp:298-329 / :541-588 through TNL's own classes
int64_t ref_spmv_symmetric( int format, int col_major, int64_t n, int64_t ncols, const int32_t* row_ptr,
                            const int32_t* cols, const double* vals, const double* x, double* y, double alpha,
                            double beta, int64_t begin, int64_t end )
{
   return dispatch< SymSpmvOp >( format, col_major, n, ncols, row_ptr, cols, vals, x, y, alpha, beta, begin, end );
}

int64_t ref_spmv_transposed( int format, int col_major, int64_t n, int64_t ncols, const int32_t* row_ptr,
                             const int32_t* cols, const double* vals, const double* x, double* y, double alpha,
                             double beta, int64_t begin, int64_t end )
{
   return dispatch< TransSpmvOp >( format, col_major, n, ncols, row_ptr, cols, vals, x, y, alpha, beta, begin, end );
}

// SparseMatrix::save of a SymmetricMatrix built from lower-triangle CSR arrays
int64_t ref_save_symmetric( int format, int col_major, int64_t n, int64_t ncols, const int32_t* row_ptr,
                            const int32_t* cols, const double* vals, const char* path )
{
   try {
      return dispatch< SymSaveOp >( format, col_major, n, ncols, row_ptr, cols, vals, path );
   }
   catch( const std::exception& e ) {
      g_err = e.what();
      return -1;
   }
}

// SparseMatrixBase::addElement / setElement (SparseMatrixBase.hpp:152-231) on a
// matrix of the given row capacities, general or symmetric
int64_t ref_add_elements( int format, int col_major, int symmetric, int64_t n, int64_t ncols, const int32_t* row_ptr,
                          const int32_t* cols, const double* vals, const int64_t* op_pos, const double* op_val,
                          int64_t nops, int64_t* failed, int32_t* out_cols, double* out_vals, int64_t cap )
{
   if( symmetric )
      return dispatch< SymElementsOp >( format, col_major, n, ncols, row_ptr, cols, vals, op_pos, op_val, nops, failed,
                                        out_cols, out_vals, cap );
   return dispatch< ElementsOp >( format, col_major, n, ncols, row_ptr, cols, vals, op_pos, op_val, nops, failed,
                                  out_cols, out_vals, cap );
}

int64_t ref_row_counts( int format, int col_major, int symmetric, int64_t n, int64_t ncols, const int32_t* row_ptr,
                        const int32_t* cols, const double* vals, int32_t* capacity, int32_t* stored, int32_t* nonzero )
{
   if( symmetric )
      return dispatch< SymCountsOp >( format, col_major, n, ncols, row_ptr, cols, vals, capacity, stored, nonzero );
   return dispatch< CountsOp >( format, col_major, n, ncols, row_ptr, cols, vals, capacity, stored, nonzero );
}

// OpenMP threads of TNL's Host algorithms in the following calls
void ref_set_threads( int threads )
{
   omp_set_num_threads( threads );
}

double ref_true_residual( int64_t n, const int32_t* rp, const int32_t* c, const double* v, const double* x,
                          const double* b )
{
   auto A = make_csr( n, rp, c, v );
   DVec bv( n ), xv( n );
   std::copy( b, b + n, bv.getData() );
   std::copy( x, x + n, xv.getData() );
   return Solvers::Linear::LinearResidueGetter::getResidue( *A, xv, bv, 0.0 );
}

}  // extern "C"
