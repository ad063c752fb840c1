// copy.hip -- copySparseMatrix on the device, between any two formats
// (Matrices/SparseOperations.hpp:600-690 with its kernels :544-597): the
// target's row capacities are the source rows' leading entries whose column
// is below the column count (padding slots, column -1, pass that test and are
// counted, as in the reference), then element k of every row is copied --
// column and value -- through both formats' getGlobalIndex maps.  One thread
// per row, grid-stride; the copy writes each target slot at most once.
#include <algorithm>

#include "segments.hpp"
#include "spmv.hpp"

namespace tnlmi {
namespace {

__global__ void copy_row_lengths( SegView v, const int32_t* __restrict__ cols, int64_t ncols, int32_t* __restrict__ len )
{
   for( int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < v.n; r += (int64_t) gridDim.x * blockDim.x ) {
      const int32_t size = segment_size( v, r );
      int32_t L = 0;
      for( int32_t k = 0; k < size; k++ ) {
         if( cols[ global_index( v, r, k ) ] < ncols )
            L++;
         else
            break;
      }
      len[ r ] = L;
   }
}

template< typename T >
__global__ void copy_clear( T* vals, int32_t* cols, int64_t storage )
{
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < storage; i += (int64_t) gridDim.x * blockDim.x ) {
      vals[ i ] = T{};
      cols[ i ] = -1;
   }
}

// (a row whose length exceeds the target's segment, or the source's, is not
// copied and sets *bad)
template< typename T >
__global__ void copy_rows( SegView a, T* __restrict__ avals, int32_t* __restrict__ acols, SegView b,
                           const T* __restrict__ bvals, const int32_t* __restrict__ bcols,
                           const int32_t* __restrict__ len, int32_t* bad )
{
   for( int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < b.n; r += (int64_t) gridDim.x * blockDim.x ) {
      const int32_t L = len[ r ];
      if( L < 0 || L > segment_size( a, r ) || L > segment_size( b, r ) ) {
         *bad = 1;
         continue;
      }
      for( int32_t k = 0; k < L; k++ ) {
         const int64_t gb = global_index( b, r, k );
         const int64_t ga = global_index( a, r, k );
         acols[ ga ] = bcols[ gb ];
         avals[ ga ] = bvals[ gb ];
      }
   }
}

dim3 rows_grid( int64_t n )
{
   return dim3( (unsigned) std::max< int64_t >( 1, std::min< int64_t >( 32 * 256, ceil_div( n, 256 ) ) ) );
}

template< typename T >
void copy_typed( const tnlmi_matrix& src, tnlmi_matrix& dst, const int32_t* len, hipStream_t s )
{
   const tnlmi_segments& ds = *dst.segments;
   if( ds.storage > 0 )
      hipLaunchKernelGGL( copy_clear< T >, dim3( (unsigned) std::min< int64_t >( 4096, ceil_div( ds.storage, 256 ) ) ),
                          dim3( 256 ), 0, s, (T*) dst.values, (int32_t*) dst.column_indexes, ds.storage );
   DeviceArray< int32_t > bad;
   bad.allocate( 1 );
   TNLMI_HIP( hipMemsetAsync( bad.ptr, 0, sizeof( int32_t ), s ) );
   if( src.rows > 0 )
      hipLaunchKernelGGL( copy_rows< T >, rows_grid( src.rows ), dim3( 256 ), 0, s, make_view( ds ), (T*) dst.values,
                          (int32_t*) dst.column_indexes, make_view( *src.segments ), (const T*) src.values,
                          src.column_indexes, len, bad.ptr );
   TNLMI_HIP( hipGetLastError() );
   int32_t h = 0;
   TNLMI_HIP( hipMemcpyAsync( &h, bad.ptr, sizeof( h ), hipMemcpyDeviceToHost, s ) );
   TNLMI_HIP( hipStreamSynchronize( s ) );
   if( h )
      fail( TNLMI_INVALID_ARGUMENT, "matrix_copy: a row length exceeds the target's (or the source's) segment size" );
}

}  // namespace
}  // namespace tnlmi

using namespace tnlmi;

extern "C" {

tnlmi_status tnlmi_matrix_row_lengths( const tnlmi_matrix* A, int32_t* d_lengths, tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( ! A || ! A->segments || ( A->rows > 0 && ! d_lengths ) )
         fail( TNLMI_INVALID_ARGUMENT, "matrix_row_lengths: null argument" );
      if( A->rows != A->segments->n )
         fail( TNLMI_INVALID_ARGUMENT, "matrix_row_lengths: matrix rows do not match the segments count" );
      hipStream_t s = as_stream( stream );
      if( A->rows > 0 )
         hipLaunchKernelGGL( copy_row_lengths, rows_grid( A->rows ), dim3( 256 ), 0, s, make_view( *A->segments ),
                             A->column_indexes, A->columns, d_lengths );
      TNLMI_HIP( hipGetLastError() );
      TNLMI_HIP( hipStreamSynchronize( s ) );
   } );
}

tnlmi_status tnlmi_matrix_copy( const tnlmi_matrix* src, tnlmi_matrix* dst, const int32_t* d_lengths,
                                tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( ! src || ! dst || ! src->segments || ! dst->segments )
         fail( TNLMI_INVALID_ARGUMENT, "matrix_copy: null argument" );
      if( src->value_type != dst->value_type )
         fail( TNLMI_INVALID_ARGUMENT, "matrix_copy: the matrices must have the same RealType" );
      if( src->matrix_type != dst->matrix_type )
         fail( TNLMI_NOT_IMPLEMENTED, "matrix_copy: copies between general and symmetric matrices are not supported" );
      if( src->rows != dst->rows || src->columns != dst->columns || src->rows != src->segments->n
          || dst->rows != dst->segments->n )
         fail( TNLMI_INVALID_ARGUMENT, "matrix_copy: the matrices' dimensions or segments differ" );
      if( src->rows > 0 && ! d_lengths )
         fail( TNLMI_INVALID_ARGUMENT, "matrix_copy: the row lengths are NULL" );
      if( dst->segments->storage > 0 && ( ! dst->values || ! dst->column_indexes ) )
         fail( TNLMI_INVALID_ARGUMENT, "matrix_copy: the target's storage arrays are NULL" );
      hipStream_t s = as_stream( stream );
      drop_auto_plans( *dst->segments, dst->values, dst->column_indexes );
      switch( src->value_type ) {
         case TNLMI_F64:
            copy_typed< double >( *src, *dst, d_lengths, s );
            break;
         case TNLMI_F32:
            copy_typed< float >( *src, *dst, d_lengths, s );
            break;
         case TNLMI_BF16:
            copy_typed< uint16_t >( *src, *dst, d_lengths, s );
            break;
      }
      TNLMI_HIP( hipStreamSynchronize( s ) );
   } );
}

}  // extern "C"
