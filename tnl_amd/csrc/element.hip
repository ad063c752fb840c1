// element.hip -- SparseMatrixBase::addElement / setElement / getElement on the
// device (Matrices/SparseMatrixBase.hpp:152-230 and :231-265).  One thread walks
// the row's getSegmentSize slots (an Ellpack row has the full width, however
// small its requested capacity) through the format's getGlobalIndex, as the
// reference's loop does with element-wise device reads and writes: a stored
// column is updated in place (value = thisElementMultiplicator * value + new value), else the new
// element goes into the first padding slot or, before the first larger column,
// with the slots after it shifted one place right (the last slot's content is
// dropped if the row is full, as in the reference).  A row with neither raises
// the reference's std::logic_error text.  The update rounds the product and the
// sum separately (no fused multiply-add), as the reference's host build does.
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "segments.hpp"
#include "spmv.hpp"

namespace tnlmi {
namespace {

__device__ __forceinline__ double update( double old, double mult, double value )
{
   return __dadd_rn( __dmul_rn( mult, old ), value );
}
__device__ __forceinline__ float update( float old, double mult, double value )
{
   return __fadd_rn( __fmul_rn( (float) mult, old ), (float) value );
}
__device__ __forceinline__ bf16 update( bf16 old, double mult, double value )
{
   bf16 r;
   r.bits = float_to_bf16( __fadd_rn( __fmul_rn( (float) mult, bf16_to_float( old.bits ) ), (float) value ) );
   return r;
}
template< typename V >
__device__ __forceinline__ V from_double( double v )
{
   return (V) v;
}
template<>
__device__ __forceinline__ bf16 from_double< bf16 >( double v )
{
   bf16 r;
   r.bits = float_to_bf16( (float) v );
   return r;
}
__device__ __forceinline__ double to_double( double v ) { return v; }
__device__ __forceinline__ double to_double( float v ) { return v; }
__device__ __forceinline__ double to_double( bf16 v ) { return bf16_to_float( v.bits ); }

// one addElement on row `row`: false if the row's capacity is exceeded
template< typename V >
__device__ bool add_one( const SegView& v, int32_t* __restrict__ cols, V* __restrict__ vals, int64_t row,
                         int32_t column, double value, double mult )
{
   const int32_t size = segment_size( v, row );  // SparseMatrixBase.hpp:184, 252
   int32_t col = -1, i = 0;
   int64_t g = 0;
   for( ; i < size; i++ ) {
      g = global_index( v, row, i );
      col = cols[ g ];
      if( col == column ) {
         vals[ g ] = update( vals[ g ], mult, value );
         return true;
      }
      if( col == -1 || col > column )
         break;
   }
   if( i == size )
      return false;
   if( col != -1 )
      for( int32_t j = size - 1; j > i; j-- ) {
         const int64_t g1 = global_index( v, row, j ), g2 = global_index( v, row, j - 1 );
         cols[ g1 ] = cols[ g2 ];
         vals[ g1 ] = vals[ g2 ];
      }
   cols[ g ] = column;
   vals[ g ] = from_double< V >( value );
   return true;
}

// status: 0 done, 1 the row's capacity is exceeded
template< typename V >
__global__ void add_element( SegView v, int32_t* __restrict__ cols, V* __restrict__ vals, int64_t row, int32_t column,
                             double value, double mult, int32_t* __restrict__ status )
{
   *status = add_one( v, cols, vals, row, column, value, mult ) ? 0 : 1;
}

// ---- a list of addElement / setElement operations in one launch.  The
// operations are sorted by row (stable: list order within a row); the thread
// of each row's first sorted operation applies that row's operations in list
// order.  Rows are independent, so the result equals the operations applied
// one by one in list order.  A row whose capacity is exceeded stops there;
// *failed receives the smallest list index of such an operation.

// key: the stored row (a symmetric matrix stores row >= column); bad[0] set
// for a position outside the matrix (or its stored triangle)
__global__ void op_keys( int64_t count, const int64_t* __restrict__ rows, const int64_t* __restrict__ columns,
                         int64_t nrows, int64_t ncols, int32_t symmetric, int32_t* __restrict__ key,
                         int32_t* __restrict__ ids, int32_t* __restrict__ bad )
{
   const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( i >= count )
      return;
   int64_t r = rows[ i ], c = columns[ i ];
   if( r < 0 || r >= nrows || c < 0 || c >= ncols ) {
      atomicExch( bad, 1 );
      r = 0;
   }
   else if( symmetric && r < c ) {  // SparseMatrixBase.hpp:178-182
      const int64_t t = r;
      r = c;
      c = t;
      if( r >= nrows || c >= ncols ) {
         atomicExch( bad, 1 );
         r = 0;
      }
   }
   key[ i ] = (int32_t) r;
   ids[ i ] = (int32_t) i;
}

template< typename V >
__global__ void apply_ops( int64_t count, const int32_t* __restrict__ skey, const int32_t* __restrict__ sid,
                           const int64_t* __restrict__ rows, const int64_t* __restrict__ columns,
                           const double* __restrict__ values, const double* __restrict__ mults, int32_t symmetric,
                           SegView v, int32_t* __restrict__ cols, V* __restrict__ vals,
                           unsigned long long* __restrict__ failed )
{
   const int64_t p = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( p >= count || ( p > 0 && skey[ p - 1 ] == skey[ p ] ) )
      return;  // not the first operation of its row
   const int32_t row = skey[ p ];
   for( int64_t q = p; q < count && skey[ q ] == row; q++ ) {
      const int32_t i = sid[ q ];
      int64_t c = columns[ i ];
      if( symmetric && rows[ i ] < c )
         c = rows[ i ];
      if( ! add_one( v, cols, vals, row, (int32_t) c, values[ i ], mults ? mults[ i ] : 0.0 ) ) {
         atomicMin( failed, (unsigned long long) i );
         return;
      }
   }
}

template< typename V >
__global__ void get_element( SegView v, const int32_t* __restrict__ cols, const V* __restrict__ vals, int64_t row,
                             int32_t column, double* __restrict__ out )
{
   const int32_t size = segment_size( v, row );  // SparseMatrixBase.hpp:184, 252
   for( int32_t i = 0; i < size; i++ ) {
      const int64_t g = global_index( v, row, i );
      if( cols[ g ] == column ) {
         *out = to_double( vals[ g ] );
         return;
      }
   }
   *out = 0.0;
}

__device__ __forceinline__ bool nonzero( double v ) { return v != 0.0; }
__device__ __forceinline__ bool nonzero( float v ) { return v != 0.0f; }
__device__ __forceinline__ bool nonzero( bf16 v ) { return ( v.bits & 0x7fffu ) != 0; }

// per row, over its getSegmentSize slots (SparseMatrixBase.hpp:55-132; padding
// slots count as the reduction's identity, :604-615): kind 0 the segment size
// (getRowCapacity), 1 the stored slots (getRowCapacities), 2 the stored
// nonzero values (getCompressedRowLengths), 3 getNonzeroElementsCount's term
// of a symmetric matrix (an off-diagonal slot inside the matrix counts twice)
template< typename V >
__global__ void row_counts( SegView v, const int32_t* __restrict__ cols, const V* __restrict__ vals, int64_t rows,
                            int64_t columns, int32_t kind, int32_t* __restrict__ out )
{
   for( int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += (int64_t) gridDim.x * blockDim.x ) {
      const int32_t size = segment_size( v, r );
      int32_t count = 0;
      if( kind == 0 )
         count = size;
      else
         for( int32_t k = 0; k < size; k++ ) {
            const int64_t g = global_index( v, r, k );
            const int32_t c = cols[ g ];
            if( c == -1 )
               continue;
            if( kind == 1 )
               count++;
            else if( kind == 2 )
               count += nonzero( vals[ g ] );
            else
               count += 1 + ( c != r && c < rows && r < columns );
         }
      out[ r ] = count;
   }
}

template< typename T >
T fetch_one( const T* d, hipStream_t s )
{
   T h;
   TNLMI_HIP( hipMemcpyAsync( &h, d, sizeof( T ), hipMemcpyDeviceToHost, s ) );
   TNLMI_HIP( hipStreamSynchronize( s ) );
   return h;
}

// checks the arguments and swaps (row, column) into the stored lower triangle
// of a symmetric matrix; false if the swapped position lies outside the matrix
bool check_element( const tnlmi_matrix* A, int64_t& row, int64_t& column, const char* who )
{
   if( ! A || ! A->segments )
      fail( TNLMI_INVALID_ARGUMENT, std::string( who ) + ": matrix or its segments is NULL" );
   if( A->rows != A->segments->n )
      fail( TNLMI_INVALID_ARGUMENT, std::string( who ) + ": matrix rows do not match the segments count" );
   if( row < 0 || row >= A->rows || column < 0 || column >= A->columns )
      fail( TNLMI_INVALID_ARGUMENT, std::string( who ) + ": element (" + std::to_string( row ) + ", "
                                        + std::to_string( column ) + ") is out of the matrix" );
   if( A->segments->storage > 0 && ( ! A->values || ! A->column_indexes ) )
      fail( TNLMI_INVALID_ARGUMENT, std::string( who ) + ": matrix storage pointer is NULL" );
   if( A->matrix_type == TNLMI_SYMMETRIC && row < column ) {  // SparseMatrixBase.hpp:178-182, 246-250
      std::swap( row, column );
      return row < A->rows && column < A->columns;
   }
   return true;
}

}  // namespace
}  // namespace tnlmi

using namespace tnlmi;

extern "C" {

tnlmi_status tnlmi_matrix_add_element( tnlmi_matrix* A, int64_t row, int64_t column, double value,
                                       double this_element_multiplicator, tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( ! check_element( A, row, column, "addElement" ) )
         fail( TNLMI_INVALID_ARGUMENT, "addElement: the transposed position is out of the symmetric matrix" );
      hipStream_t s = as_stream( stream );
      drop_auto_plans( *A->segments, A->values, A->column_indexes );
      DeviceArray< int32_t > status;
      status.allocate( 1 );
      const SegView v = make_view( *A->segments );
      int32_t* cols = const_cast< int32_t* >( A->column_indexes );
      switch( A->value_type ) {
         case TNLMI_F64:
            hipLaunchKernelGGL( add_element< double >, dim3( 1 ), dim3( 1 ), 0, s, v, cols, (double*) A->values, row,
                                (int32_t) column, value, this_element_multiplicator, status.ptr );
            break;
         case TNLMI_F32:
            hipLaunchKernelGGL( add_element< float >, dim3( 1 ), dim3( 1 ), 0, s, v, cols, (float*) A->values, row,
                                (int32_t) column, value, this_element_multiplicator, status.ptr );
            break;
         case TNLMI_BF16:
            hipLaunchKernelGGL( add_element< bf16 >, dim3( 1 ), dim3( 1 ), 0, s, v, cols, (bf16*) A->values, row,
                                (int32_t) column, value, this_element_multiplicator, status.ptr );
            break;
         default:
            fail( TNLMI_INVALID_ARGUMENT, "addElement: unknown value type" );
      }
      TNLMI_HIP( hipGetLastError() );
      if( fetch_one( status.ptr, s ) != 0 )  // SparseMatrixBase.hpp:198-206
         fail( TNLMI_LOGIC_ERROR,
               "The capacity of the sparse matrix row number " + std::to_string( row ) + " was exceeded." );
   } );
}

tnlmi_status tnlmi_matrix_add_elements( tnlmi_matrix* A, int64_t count, const int64_t* d_rows,
                                        const int64_t* d_columns, const double* d_values,
                                        const double* d_multiplicators, int64_t* failed, tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( failed )
         *failed = -1;
      if( ! A || ! A->segments )
         fail( TNLMI_INVALID_ARGUMENT, "addElements: matrix or its segments is NULL" );
      if( A->rows != A->segments->n )
         fail( TNLMI_INVALID_ARGUMENT, "addElements: matrix rows do not match the segments count" );
      if( count < 0 || count > INT32_MAX )
         fail( TNLMI_INVALID_ARGUMENT, "addElements: operation count out of range" );
      if( count == 0 )
         return;
      if( ! d_rows || ! d_columns || ! d_values )
         fail( TNLMI_INVALID_ARGUMENT, "addElements: operation array is NULL" );
      if( A->segments->storage > 0 && ( ! A->values || ! A->column_indexes ) )
         fail( TNLMI_INVALID_ARGUMENT, "addElements: matrix storage pointer is NULL" );
      hipStream_t s = as_stream( stream );
      drop_auto_plans( *A->segments, A->values, A->column_indexes );
      const int32_t sym = A->matrix_type == TNLMI_SYMMETRIC;
      DeviceArray< int32_t > key, skey, ids, sid, bad;
      DeviceArray< unsigned long long > first;
      for( auto* a : { &key, &skey, &ids, &sid } )
         a->allocate( count );
      bad.allocate( 1 );
      first.allocate( 1 );
      TNLMI_HIP( hipMemsetAsync( bad.ptr, 0, sizeof( int32_t ), s ) );
      TNLMI_HIP( hipMemsetAsync( first.ptr, 0xff, sizeof( unsigned long long ), s ) );
      const dim3 grid( (unsigned) ( ( count + 255 ) / 256 ) );
      hipLaunchKernelGGL( op_keys, grid, dim3( 256 ), 0, s, count, d_rows, d_columns, A->rows, A->columns, sym,
                          key.ptr, ids.ptr, bad.ptr );
      TNLMI_HIP( hipGetLastError() );
      if( fetch_one( bad.ptr, s ) != 0 )
         fail( TNLMI_INVALID_ARGUMENT, "addElements: an element lies outside the matrix" );
      int bits = 1;
      while( bits < 31 && ( int64_t( 1 ) << bits ) < A->rows )
         bits++;
      size_t bytes = 0;
      TNLMI_HIP( hipcub::DeviceRadixSort::SortPairs( nullptr, bytes, key.ptr, skey.ptr, ids.ptr, sid.ptr, (int) count,
                                                     0, bits, s ) );
      DeviceArray< char > tmp;
      tmp.allocate( (int64_t) bytes + 1 );
      TNLMI_HIP( hipcub::DeviceRadixSort::SortPairs( tmp.ptr, bytes, key.ptr, skey.ptr, ids.ptr, sid.ptr, (int) count,
                                                     0, bits, s ) );
      const SegView v = make_view( *A->segments );
      int32_t* cols = const_cast< int32_t* >( A->column_indexes );
      switch( A->value_type ) {
         case TNLMI_F64:
            hipLaunchKernelGGL( apply_ops< double >, grid, dim3( 256 ), 0, s, count, skey.ptr, sid.ptr, d_rows,
                                d_columns, d_values, d_multiplicators, sym, v, cols, (double*) A->values,
                                first.ptr );
            break;
         case TNLMI_F32:
            hipLaunchKernelGGL( apply_ops< float >, grid, dim3( 256 ), 0, s, count, skey.ptr, sid.ptr, d_rows,
                                d_columns, d_values, d_multiplicators, sym, v, cols, (float*) A->values, first.ptr );
            break;
         case TNLMI_BF16:
            hipLaunchKernelGGL( apply_ops< bf16 >, grid, dim3( 256 ), 0, s, count, skey.ptr, sid.ptr, d_rows,
                                d_columns, d_values, d_multiplicators, sym, v, cols, (bf16*) A->values, first.ptr );
            break;
         default:
            fail( TNLMI_INVALID_ARGUMENT, "addElements: unknown value type" );
      }
      TNLMI_HIP( hipGetLastError() );
      const unsigned long long f = fetch_one( first.ptr, s );
      if( f != ~0ull ) {
         if( failed )
            *failed = (int64_t) f;
         int64_t r = 0, c = 0;
         TNLMI_HIP( hipMemcpy( &r, d_rows + f, sizeof( r ), hipMemcpyDeviceToHost ) );
         TNLMI_HIP( hipMemcpy( &c, d_columns + f, sizeof( c ), hipMemcpyDeviceToHost ) );
         if( sym && r < c )
            r = c;
         fail( TNLMI_LOGIC_ERROR,
               "The capacity of the sparse matrix row number " + std::to_string( r ) + " was exceeded." );
      }
   } );
}

tnlmi_status tnlmi_matrix_row_counts( const tnlmi_matrix* A, int32_t kind, int32_t* d_counts, tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( ! A || ! A->segments || ( A->rows > 0 && ! d_counts ) )
         fail( TNLMI_INVALID_ARGUMENT, "matrix_row_counts: null argument" );
      if( A->rows != A->segments->n )
         fail( TNLMI_INVALID_ARGUMENT, "matrix_row_counts: matrix rows do not match the segments count" );
      if( kind < 0 || kind > 3 )
         fail( TNLMI_INVALID_ARGUMENT, "matrix_row_counts: unknown kind" );
      if( kind > 0 && A->segments->storage > 0 && ( ! A->values || ! A->column_indexes ) )
         fail( TNLMI_INVALID_ARGUMENT, "matrix_row_counts: matrix storage pointer is NULL" );
      hipStream_t s = as_stream( stream );
      if( A->rows > 0 ) {
         const SegView v = make_view( *A->segments );
         const dim3 grid( (unsigned) std::min< int64_t >( 4096, ( A->rows + 255 ) / 256 ) );
         switch( A->value_type ) {
            case TNLMI_F64:
               hipLaunchKernelGGL( row_counts< double >, grid, dim3( 256 ), 0, s, v, A->column_indexes,
                                   (const double*) A->values, A->rows, A->columns, kind, d_counts );
               break;
            case TNLMI_F32:
               hipLaunchKernelGGL( row_counts< float >, grid, dim3( 256 ), 0, s, v, A->column_indexes,
                                   (const float*) A->values, A->rows, A->columns, kind, d_counts );
               break;
            case TNLMI_BF16:
               hipLaunchKernelGGL( row_counts< bf16 >, grid, dim3( 256 ), 0, s, v, A->column_indexes,
                                   (const bf16*) A->values, A->rows, A->columns, kind, d_counts );
               break;
            default:
               fail( TNLMI_INVALID_ARGUMENT, "matrix_row_counts: unknown value type" );
         }
         TNLMI_HIP( hipGetLastError() );
      }
      TNLMI_HIP( hipStreamSynchronize( s ) );
   } );
}

tnlmi_status tnlmi_matrix_get_element( const tnlmi_matrix* A, int64_t row, int64_t column, double* value,
                                       tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( ! value )
         fail( TNLMI_INVALID_ARGUMENT, "getElement: value pointer is NULL" );
      if( ! check_element( A, row, column, "getElement" ) ) {
         *value = 0.0;
         return;
      }
      hipStream_t s = as_stream( stream );
      DeviceArray< double > out;
      out.allocate( 1 );
      const SegView v = make_view( *A->segments );
      switch( A->value_type ) {
         case TNLMI_F64:
            hipLaunchKernelGGL( get_element< double >, dim3( 1 ), dim3( 1 ), 0, s, v, A->column_indexes,
                                (const double*) A->values, row, (int32_t) column, out.ptr );
            break;
         case TNLMI_F32:
            hipLaunchKernelGGL( get_element< float >, dim3( 1 ), dim3( 1 ), 0, s, v, A->column_indexes,
                                (const float*) A->values, row, (int32_t) column, out.ptr );
            break;
         case TNLMI_BF16:
            hipLaunchKernelGGL( get_element< bf16 >, dim3( 1 ), dim3( 1 ), 0, s, v, A->column_indexes,
                                (const bf16*) A->values, row, (int32_t) column, out.ptr );
            break;
         default:
            fail( TNLMI_INVALID_ARGUMENT, "getElement: unknown value type" );
      }
      TNLMI_HIP( hipGetLastError() );
      *value = fetch_one( out.ptr, s );
   } );
}

}  // extern "C"
