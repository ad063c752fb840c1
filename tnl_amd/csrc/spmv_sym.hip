// spmv_sym.hip -- the two products that scatter into y: a symmetric matrix's
// vectorProduct (SparseMatrixBase.hpp:298-329; the storage holds the lower
// triangle and the diagonal) and transposedVectorProduct (:541-588).
//
// Both add into y with atomics, as the reference does on the GPU, so the
// order in which a y entry collects its terms is not fixed.  y is first
// scaled by beta over the whole vector (beta != 1), in a kernel of its own
// that precedes every add on the stream; beta == 0 clears y, as the general
// product ignores y then.
//
// Symmetric product: one thread per row, a workgroup per kSymRows consecutive
// rows.  A row's own sum is formed in registers; its mirrored terms
// v * x[r] go to y[c] for every stored c < r -- into an LDS accumulator when c
// is a row of the same workgroup (the near-diagonal terms of a banded matrix:
// no global atomic), else by a global atomic add.  After a barrier each row
// adds alpha * (its sum + its LDS accumulator) to y[r] with one global atomic
// (later workgroups may still be adding their mirrored terms to it).
#include <algorithm>

#include "segments.hpp"
#include "spmv.hpp"

namespace tnlmi {
namespace {

constexpr int kSymRows = 1024;  // rows per workgroup: a 3-D stencil's plane-1 neighbours (r - 256 on a 256^3 grid) stay in LDS for 3/4 of the rows

__device__ __forceinline__ void atomic_add( double* p, double v )
{
   unsafeAtomicAdd( p, v );  // global_atomic_add_f64 (coarse-grained device memory)
}
__device__ __forceinline__ void atomic_add( float* p, float v )
{
   unsafeAtomicAdd( p, v );
}

// y *= beta; beta == 0 writes zeros (the reference multiplies, which keeps a
// NaN of an uninitialised y -- a solver's fresh work vector -- alive)
template< typename X, typename C >
__global__ void scale_y( int64_t n, X* __restrict__ y, C beta )
{
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x )
      y[ i ] = beta == (C) 0 ? (X) 0 : (X) ( (C) y[ i ] * beta );
}

template< typename V, typename X, typename C >
__global__ void __launch_bounds__( kSymRows )
   spmv_symmetric( SegView v, const int32_t* __restrict__ cols, const V* __restrict__ vals, const X* __restrict__ x,
                   X* __restrict__ y, C alpha, int64_t begin, int64_t end )
{
   __shared__ C acc[ kSymRows ];
   const int t = threadIdx.x;
   const int64_t rb = begin + (int64_t) blockIdx.x * kSymRows, r = rb + t;
   acc[ t ] = 0;
   __syncthreads();
   C sum = 0;
   if( r < end ) {
      const C xr = (C) x[ r ];
      const int32_t size = segment_size( v, r );
      for( int32_t k = 0; k < size; k++ ) {
         const int64_t g = global_index( v, r, k );
         const int32_t c = cols[ g ];
         if( c < 0 )
            continue;  // padding slot
         const C val = load_val< V, C >( vals, g );
         if( c < r ) {
            const C mirrored = val * xr;
            if( c >= rb )
               atomicAdd( &acc[ c - rb ], mirrored );
            else
               atomic_add( y + c, (X) ( alpha * mirrored ) );
         }
         sum += val * (C) x[ c ];
      }
   }
   __syncthreads();
   if( r < end )
      atomic_add( y + r, (X) ( alpha * ( sum + acc[ t ] ) ) );
}

template< typename V, typename X, typename C >
__global__ void __launch_bounds__( 256 )
   spmv_transposed( SegView v, const int32_t* __restrict__ cols, const V* __restrict__ vals, const X* __restrict__ x,
                    X* __restrict__ y, C alpha, int64_t begin, int64_t end )
{
   for( int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < v.n; r += (int64_t) gridDim.x * blockDim.x ) {
      const C xr = (C) x[ r ];
      const int32_t size = segment_size( v, r );
      for( int32_t k = 0; k < size; k++ ) {
         const int64_t g = global_index( v, r, k );
         const int32_t c = cols[ g ];
         if( c >= begin && c < end )  // padding (-1) lies below every begin
            atomic_add( y + c, (X) ( alpha * xr * load_val< V, C >( vals, g ) ) );
      }
   }
}

template< typename X, typename C >
void scale( int64_t n, X* y, double beta, hipStream_t s )
{
   if( beta != 1.0 && n > 0 )
      hipLaunchKernelGGL( ( scale_y< X, C > ), dim3( (unsigned) std::min< int64_t >( 4096, ceil_div( n, 256 ) ) ),
                          dim3( 256 ), 0, s, n, y, (C) beta );
}

template< typename V, typename X >
void symmetric_typed( const tnlmi_matrix& A, const void* x, void* y, double alpha, double beta, int64_t begin,
                      int64_t end, hipStream_t s )
{
   using C = X;
   scale< X, C >( A.rows, (X*) y, beta, s );
   if( end > begin )
      hipLaunchKernelGGL( ( spmv_symmetric< V, X, C > ), dim3( (unsigned) ceil_div( end - begin, kSymRows ) ),
                          dim3( kSymRows ), 0, s, make_view( *A.segments ), A.column_indexes, (const V*) A.values,
                          (const X*) x, (X*) y, (C) alpha, begin, end );
   TNLMI_HIP( hipGetLastError() );
}

template< typename V, typename X >
void transposed_typed( const tnlmi_matrix& A, const void* x, void* y, double alpha, double beta, int64_t begin,
                       int64_t end, hipStream_t s )
{
   using C = X;
   scale< X, C >( A.columns, (X*) y, beta, s );
   if( A.rows > 0 && end > begin )
      hipLaunchKernelGGL( ( spmv_transposed< V, X, C > ),
                          dim3( (unsigned) std::min< int64_t >( 8192, ceil_div( A.rows, 256 ) ) ), dim3( 256 ), 0, s,
                          make_view( *A.segments ), A.column_indexes, (const V*) A.values, (const X*) x, (X*) y,
                          (C) alpha, begin, end );
   TNLMI_HIP( hipGetLastError() );
}

// the (value, vector) pairs of tnlmi_spmv, compute type = vector type
template< template< typename, typename > class F >
void by_type( const tnlmi_matrix& A, tnlmi_dtype xt, tnlmi_dtype ct, const char* who, const void* x, void* y,
              double alpha, double beta, int64_t begin, int64_t end, hipStream_t s )
{
   const tnlmi_dtype vt = A.value_type;
   if( ct != xt )
      fail( TNLMI_NOT_IMPLEMENTED, std::string( who ) + ": the compute type must be the vector type (atomic adds)" );
   if( vt == TNLMI_F64 && xt == TNLMI_F64 )
      F< double, double >::run( A, x, y, alpha, beta, begin, end, s );
   else if( vt == TNLMI_F32 && xt == TNLMI_F32 )
      F< float, float >::run( A, x, y, alpha, beta, begin, end, s );
   else if( vt == TNLMI_F32 && xt == TNLMI_F64 )
      F< float, double >::run( A, x, y, alpha, beta, begin, end, s );
   else if( vt == TNLMI_BF16 && xt == TNLMI_F32 )
      F< bf16, float >::run( A, x, y, alpha, beta, begin, end, s );
   else if( vt == TNLMI_BF16 && xt == TNLMI_F64 )
      F< bf16, double >::run( A, x, y, alpha, beta, begin, end, s );
   else
      fail( TNLMI_NOT_IMPLEMENTED, std::string( who ) + ": unsupported (value, vector) type combination" );
}

template< typename V, typename X >
struct SymmetricOp
{
   static void run( const tnlmi_matrix& A, const void* x, void* y, double alpha, double beta, int64_t begin,
                    int64_t end, hipStream_t s )
   {
      symmetric_typed< V, X >( A, x, y, alpha, beta, begin, end, s );
   }
};

template< typename V, typename X >
struct TransposedOp
{
   static void run( const tnlmi_matrix& A, const void* x, void* y, double alpha, double beta, int64_t begin,
                    int64_t end, hipStream_t s )
   {
      transposed_typed< V, X >( A, x, y, alpha, beta, begin, end, s );
   }
};

}  // namespace

void spmv_symmetric_dispatch( const tnlmi_matrix& A, tnlmi_dtype xt, const void* x, void* y, tnlmi_dtype ct,
                              double alpha, double beta, int64_t begin, int64_t end, hipStream_t s )
{
   if( A.rows != A.columns )
      fail( TNLMI_INVALID_ARGUMENT, "vectorProduct: a symmetric matrix must be square" );
   by_type< SymmetricOp >( A, xt, ct, "vectorProduct (symmetric)", x, y, alpha, beta, begin, end, s );
}

}  // namespace tnlmi

using namespace tnlmi;

extern "C" tnlmi_status tnlmi_spmv_transposed( const tnlmi_matrix* A, tnlmi_dtype x_type, const void* d_x, void* d_y,
                                               double alpha, double beta, int64_t begin, int64_t end,
                                               tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( ! A || ! A->segments )
         fail( TNLMI_INVALID_ARGUMENT, "transposedVectorProduct: matrix or its segments is NULL" );
      if( A->rows != A->segments->n )
         fail( TNLMI_INVALID_ARGUMENT, "transposedVectorProduct: matrix rows do not match the segments count" );
      if( ( A->rows && ! d_x ) || ( A->columns && ! d_y ) )
         fail( TNLMI_INVALID_ARGUMENT, "transposedVectorProduct: vector pointer is NULL" );
      if( A->segments->storage && ( ! A->values || ! A->column_indexes ) )
         fail( TNLMI_INVALID_ARGUMENT, "transposedVectorProduct: matrix storage pointer is NULL" );
      hipStream_t s = as_stream( stream );
      if( A->matrix_type == TNLMI_SYMMETRIC ) {  // SparseMatrixBase.hpp:559-562
         if( end == 0 )
            end = A->rows;
         if( begin < 0 || end > A->rows || begin > end )
            fail( TNLMI_INVALID_ARGUMENT, "vectorProduct: row range [begin, end) is out of bounds" );
         spmv_symmetric_dispatch( *A, x_type, d_x, d_y, x_type, alpha, beta, begin, end, s );
         return;
      }
      if( end == 0 )
         end = A->columns;  // SparseMatrixBase.hpp:573-574
      if( begin < 0 || end > A->columns || begin > end )
         fail( TNLMI_INVALID_ARGUMENT, "transposedVectorProduct: column range [begin, end) is out of bounds" );
      by_type< TransposedOp >( *A, x_type, x_type, "transposedVectorProduct", d_x, d_y, alpha, beta, begin, end, s );
   } );
}
