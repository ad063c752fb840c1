// blas1.hip -- deterministic device reductions and vector kernels (see blas1.hpp)
// plus their C ABI (TNL's Containers/Expressions BLAS-1 on the new device).
#include <algorithm>
#include <cmath>
#include <mutex>

#include "blas1.hpp"

namespace tnlmi {
namespace {

inline int red_grid( int64_t n )
{
   return (int) std::max< int64_t >( 1, std::min< int64_t >( kRedBlocks, ceil_div( n, 256 * 4 ) ) );
}

// first pass: one partial per workgroup (grid-stride, 4 independent sums per thread)
template< typename F >
__global__ void __launch_bounds__( 256 ) red_pass1( int64_t n, F f, double* part )
{
   __shared__ double sm[ 4 ];
   double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
   const int64_t stride = (int64_t) gridDim.x * 256;
   int64_t i = (int64_t) blockIdx.x * 256 + threadIdx.x;
   for( ; i + 3 * stride < n; i += 4 * stride ) {
      a0 += f( i );
      a1 += f( i + stride );
      a2 += f( i + 2 * stride );
      a3 += f( i + 3 * stride );
   }
   for( ; i < n; i += stride )
      a0 += f( i );
   const double v = block_sum< 256 >( ( a0 + a1 ) + ( a2 + a3 ), sm );
   if( threadIdx.x == 0 )
      part[ blockIdx.x ] = v;
}

// second pass: fixed-order sum of the partials; mode 0: sum, 1: sqrt(sum)
__global__ void __launch_bounds__( 256 ) red_pass2( const double* part, int np, double* out, int mode )
{
   __shared__ double sm[ 4 ];
   double a = 0;
   for( int i = threadIdx.x; i < np; i += 256 )
      a += part[ i ];
   a = block_sum< 256 >( a, sm );
   if( threadIdx.x == 0 )
      out[ 0 ] = mode == 1 ? sqrt( a ) : a;
}

struct DotF
{
   const double *a, *b;
   __device__ double operator()( int64_t i ) const { return a[ i ] * b[ i ]; }
};
struct SqF
{
   const double* a;
   __device__ double operator()( int64_t i ) const { return a[ i ] * a[ i ]; }
};

constexpr int kMaxCols = 32;
constexpr int kDotGroup = 8;  // columns of V per workgroup of the multi-dot's first pass

// A wave's 64 lanes hold the partial sums of kDotGroup columns each; three
// halving steps (4 + 2 + 1 exchanges) leave lane l with column l >> 3 summed
// over the 8 lanes that share bits 5..3 of l, and three more exchanges sum
// those 8 lanes.
__device__ __forceinline__ double wave_sum_group( double ( &acc )[ kDotGroup ] )
{
   static_assert( kDotGroup == 8, "three halving steps" );
   const int lane = threadIdx.x & 63;
#pragma unroll
   for( int o = 32, m = kDotGroup / 2; m >= 1; o /= 2, m /= 2 ) {
      const bool upper = lane & o;
#pragma unroll
      for( int j = 0; j < m; j++ ) {
         // keep one half of the columns, send the other to the partner lane
         const double keep = upper ? acc[ j + m ] : acc[ j ];
         const double send = upper ? acc[ j ] : acc[ j + m ];
         acc[ j ] = keep + __shfl_xor( send, o, 64 );
      }
   }
   double t = acc[ 0 ];
   t += __shfl_xor( t, 4, 64 );
   t += __shfl_xor( t, 2, 64 );
   t += __shfl_xor( t, 1, 64 );
   return t;
}

// part[k * kRedBlocks + b] = sum over the rows of workgroup column b of
// V[i, k] w[i], for the kDotGroup columns k of workgroup row blockIdx.y:
// eight accumulators per thread keep the loads of several rows in flight at
// full occupancy (w is read once per group of eight columns)
__global__ void __launch_bounds__( 256 ) multi_dot_pass1( int64_t n, int cols, const double* __restrict__ V,
                                                          int64_t ld, const double* __restrict__ w,
                                                          double* __restrict__ part )
{
   __shared__ double sm[ 4 ][ kDotGroup ];
   const int c0 = blockIdx.y * kDotGroup;
   const int nc = min( kDotGroup, cols - c0 );
   const double* __restrict__ Vc = V + (int64_t) c0 * ld;
   double acc[ kDotGroup ];
#pragma unroll
   for( int k = 0; k < kDotGroup; k++ )
      acc[ k ] = 0;
   const int64_t stride = (int64_t) gridDim.x * 256;
   // every column load is issued (a group's columns past cols read its last
   // column again and add nothing): loads under a condition would each wait
   // for the ones before them
   for( int64_t i = (int64_t) blockIdx.x * 256 + threadIdx.x; i < n; i += stride ) {
      const double wi = w[ i ];
      double v[ kDotGroup ];
#pragma unroll
      for( int k = 0; k < kDotGroup; k++ )
         v[ k ] = Vc[ i + (int64_t) min( k, nc - 1 ) * ld ];
#pragma unroll
      for( int k = 0; k < kDotGroup; k++ )
         acc[ k ] = fma( k < nc ? v[ k ] : 0.0, wi, acc[ k ] );  // (the contracted multiply-add)
   }
   // every column's wave total at once, then one barrier and a fixed-order
   // sum of the four waves' totals
   const double t = wave_sum_group( acc );
   const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
   if( ( lane & 7 ) == 0 )
      sm[ wv ][ lane >> 3 ] = t;
   __syncthreads();
   if( (int) threadIdx.x < nc ) {
      const int k = threadIdx.x;
      part[ (int64_t) ( c0 + k ) * kRedBlocks + blockIdx.x ] =
         ( ( sm[ 0 ][ k ] + sm[ 1 ][ k ] ) + sm[ 2 ][ k ] ) + sm[ 3 ][ k ];
   }
}

// out[k] = sum of the np partials of column k: one wavefront per column
// (lanes stride the partials, then a wave butterfly), four per workgroup
__global__ void __launch_bounds__( 256 ) multi_dot_pass2( const double* part, int np, int cols, double* out )
{
   const int lane = threadIdx.x & 63;
   const int k = blockIdx.x * 4 + ( threadIdx.x >> 6 );
   if( k >= cols )
      return;  // the whole wavefront
   double a = 0;
   for( int i = lane; i < np; i += 64 )
      a += part[ (int64_t) k * kRedBlocks + i ];
   a = group_sum< 64 >( a );
   if( lane == 0 )
      out[ k ] = a;
}

__global__ void axpby_k( int64_t n, double alpha, const double* x, double beta, double* y )
{
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x )
      y[ i ] = beta == 0.0 ? alpha * x[ i ] : alpha * x[ i ] + beta * y[ i ];
}

__global__ void scale_ratio_k( int64_t n, const double* x, const double* num, const double* den, double* y )
{
   const double d = den[ 0 ];
   const double sc = d != 0.0 ? num[ 0 ] / d : 0.0;
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x )
      y[ i ] = sc * x[ i ];
}

template< int SIGN >
__global__ void gemv_k( int64_t n, int cols, const double* __restrict__ V, int64_t ld, const double* __restrict__ h,
                        double* __restrict__ y )
{
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x ) {
      double acc = 0;
      for( int k = 0; k < cols; k++ )
         acc += V[ i + k * ld ] * h[ k ];
      y[ i ] = SIGN > 0 ? y[ i ] + acc : y[ i ] - acc;
   }
}

__global__ void divide_k( int64_t n, const double* x, const double* d, double* y )
{
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x )
      y[ i ] = x[ i ] / d[ i ];
}

__global__ void rsub_k( int64_t n, const double* x, double* y )
{
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x )
      y[ i ] = x[ i ] - y[ i ];
}

__global__ void copy_k( int64_t n, const double* x, double* y )
{
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x )
      y[ i ] = x[ i ];
}

__global__ void fill_k( int64_t n, double v, double* y )
{
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x )
      y[ i ] = v;
}

inline dim3 vgrid( int64_t n )
{
   return dim3( (unsigned) std::max< int64_t >( 1, std::min< int64_t >( 8192, ceil_div( n, 256 ) ) ) );
}

}  // namespace

void dot( int64_t n, const double* a, const double* b, double* out, RedWork& w, hipStream_t s, int slot )
{
   w.ensure( slot + 1 );
   const int g = red_grid( n );
   hipLaunchKernelGGL( red_pass1< DotF >, dim3( g ), dim3( 256 ), 0, s, n, DotF{ a, b }, w.slot( slot ) );
   hipLaunchKernelGGL( red_pass2, dim3( 1 ), dim3( 256 ), 0, s, w.slot( slot ), g, out, 0 );
}

void nrm2( int64_t n, const double* a, double* out, RedWork& w, hipStream_t s, int slot )
{
   w.ensure( slot + 1 );
   const int g = red_grid( n );
   hipLaunchKernelGGL( red_pass1< SqF >, dim3( g ), dim3( 256 ), 0, s, n, SqF{ a }, w.slot( slot ) );
   hipLaunchKernelGGL( red_pass2, dim3( 1 ), dim3( 256 ), 0, s, w.slot( slot ), g, out, 1 );
}

void multi_dot( int64_t n, int cols, const double* V, int64_t ld, const double* w, double* out, RedWork& rw,
                hipStream_t s )
{
   if( cols > kMaxCols )
      fail( TNLMI_NOT_IMPLEMENTED, "multi_dot: at most 32 columns" );
   if( cols <= 0 )
      return;
   rw.ensure( kMaxCols );
   const int g = red_grid( n );
   hipLaunchKernelGGL( multi_dot_pass1, dim3( g, (unsigned) ceil_div( cols, kDotGroup ) ), dim3( 256 ), 0, s, n,
                       cols, V, ld, w, rw.slot( 0 ) );
   hipLaunchKernelGGL( multi_dot_pass2, dim3( (unsigned) ceil_div( cols, 4 ) ), dim3( 256 ), 0, s, rw.slot( 0 ), g,
                       cols, out );
}

void axpby( int64_t n, double alpha, const double* x, double beta, double* y, hipStream_t s )
{
   hipLaunchKernelGGL( axpby_k, vgrid( n ), dim3( 256 ), 0, s, n, alpha, x, beta, y );
}

void scale_by_ratio( int64_t n, const double* x, const double* num, const double* den, double* y, hipStream_t s )
{
   hipLaunchKernelGGL( scale_ratio_k, vgrid( n ), dim3( 256 ), 0, s, n, x, num, den, y );
}

void gemv_sub( int64_t n, int cols, const double* V, int64_t ld, const double* h, double* y, hipStream_t s )
{
   hipLaunchKernelGGL( gemv_k< -1 >, vgrid( n ), dim3( 256 ), 0, s, n, cols, V, ld, h, y );
}

void gemv_add( int64_t n, int cols, const double* V, int64_t ld, const double* c, double* y, hipStream_t s )
{
   hipLaunchKernelGGL( gemv_k< 1 >, vgrid( n ), dim3( 256 ), 0, s, n, cols, V, ld, c, y );
}

void divide( int64_t n, const double* x, const double* d, double* y, hipStream_t s )
{
   hipLaunchKernelGGL( divide_k, vgrid( n ), dim3( 256 ), 0, s, n, x, d, y );
}

void rsub( int64_t n, const double* x, double* y, hipStream_t s )
{
   hipLaunchKernelGGL( rsub_k, vgrid( n ), dim3( 256 ), 0, s, n, x, y );
}

void copy( int64_t n, const double* x, double* y, hipStream_t s )
{
   hipLaunchKernelGGL( copy_k, vgrid( n ), dim3( 256 ), 0, s, n, x, y );
}

void fill( int64_t n, double v, double* y, hipStream_t s )
{
   hipLaunchKernelGGL( fill_k, vgrid( n ), dim3( 256 ), 0, s, n, v, y );
}

// ---------------------------------------------- allocation-free (a, b)
// The first pass's partial sums of tnlmi_dot_noalloc live in this file-scope
// device array: kRedBlocks FP64 (8 KiB), one copy per GPU, held until the
// process ends.  Host code never passes `scratch` itself to a kernel (in host
// code the name is a stand-in, not a device address): it passes the address
// hipGetSymbolAddress returns for the current device, looked up once per
// device and kept in scratch_dev.
static __device__ __align__( 256 ) double scratch[ kRedBlocks ];

namespace {

constexpr int kMaxDevices = 64;
double* scratch_dev[ kMaxDevices ] = {};
std::mutex scratch_mutex;

double* scratch_ptr()
{
   int dev = 0;
   TNLMI_HIP( hipGetDevice( &dev ) );
   if( dev < 0 || dev >= kMaxDevices )
      fail( TNLMI_RUNTIME_ERROR, "dot_noalloc: device ordinal out of range" );
   std::lock_guard< std::mutex > lock( scratch_mutex );
   if( ! scratch_dev[ dev ] ) {
      void* p = nullptr;
      TNLMI_HIP( hipGetSymbolAddress( &p, HIP_SYMBOL( scratch ) ) );
      if( ! p )
         fail( TNLMI_RUNTIME_ERROR, "dot_noalloc: no device address for the scratch array" );
      scratch_dev[ dev ] = (double*) p;
   }
   return scratch_dev[ dev ];
}

}  // namespace

}  // namespace tnlmi

using namespace tnlmi;

extern "C" {

// (a, b) and lpNorm(a, 2) of FP64 device vectors, result to device memory
// (asynchronous) -- Containers/Expressions/ExpressionTemplates.h:372-384, :442-460.
tnlmi_status tnlmi_dot( int64_t n, const double* d_a, const double* d_b, double* d_result, tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( n < 0 || ( n && ( ! d_a || ! d_b ) ) || ! d_result )
         fail( TNLMI_INVALID_ARGUMENT, "dot: invalid argument" );
      static thread_local RedWork w;
      dot( n, d_a, d_b, d_result, w, as_stream( stream ) );
      TNLMI_HIP( hipGetLastError() );
   } );
}

// tnlmi_dot without any device allocation: the same two kernels, the same
// fixed grid and order (bit-identical results), with the partial sums in the
// library's static device array instead of a workspace it allocates.  The
// first pass's grid is at most kRedBlocks workgroups whatever n is, so every
// n fits the array; a grid beyond it is refused all the same.  Calls that
// overlap on one GPU share the array and race.
tnlmi_status tnlmi_dot_noalloc( int64_t n, const double* d_a, const double* d_b, double* d_result,
                                tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( n < 0 || ( n && ( ! d_a || ! d_b ) ) || ! d_result )
         fail( TNLMI_INVALID_ARGUMENT, "dot_noalloc: invalid argument" );
      const int g = red_grid( n );
      if( g > kRedBlocks )
         fail( TNLMI_INVALID_ARGUMENT, "dot_noalloc: more partial sums than the static scratch array holds" );
      double* part = scratch_ptr();
      const hipStream_t s = as_stream( stream );
      // every partial that pass 2 reads (the first g) is written by pass 1:
      // nothing of an earlier call's contents is read
      hipLaunchKernelGGL( red_pass1< DotF >, dim3( g ), dim3( 256 ), 0, s, n, DotF{ d_a, d_b }, part );
      hipLaunchKernelGGL( red_pass2, dim3( 1 ), dim3( 256 ), 0, s, part, g, d_result, 0 );
      TNLMI_HIP( hipGetLastError() );
   } );
}

tnlmi_status tnlmi_nrm2( int64_t n, const double* d_a, double* d_result, tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( n < 0 || ( n && ! d_a ) || ! d_result )
         fail( TNLMI_INVALID_ARGUMENT, "nrm2: invalid argument" );
      static thread_local RedWork w;
      nrm2( n, d_a, d_result, w, as_stream( stream ) );
      TNLMI_HIP( hipGetLastError() );
   } );
}

// y = alpha*x + beta*y (VectorAssignment of x += a*p, p = z + b*p, ...)
tnlmi_status tnlmi_axpby( int64_t n, double alpha, const double* d_x, double beta, double* d_y, tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( n < 0 || ( n && ( ! d_x || ! d_y ) ) )
         fail( TNLMI_INVALID_ARGUMENT, "axpby: invalid argument" );
      if( n )
         axpby( n, alpha, d_x, beta, d_y, as_stream( stream ) );
      TNLMI_HIP( hipGetLastError() );
   } );
}

}  // extern "C"
