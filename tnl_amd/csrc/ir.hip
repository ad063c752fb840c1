// ir.hip -- mixed-precision CG: FP64 iterative refinement around an inner CG
// whose matrix values are FP32 or BF16 and whose vectors and SpMV
// accumulation are FP32 (the inner dot products accumulate in FP64, which
// costs nothing and keeps the inner recurrences stable).
//
// Outer loop (FP64, on the GPU):  r = b - A x;  stop when ||r|| / ||b|| <=
// tolerance (the true residual of LinearResidueGetter.hpp:11-23);  solve
// A_lo d = r approximately (Jacobi-preconditioned CG in FP32, relative
// tolerance `inner_tolerance`, the right-hand side scaled to unit norm);
// x += d.
//
// No reference counterpart (TNL has no BF16 type and no mixed-precision
// solver); the inner iteration restates CG.hpp:10-108 in FP32.
#include <algorithm>
#include <cmath>

#include "blas1.hpp"
#include "krylov.hpp"
#include "segments.hpp"
#include "spmv.hpp"

using namespace tnlmi;

namespace tnlmi {
namespace {

inline dim3 vgrid( int64_t n )
{
   return dim3( (unsigned) std::max< int64_t >( 1, std::min< int64_t >( 8192, ceil_div( n, 256 ) ) ) );
}

// ---- FP32 vector kernels with FP64-accumulated reductions
template< typename F >
__global__ void __launch_bounds__( 256 ) red32( int64_t n, F f, double* part )
{
   __shared__ double sm[ 4 ];
   double a = 0;
   for( int64_t i = (int64_t) blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t) gridDim.x * 256 )
      a += f( i );
   a = block_sum< 256 >( a, sm );
   if( threadIdx.x == 0 )
      part[ blockIdx.x ] = a;
}
__global__ void __launch_bounds__( 256 ) red32_fin( const double* part, int np, double* out )
{
   __shared__ double sm[ 4 ];
   double a = 0;
   for( int i = threadIdx.x; i < np; i += 256 )
      a += part[ i ];
   a = block_sum< 256 >( a, sm );
   if( threadIdx.x == 0 )
      out[ 0 ] = a;
}
struct Dot32
{
   const float *a, *b;
   __device__ double operator()( int64_t i ) const { return (double) a[ i ] * (double) b[ i ]; }
};

void dot32( int64_t n, const float* a, const float* b, double* out, double* part, hipStream_t s )
{
   const int g = (int) std::max< int64_t >( 1, std::min< int64_t >( kRedBlocks, ceil_div( n, 1024 ) ) );
   hipLaunchKernelGGL( red32< Dot32 >, dim3( g ), dim3( 256 ), 0, s, n, Dot32{ a, b }, part );
   hipLaunchKernelGGL( red32_fin, dim3( 1 ), dim3( 256 ), 0, s, part, g, out );
}

// FP64 residual -> FP32 scaled right-hand side: rs = r / ||r||; inner x = 0
__global__ void ir_scale_rhs( int64_t n, const double* __restrict__ r, const double* __restrict__ sc,
                              float* __restrict__ rs, float* __restrict__ d )
{
   const double inv = sc[ 1 ] > 0.0 ? 1.0 / sc[ 1 ] : 0.0;
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x ) {
      rs[ i ] = (float) ( r[ i ] * inv );
      d[ i ] = 0.0f;
   }
}

// x += ||r|| * d  (FP64)
__global__ void ir_update( int64_t n, const float* __restrict__ d, const double* __restrict__ sc,
                           double* __restrict__ x )
{
   const double s = sc[ 1 ];
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x )
      x[ i ] += s * (double) d[ i ];
}

__global__ void residual64( int64_t n, const double* __restrict__ b, double* __restrict__ r )
{
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x )
      r[ i ] = b[ i ] - r[ i ];
}

// inner CG scalars ic[]: 0 s1, 1 s2, 2 alpha, 3 beta, 4 s1new, 5 res^2 (relative to the unit rhs)
__global__ void icg_alpha( double* ic )
{
   ic[ 2 ] = ic[ 1 ] != 0.0 ? ic[ 0 ] / ic[ 1 ] : 0.0;
}
__global__ void icg_update( int64_t n, const float* __restrict__ p, const float* __restrict__ Ap,
                            float* __restrict__ x, float* __restrict__ r, float* __restrict__ z,
                            const float* __restrict__ dinv, const double* __restrict__ ic )
{
   const float a = (float) ic[ 2 ];
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x ) {
      x[ i ] += a * p[ i ];
      const float ri = r[ i ] - a * Ap[ i ];
      r[ i ] = ri;
      z[ i ] = ri * dinv[ i ];
   }
}
__global__ void icg_beta( double* ic )
{
   const double s2 = ic[ 0 ], s1 = ic[ 4 ];
   ic[ 1 ] = s2;
   ic[ 0 ] = s1;
   ic[ 3 ] = s2 == 0.0 ? 0.0 : s1 / s2;
}
__global__ void icg_dir( int64_t n, const float* __restrict__ z, float* __restrict__ p, const double* __restrict__ ic )
{
   const float b = (float) ic[ 3 ];
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x )
      p[ i ] = z[ i ] + b * p[ i ];
}
__global__ void icg_init( int64_t n, const float* __restrict__ r, const float* __restrict__ dinv, float* __restrict__ z,
                          float* __restrict__ p )
{
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x ) {
      const float zi = r[ i ] * dinv[ i ];
      z[ i ] = zi;
      p[ i ] = zi;
   }
}

template< typename V >
__global__ void inv_diag32( SegView v, const int32_t* __restrict__ cols, const V* __restrict__ vals, float* dinv )
{
   const int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( r >= v.n )
      return;
   float d = 0.0f;
   const int32_t size = segment_size( v, r );
   for( int32_t k = 0; k < size; k++ ) {
      const int64_t g = global_index( v, r, k );
      if( cols[ g ] == r ) {
         d = load_val< V, float >( vals, g );
         break;
      }
   }
   dinv[ r ] = d != 0.0f ? 1.0f / d : 1.0f;
}

struct Pinned
{
   double* h = nullptr;
   Pinned() { TNLMI_HIP( hipHostMalloc( (void**) &h, 16 * sizeof( double ), hipHostMallocDefault ) ); }
   ~Pinned()
   {
      if( h )
         (void) hipHostFree( h );
   }
   const double* read( const double* d, int k, hipStream_t s )
   {
      TNLMI_HIP( hipMemcpyAsync( h, d, k * sizeof( double ), hipMemcpyDeviceToHost, s ) );
      TNLMI_HIP( hipStreamSynchronize( s ) );
      return h;
   }
};

}  // namespace
}  // namespace tnlmi

extern "C" tnlmi_status tnlmi_ir_cg_solve( const tnlmi_matrix* A, const tnlmi_matrix* A_lo,
                                           const tnlmi_spmv_plan* plan, const tnlmi_spmv_plan* plan_lo,
                                           double inner_tolerance, int64_t inner_max_iterations, const double* d_b,
                                           double* d_x, tnlmi_solver_state* st, int64_t* inner_iterations,
                                           tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( ! A || ! A_lo || ! st || ! A->segments || ! A_lo->segments )
         fail( TNLMI_INVALID_ARGUMENT, "ir_cg_solve: NULL argument" );
      if( A->rows != A->columns || A_lo->rows != A->rows || A_lo->columns != A->columns )
         fail( TNLMI_INVALID_ARGUMENT, "ir_cg_solve: matrix dimensions do not match" );
      if( A_lo->value_type == TNLMI_F64 )
         fail( TNLMI_INVALID_ARGUMENT, "ir_cg_solve: the inner matrix must hold FP32 or BF16 values" );
      hipStream_t s = as_stream( stream );
      const int64_t n = A->rows;
      const Operator op{ A, plan, n };
      DeviceArray< double > r, sc, part;
      DeviceArray< float > rs, d, ri, p, Ap, z, dinv;
      r.allocate( n );
      sc.allocate( 8 );
      part.allocate( kRedBlocks + 16 );
      for( auto* a : { &rs, &d, &ri, &p, &Ap, &z, &dinv } )
         a->allocate( n );
      double* ic = part.ptr + kRedBlocks;  // inner scalars
      RedWork w;
      Pinned hs;
      const tnlmi_segments& seg = *A_lo->segments;
      if( A_lo->value_type == TNLMI_F32 )
         hipLaunchKernelGGL( inv_diag32< float >, dim3( (unsigned) ceil_div( n, 256 ) ), dim3( 256 ), 0, s,
                             make_view( seg ), A_lo->column_indexes, (const float*) A_lo->values, dinv.ptr );
      else
         hipLaunchKernelGGL( inv_diag32< bf16 >, dim3( (unsigned) ceil_div( n, 256 ) ), dim3( 256 ), 0, s,
                             make_view( seg ), A_lo->column_indexes, (const bf16*) A_lo->values, dinv.ptr );
      nrm2( n, d_b, sc.ptr + 0, w, s );
      const double bnorm = hs.read( sc.ptr, 1, s )[ 0 ];
      int64_t outer = 0, inner_total = 0;
      double res = bnorm == 0.0 ? 0.0 : INFINITY;
      if( bnorm == 0.0 )
         fill( n, 0.0, d_x, s );
      const tnlmi_launch cfg{ TNLMI_MAP_AUTO, 1 };
      auto apply_lo = [ & ]( const float* xin, float* yout ) {
         if( plan_lo )
            plan_execute( *plan_lo, TNLMI_F32, xin, yout, TNLMI_F32, 1.0, 0.0, 0, n, s );
         else
            spmv_dispatch( *A_lo, TNLMI_F32, xin, yout, TNLMI_F32, 1.0, 0.0, 0, n, cfg, s );
      };
      while( bnorm != 0.0 ) {
         // r = b - A x in FP64 and the true residual
         op.apply( d_x, r.ptr, s );
         hipLaunchKernelGGL( residual64, vgrid( n ), dim3( 256 ), 0, s, n, d_b, r.ptr );
         nrm2( n, r.ptr, sc.ptr + 1, w, s );
         res = hs.read( sc.ptr + 1, 1, s )[ 0 ] / bnorm;
         if( std::isnan( res ) || res < st->convergence_residue || outer >= st->max_iterations
             || res > st->divergence_residue )
            break;
         outer++;
         // inner Jacobi-CG in FP32 on A_lo d = r / ||r||, from d = 0
         hipLaunchKernelGGL( ir_scale_rhs, vgrid( n ), dim3( 256 ), 0, s, n, r.ptr, sc.ptr, ri.ptr, d.ptr );
         hipLaunchKernelGGL( icg_init, vgrid( n ), dim3( 256 ), 0, s, n, ri.ptr, dinv.ptr, z.ptr, p.ptr );
         dot32( n, ri.ptr, z.ptr, ic + 0, part.ptr, s );
         const double s10 = hs.read( ic, 1, s )[ 0 ];
         for( int64_t k = 0; k < inner_max_iterations; k++ ) {
            apply_lo( p.ptr, Ap.ptr );
            dot32( n, Ap.ptr, p.ptr, ic + 1, part.ptr, s );
            hipLaunchKernelGGL( icg_alpha, dim3( 1 ), dim3( 1 ), 0, s, ic );
            hipLaunchKernelGGL( icg_update, vgrid( n ), dim3( 256 ), 0, s, n, p.ptr, Ap.ptr, d.ptr, ri.ptr, z.ptr,
                                dinv.ptr, ic );
            dot32( n, ri.ptr, z.ptr, ic + 4, part.ptr, s );
            hipLaunchKernelGGL( icg_beta, dim3( 1 ), dim3( 1 ), 0, s, ic );
            hipLaunchKernelGGL( icg_dir, vgrid( n ), dim3( 256 ), 0, s, n, z.ptr, p.ptr, ic );
            inner_total++;
            const double s1 = hs.read( ic, 1, s )[ 0 ];
            // preconditioned residual relative to the start of the inner solve
            if( !( s1 > inner_tolerance * inner_tolerance * s10 ) )
               break;
         }
         hipLaunchKernelGGL( ir_update, vgrid( n ), dim3( 256 ), 0, s, n, d.ptr, sc.ptr, d_x );
      }
      st->iterations = outer;
      st->residue = res;
      st->converged = ! std::isnan( res ) && res <= st->convergence_residue;
      if( inner_iterations )
         *inner_iterations = inner_total;
      TNLMI_HIP( hipGetLastError() );
   } );
}
