// krylov_ns.hip -- the nonsymmetric Krylov solvers beside GMRES: BiCGStab
// (Solvers/Linear/BICGStab.hpp:33-127) and TFQMR (Solvers/Linear/TFQMR.hpp:
// 12-121), on the library's SpMV (matrix or plan), preconditioners and
// deterministic BLAS-1 reductions.
//
// Every scalar stays on the device; the host reads a few of them once per
// iteration to evaluate TNL's stopping rule (IterativeSolver.hpp:98-143).
// BiCGStab computes the next direction into a second buffer while it reads
// its restart test ((r, r*) too small against ||r*||^2): in the rare restart
// the direction is recomputed from the kept one after r* is reset, so the
// reference's order of operations holds with one host read per iteration.
// TFQMR reads its quasi-residual tau sqrt(it + 1) / ||M^-1 b|| once per
// iteration: its mid-iteration convergence break and the loop check both
// use it.
#include <cmath>
#include <limits>

#include "blas1.hpp"
#include "krylov.hpp"

namespace tnlmi {
namespace {

inline dim3 vg( int64_t n )
{
   return dim3( (unsigned) std::max< int64_t >( 1, std::min< int64_t >( 8192, ceil_div( n, 256 ) ) ) );
}

#define GRID_STRIDE( i, n )                                                                  \
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < ( n );              \
        i += (int64_t) gridDim.x * blockDim.x )

// dst = M^{-1} A src (A src without a preconditioner)
void pmatvec( const Operator& A, const tnlmi_preconditioner* M, const double* src, double* dst, double* tmp,
              hipStream_t s )
{
   if( M && precond_kind( M ) != TNLMI_PRECOND_NONE ) {
      A.apply( src, tmp, s );
      precond_apply( M, A.rows, tmp, dst, s );
   }
   else
      A.apply( src, dst, s );
}

// r = M^{-1} (b - A x)
void presidue( const Operator& A, const tnlmi_preconditioner* M, const double* x, const double* b, double* r,
               double* tmp, hipStream_t s )
{
   const int64_t n = A.rows;
   if( M && precond_kind( M ) != TNLMI_PRECOND_NONE ) {
      A.apply( x, tmp, s );
      rsub( n, b, tmp, s );
      precond_apply( M, n, tmp, r, s );
   }
   else {
      A.apply( x, r, s );
      rsub( n, b, r, s );
   }
}

void finish( tnlmi_solver_state* st, const Stop& S )
{
   st->iterations = S.it;
   st->residue = S.res;
   st->converged = S.converged();
}

// ------------------------------------------------------------------ BiCGStab
// scalars: 0 rho, 1 rho_old, 2 alpha, 3 omega, 4 beta, 5 b_norm, 6 r_ast_sqnorm,
// 7 (Ap, r*), 8 (As, s), 9 (As, As), 10 (r, r*) new, 11 ||r|| (or exact), 12 restart flag
enum { kRho, kRhoOld, kAlpha, kOmega, kBeta, kBnorm, kRast, kApR, kAsS, kAsAs, kRhoNew, kRnorm, kFlag, kBsSlots };

__global__ void bs_alpha( double* sc )
{
   sc[ kAlpha ] = sc[ kRho ] / sc[ kApR ];
}

__global__ void bs_s( int64_t n, const double* __restrict__ r, const double* __restrict__ Ap, double* __restrict__ s_,
                      const double* sc )
{
   const double a = sc[ kAlpha ];
   GRID_STRIDE( i, n )
   s_[ i ] = r[ i ] - a * Ap[ i ];
}

__global__ void bs_omega( double* sc )
{
   sc[ kOmega ] = sc[ kAsS ] / sc[ kAsAs ];
}

// x += alpha p + omega s;  r = s - omega As
__global__ void bs_xr( int64_t n, const double* __restrict__ p, const double* __restrict__ s_,
                       const double* __restrict__ As, double* __restrict__ x, double* __restrict__ r, const double* sc )
{
   const double a = sc[ kAlpha ], w = sc[ kOmega ];
   GRID_STRIDE( i, n )
   {
      x[ i ] += a * p[ i ] + w * s_[ i ];
      r[ i ] = s_[ i ] - w * As[ i ];
   }
}

// rho_old = rho; rho = (r, r*); restart test; beta
__global__ void bs_rho( double* sc, double eps2 )
{
   sc[ kRhoOld ] = sc[ kRho ];
   sc[ kRho ] = sc[ kRhoNew ];
   sc[ kFlag ] = fabs( sc[ kRho ] ) < eps2 * sc[ kRast ] ? 1.0 : 0.0;
   sc[ kBeta ] = ( sc[ kRho ] / sc[ kRhoOld ] ) * ( sc[ kAlpha ] / sc[ kOmega ] );
}

// after a restart: r* = r, rho = ||r*||^2 = (r, r) (in kRhoNew), beta again
__global__ void bs_restart( double* sc )
{
   sc[ kRho ] = sc[ kRhoNew ];
   sc[ kRast ] = sc[ kRhoNew ];
   sc[ kBeta ] = ( sc[ kRho ] / sc[ kRhoOld ] ) * ( sc[ kAlpha ] / sc[ kOmega ] );
}

// pn = r + beta p - (beta omega) Ap
__global__ void bs_dir( int64_t n, const double* __restrict__ r, const double* __restrict__ p,
                        const double* __restrict__ Ap, double* __restrict__ pn, const double* sc )
{
   const double b = sc[ kBeta ], bw = sc[ kBeta ] * sc[ kOmega ];
   GRID_STRIDE( i, n )
   pn[ i ] = r[ i ] + b * p[ i ] - bw * Ap[ i ];
}

__global__ void bs_init( double* sc )
{
   sc[ kRho ] = sc[ kRhoNew ];
   sc[ kRast ] = sc[ kRhoNew ];
}

void bicgstab_solve( const Operator& A, const tnlmi_preconditioner* M, bool exact, const double* b, double* x,
                     tnlmi_solver_state* st, hipStream_t s )
{
   const int64_t n = A.rows;
   DeviceArray< double > r, rast, p, pn, s_, Ap, As, tmp, sc;
   for( auto* a : { &r, &rast, &p, &pn, &s_, &Ap, &As, &tmp } )
      a->allocate( std::max< int64_t >( 1, n ) );
   sc.allocate( kBsSlots );
   RedWork w;
   HostScalars hs;
   Stop S{ st };
   const bool pre = M && precond_kind( M ) != TNLMI_PRECOND_NONE;
   // the norm of the preconditioned right-hand side (BICGStab.hpp:47-53)
   if( pre ) {
      precond_apply( M, n, b, tmp.ptr, s );
      nrm2( n, tmp.ptr, sc.ptr + kBnorm, w, s );
   }
   else
      nrm2( n, b, sc.ptr + kBnorm, w, s );
   const double b_norm = hs.read( sc.ptr + kBnorm, 1, s )[ 0 ];
   if( b_norm == 0 ) {
      fill( n, 0.0, x, s );
      st->iterations = 0;
      st->residue = 0.0;
      st->converged = 1;
      return;
   }
   presidue( A, M, x, b, r.ptr, tmp.ptr, s );
   copy( n, r.ptr, p.ptr, s );
   copy( n, r.ptr, rast.ptr, s );
   fill( n, 0.0, s_.ptr, s );
   dot( n, r.ptr, rast.ptr, sc.ptr + kRhoNew, w, s );
   hipLaunchKernelGGL( bs_init, dim3( 1 ), dim3( 1 ), 0, s, sc.ptr );
   const double eps2 = std::numeric_limits< double >::epsilon() * std::numeric_limits< double >::epsilon();
   S.res = std::sqrt( hs.read( sc.ptr + kRho, 1, s )[ 0 ] ) / b_norm;
   while( S.next() ) {
      pmatvec( A, M, p.ptr, Ap.ptr, tmp.ptr, s );
      dot( n, Ap.ptr, rast.ptr, sc.ptr + kApR, w, s );
      hipLaunchKernelGGL( bs_alpha, dim3( 1 ), dim3( 1 ), 0, s, sc.ptr );
      hipLaunchKernelGGL( bs_s, vg( n ), dim3( 256 ), 0, s, n, r.ptr, Ap.ptr, s_.ptr, sc.ptr );
      pmatvec( A, M, s_.ptr, As.ptr, tmp.ptr, s );
      dot( n, As.ptr, s_.ptr, sc.ptr + kAsS, w, s );
      dot( n, As.ptr, As.ptr, sc.ptr + kAsAs, w, s );
      hipLaunchKernelGGL( bs_omega, dim3( 1 ), dim3( 1 ), 0, s, sc.ptr );
      hipLaunchKernelGGL( bs_xr, vg( n ), dim3( 256 ), 0, s, n, p.ptr, s_.ptr, As.ptr, x, r.ptr, sc.ptr );
      dot( n, r.ptr, rast.ptr, sc.ptr + kRhoNew, w, s );
      hipLaunchKernelGGL( bs_rho, dim3( 1 ), dim3( 1 ), 0, s, sc.ptr, eps2 );
      hipLaunchKernelGGL( bs_dir, vg( n ), dim3( 256 ), 0, s, n, r.ptr, p.ptr, Ap.ptr, pn.ptr, sc.ptr );
      if( ! exact )
         nrm2( n, r.ptr, sc.ptr + kRnorm, w, s );
      const double* h = hs.read( sc.ptr + kRnorm, 2, s );  // ||r||, restart flag
      double rnorm = h[ 0 ];
      if( h[ 1 ] != 0.0 ) {
         // r has become too orthogonal to r*: restart with r* = r = M^{-1}(b - A x)
         presidue( A, M, x, b, r.ptr, tmp.ptr, s );
         copy( n, r.ptr, rast.ptr, s );
         dot( n, r.ptr, r.ptr, sc.ptr + kRhoNew, w, s );
         hipLaunchKernelGGL( bs_restart, dim3( 1 ), dim3( 1 ), 0, s, sc.ptr );
         hipLaunchKernelGGL( bs_dir, vg( n ), dim3( 256 ), 0, s, n, r.ptr, p.ptr, Ap.ptr, pn.ptr, sc.ptr );
         if( ! exact ) {
            nrm2( n, r.ptr, sc.ptr + kRnorm, w, s );
            rnorm = hs.read( sc.ptr + kRnorm, 1, s )[ 0 ];
         }
      }
      std::swap( p.ptr, pn.ptr );
      if( exact ) {
         // the exact preconditioned residue, computed into s
         presidue( A, M, x, b, s_.ptr, tmp.ptr, s );
         nrm2( n, s_.ptr, sc.ptr + kRnorm, w, s );
         rnorm = hs.read( sc.ptr + kRnorm, 1, s )[ 0 ];
      }
      S.res = rnorm / b_norm;
   }
   finish( st, S );
}

// ------------------------------------------------------------------- TFQMR
// scalars: 0 tau, 1 theta, 2 eta, 3 rho, 4 alpha, 5 b_norm, 6 ||w||, 7 (v, r*),
// 8 (w, r*), 9 beta, 10 residue, 11 d coefficient
enum { kTau, kTheta, kEta, kTRho, kTAlpha, kTBnorm, kWnorm, kVR, kWR, kTBeta, kTRes, kDc, kTfSlots };

__global__ void tf_alpha( double* sc )
{
   sc[ kTAlpha ] = sc[ kTRho ] / sc[ kVR ];
}

// w -= alpha Au;  d = u + (theta^2 eta / alpha) d
__global__ void tf_wd( int64_t n, const double* __restrict__ Au, const double* __restrict__ u, double* __restrict__ w,
                       double* __restrict__ d, const double* sc )
{
   const double a = sc[ kTAlpha ];
   const double c = sc[ kTheta ] * sc[ kTheta ] * sc[ kEta ] / a;
   GRID_STRIDE( i, n )
   {
      w[ i ] -= a * Au[ i ];
      d[ i ] = u[ i ] + c * d[ i ];
   }
}

// theta, c, tau, eta from ||w||; the quasi-residual of iteration it
__global__ void tf_scalars( double* sc, int64_t it )
{
   const double theta = sc[ kWnorm ] / sc[ kTau ];
   const double c = 1.0 / sqrt( 1.0 + theta * theta );
   sc[ kTheta ] = theta;
   sc[ kTau ] = sc[ kTau ] * theta * c;
   sc[ kEta ] = c * c * sc[ kTAlpha ];
   sc[ kTRes ] = sc[ kTau ] * sqrt( (double) ( it + 1 ) ) / sc[ kTBnorm ];
}

__global__ void tf_x( int64_t n, const double* __restrict__ d, double* __restrict__ x, const double* sc )
{
   const double e = sc[ kEta ];
   GRID_STRIDE( i, n )
   x[ i ] += e * d[ i ];
}

__global__ void tf_beta( double* sc )
{
   const double rho_new = sc[ kWR ];
   sc[ kTBeta ] = rho_new / sc[ kTRho ];
   sc[ kTRho ] = rho_new;
}

// u = w + beta u;  v = beta Au + beta^2 v
__global__ void tf_uv( int64_t n, const double* __restrict__ w, const double* __restrict__ Au, double* __restrict__ u,
                       double* __restrict__ v, const double* sc )
{
   const double b = sc[ kTBeta ];
   GRID_STRIDE( i, n )
   {
      u[ i ] = w[ i ] + b * u[ i ];
      v[ i ] = b * Au[ i ] + ( b * b ) * v[ i ];
   }
}

__global__ void tf_vadd( int64_t n, const double* __restrict__ Au, double* __restrict__ v )
{
   GRID_STRIDE( i, n )
   v[ i ] += Au[ i ];
}

__global__ void tf_usub( int64_t n, const double* __restrict__ v, double* __restrict__ u, const double* sc )
{
   const double a = sc[ kTAlpha ];
   GRID_STRIDE( i, n )
   u[ i ] -= a * v[ i ];
}

__global__ void tf_init( double* sc )
{
   sc[ kTheta ] = 0.0;
   sc[ kEta ] = 0.0;
   sc[ kTAlpha ] = 0.0;
   sc[ kTRho ] = sc[ kWR ];  // (r*, r) with r* = r
}

void tfqmr_solve( const Operator& A, const tnlmi_preconditioner* M, const double* b, double* x,
                  tnlmi_solver_state* st, hipStream_t s )
{
   const int64_t n = A.rows;
   DeviceArray< double > r, rast, w_, u, v, d, Au, tmp, sc;
   for( auto* a : { &r, &rast, &w_, &u, &v, &d, &Au, &tmp } )
      a->allocate( std::max< int64_t >( 1, n ) );
   sc.allocate( kTfSlots );
   RedWork rw;
   HostScalars hs;
   Stop S{ st };
   const bool pre = M && precond_kind( M ) != TNLMI_PRECOND_NONE;
   if( pre ) {
      precond_apply( M, n, b, tmp.ptr, s );
      nrm2( n, tmp.ptr, sc.ptr + kTBnorm, rw, s );
   }
   else
      nrm2( n, b, sc.ptr + kTBnorm, rw, s );
   presidue( A, M, x, b, r.ptr, tmp.ptr, s );
   const double b_norm = hs.read( sc.ptr + kTBnorm, 1, s )[ 0 ];
   if( b_norm == 0 ) {
      fill( n, 0.0, x, s );
      st->iterations = 0;
      st->residue = 0.0;
      st->converged = 1;
      return;
   }
   copy( n, r.ptr, w_.ptr, s );
   copy( n, r.ptr, u.ptr, s );
   pmatvec( A, M, u.ptr, Au.ptr, tmp.ptr, s );
   copy( n, Au.ptr, v.ptr, s );
   fill( n, 0.0, d.ptr, s );
   nrm2( n, r.ptr, sc.ptr + kTau, rw, s );
   copy( n, r.ptr, rast.ptr, s );
   dot( n, rast.ptr, r.ptr, sc.ptr + kWR, rw, s );
   hipLaunchKernelGGL( tf_init, dim3( 1 ), dim3( 1 ), 0, s, sc.ptr );
   S.res = hs.read( sc.ptr + kTau, 1, s )[ 0 ] / b_norm;
   while( S.next() ) {
      const int64_t it = S.it;
      if( it % 2 == 1 ) {
         dot( n, v.ptr, rast.ptr, sc.ptr + kVR, rw, s );
         hipLaunchKernelGGL( tf_alpha, dim3( 1 ), dim3( 1 ), 0, s, sc.ptr );
      }
      else
         pmatvec( A, M, u.ptr, Au.ptr, tmp.ptr, s );
      hipLaunchKernelGGL( tf_wd, vg( n ), dim3( 256 ), 0, s, n, Au.ptr, u.ptr, w_.ptr, d.ptr, sc.ptr );
      nrm2( n, w_.ptr, sc.ptr + kWnorm, rw, s );
      hipLaunchKernelGGL( tf_scalars, dim3( 1 ), dim3( 1 ), 0, s, sc.ptr, it );
      hipLaunchKernelGGL( tf_x, vg( n ), dim3( 256 ), 0, s, n, d.ptr, x, sc.ptr );
      S.res = hs.read( sc.ptr + kTRes, 1, s )[ 0 ];
      if( it > st->min_iterations && S.res < st->convergence_residue )
         break;
      if( it % 2 == 0 ) {
         dot( n, w_.ptr, rast.ptr, sc.ptr + kWR, rw, s );
         hipLaunchKernelGGL( tf_beta, dim3( 1 ), dim3( 1 ), 0, s, sc.ptr );
         hipLaunchKernelGGL( tf_uv, vg( n ), dim3( 256 ), 0, s, n, w_.ptr, Au.ptr, u.ptr, v.ptr, sc.ptr );
         pmatvec( A, M, u.ptr, Au.ptr, tmp.ptr, s );
         hipLaunchKernelGGL( tf_vadd, vg( n ), dim3( 256 ), 0, s, n, Au.ptr, v.ptr );
      }
      else
         hipLaunchKernelGGL( tf_usub, vg( n ), dim3( 256 ), 0, s, n, v.ptr, u.ptr, sc.ptr );
   }
   finish( st, S );
}

Operator make_operator( const tnlmi_matrix* A, const tnlmi_spmv_plan* plan )
{
   if( ! A || ! A->segments )
      fail( TNLMI_INVALID_ARGUMENT, "solver: matrix is NULL" );
   if( A->rows != A->columns )
      fail( TNLMI_INVALID_ARGUMENT, "solver: the matrix must be square" );
   return Operator{ A, plan, A->rows };
}

}  // namespace
}  // namespace tnlmi

using namespace tnlmi;

extern "C" {

tnlmi_status tnlmi_bicgstab_solve( const tnlmi_matrix* A, const tnlmi_spmv_plan* plan, const tnlmi_preconditioner* M,
                                   int32_t exact_residue, const double* d_b, double* d_x, tnlmi_solver_state* st,
                                   tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( ! st )
         fail( TNLMI_INVALID_ARGUMENT, "solver: state is NULL" );
      const Operator op = make_operator( A, plan );
      bicgstab_solve( op, M, exact_residue != 0, d_b, d_x, st, as_stream( stream ) );
      TNLMI_HIP( hipGetLastError() );
   } );
}

tnlmi_status tnlmi_tfqmr_solve( const tnlmi_matrix* A, const tnlmi_spmv_plan* plan, const tnlmi_preconditioner* M,
                                const double* d_b, double* d_x, tnlmi_solver_state* st, tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( ! st )
         fail( TNLMI_INVALID_ARGUMENT, "solver: state is NULL" );
      const Operator op = make_operator( A, plan );
      tfqmr_solve( op, M, d_b, d_x, st, as_stream( stream ) );
      TNLMI_HIP( hipGetLastError() );
   } );
}

}  // extern "C"
