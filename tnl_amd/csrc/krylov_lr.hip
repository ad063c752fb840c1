// krylov_lr.hip -- BiCGStab(ell) (Solvers/Linear/BICGStabL.hpp:37-257) and
// IDR(s) (Solvers/Linear/IDRs.hpp:47-386) on the library's SpMV (matrix or
// plan), preconditioners and deterministic reductions.
//
// Both solvers keep their vectors on the device as columns of one array
// (R, U for BiCGStab(ell); P, G, U for IDR(s)) and their small dense algebra
// on the host, as the reference does: the host reads each inner product it
// needs (the Bi-CG scalars, the Gram-Schmidt coefficients, IDR(s)'s projected
// system) and passes the resulting coefficients to the update kernels by
// value.  Vector updates evaluate the reference's expressions element by
// element in its order (geam: 1 * A + beta * B; gemv: the column sum, then
// + y), so a GPU solve follows the host solve step for step.
//
// IDR(s)'s shadow space P is drawn on the host exactly as the reference
// draws it (std::mt19937 seeded 0, std::normal_distribution<double>(0, 1),
// then for each column the scalar (P_i, P_j) subtracted for j < i and the
// column scaled to unit norm) and uploaded once per solve.
#include <cmath>
#include <random>
#include <vector>

#include "blas1.hpp"
#include "krylov.hpp"

namespace tnlmi {
namespace {

constexpr int kMaxCols = 32;  // ell of BiCGStab(ell), s of IDR(s)

struct Coef
{
   double c[ kMaxCols ];
};

inline dim3 vg( int64_t n )
{
   return dim3( (unsigned) std::max< int64_t >( 1, std::min< int64_t >( 8192, ceil_div( n, 256 ) ) ) );
}

#define GRID_STRIDE( i, n )                                                                  \
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < ( n );              \
        i += (int64_t) gridDim.x * blockDim.x )

// out = a x + b y
__global__ void k_lin( int64_t n, double a, const double* x, double b, const double* y, double* out )
{
   GRID_STRIDE( i, n )
   out[ i ] = a * x[ i ] + b * y[ i ];
}

// y = y + a x
__global__ void k_axpy( int64_t n, double a, const double* __restrict__ x, double* __restrict__ y )
{
   GRID_STRIDE( i, n )
   y[ i ] = y[ i ] + a * x[ i ];
}

// y = y - a x (and z = z - a w, when w is given)
__global__ void k_axmy2( int64_t n, double a, const double* __restrict__ x, double* __restrict__ y,
                         const double* __restrict__ w, double* __restrict__ z )
{
   GRID_STRIDE( i, n )
   {
      y[ i ] = y[ i ] - a * x[ i ];
      if( w )
         z[ i ] = z[ i ] - a * w[ i ];
   }
}

// y = x / d
__global__ void k_div( int64_t n, const double* __restrict__ x, double d, double* __restrict__ y )
{
   GRID_STRIDE( i, n )
   y[ i ] = x[ i ] / d;
}

// C_c = 1 * A_c + beta * B_c for the columns c < cols (MatrixOperations::geam)
__global__ void k_geam( int64_t n, int cols, int64_t ld, const double* A, double beta, const double* B, double* C )
{
   GRID_STRIDE( i, n )
   for( int c = 0; c < cols; c++ )
      C[ c * ld + i ] = 1.0 * A[ c * ld + i ] + beta * B[ c * ld + i ];
}

// out = sum_c V_c * (alpha coef_c) + y (y == NULL: the sum alone) -- MatrixOperations::gemv;
// every entry of out is written after all its reads, so out may be one of the columns
__global__ void k_gemv( int64_t n, int cols, int64_t ld, const double* V, double alpha, Coef coef, const double* y,
                        double* out )
{
   GRID_STRIDE( i, n )
   {
      double acc = 0.0;
      for( int c = 0; c < cols; c++ )
         acc += V[ c * ld + i ] * ( alpha * coef.c[ c ] );
      out[ i ] = y ? acc + 1.0 * y[ i ] : acc;
   }
}

// residual smoothing: r_s = r_s - gamma t; x_s = x_s - gamma (x_s - x)
__global__ void k_smooth( int64_t n, double gamma, const double* __restrict__ t, const double* __restrict__ x,
                          double* __restrict__ r_s, double* __restrict__ x_s )
{
   GRID_STRIDE( i, n )
   {
      r_s[ i ] = r_s[ i ] - gamma * t[ i ];
      x_s[ i ] = x_s[ i ] - gamma * ( x_s[ i ] - x[ i ] );
   }
}

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

// device inner products read on the host
struct Dots
{
   int64_t n;
   hipStream_t s;
   DeviceArray< double > d;
   RedWork w;
   HostScalars hs;
   Dots( int64_t n_, hipStream_t s_ ) : n( n_ ), s( s_ ) { d.allocate( 2 * kMaxCols + 4 ); }
   double dot1( const double* a, const double* b )
   {
      dot( n, a, b, d.ptr, w, s );
      return hs.read( d.ptr, 1, s )[ 0 ];
   }
   void dot2( const double* a, const double* b, const double* c, const double* e, double& ab, double& ce )
   {
      dot( n, a, b, d.ptr, w, s, 0 );
      dot( n, c, e, d.ptr + 1, w, s, 1 );
      const double* h = hs.read( d.ptr, 2, s );
      ab = h[ 0 ];
      ce = h[ 1 ];
   }
   double norm( const double* a )
   {
      nrm2( n, a, d.ptr, w, s );
      return hs.read( d.ptr, 1, s )[ 0 ];
   }
   // out[k] = (V_k, x) for k < cols
   void multi( int cols, const double* V, int64_t ld, const double* x, double* out )
   {
      multi_dot( n, cols, V, ld, x, d.ptr, w, s );
      const double* h = hs.read( d.ptr, cols, s );
      for( int k = 0; k < cols; k++ )
         out[ k ] = h[ k ];
   }
};

void finish( tnlmi_solver_state* st, const Stop& S, bool converged )
{
   st->iterations = S.it;
   st->residue = S.res;
   st->converged = converged ? 1 : 0;
}

// ---------------------------------------------------------- BiCGStab(ell)
void bicgstabl_solve( const Operator& A, const tnlmi_preconditioner* M, int ell, bool exact, const double* b,
                      double* x, tnlmi_solver_state* st, hipStream_t s )
{
   const int64_t n = A.rows;
   const int64_t ld = std::max< int64_t >( 1, n );
   DeviceArray< double > R, U, r_ast, tmp, res;
   R.allocate( ( ell + 1 ) * ld );
   U.allocate( ( ell + 1 ) * ld );
   for( auto* a : { &r_ast, &tmp, &res } )
      a->allocate( ld );
   auto Rc = [ & ]( int j ) { return R.ptr + j * ld; };
   auto Uc = [ & ]( int j ) { return U.ptr + j * ld; };
   Dots D( n, s );
   Stop S{ st };
   std::vector< double > T( ell * ell ), sigma( ell + 1 ), g0( ell + 1 ), g1( ell + 1 ), g2( ell + 1 );
   const bool pre = M && precond_kind( M ) != TNLMI_PRECOND_NONE;
   double b_norm;
   if( pre ) {
      precond_apply( M, n, b, tmp.ptr, s );
      b_norm = D.norm( tmp.ptr );
   }
   else
      b_norm = D.norm( b );
   if( b_norm == 0 ) {
      fill( n, 0.0, x, s );
      finish( st, S, true );
      return;
   }
   presidue( A, M, x, b, Rc( 0 ), tmp.ptr, s );
   sigma[ 0 ] = D.norm( Rc( 0 ) );
   if( std::isnan( sigma[ 0 ] ) )
      fail( TNLMI_RUNTIME_ERROR, "BiCGstab(ell): initial residue is NAN" );
   hipLaunchKernelGGL( k_div, vg( n ), dim3( 256 ), 0, s, n, Rc( 0 ), sigma[ 0 ], r_ast.ptr );
   double rho0 = 1.0, alpha = 0.0, omega = 1.0;
   fill( n, 0.0, Uc( 0 ), s );
   S.res = sigma[ 0 ] / b_norm;
   while( S.check() ) {
      rho0 = -omega * rho0;
      // Bi-CG part
      for( int j = 0; j < ell; j++ ) {
         S.next();
         const double rho1 = D.dot1( r_ast.ptr, Rc( j ) );
         const double beta = alpha * rho1 / rho0;
         rho0 = rho1;
         hipLaunchKernelGGL( k_geam, vg( n ), dim3( 256 ), 0, s, n, j + 1, ld, R.ptr, -beta, U.ptr, U.ptr );
         pmatvec( A, M, Uc( j ), Uc( j + 1 ), tmp.ptr, s );
         const double gamma = D.dot1( r_ast.ptr, Uc( j + 1 ) );
         alpha = rho0 / gamma;
         hipLaunchKernelGGL( k_geam, vg( n ), dim3( 256 ), 0, s, n, j + 1, ld, R.ptr, -alpha, Uc( 1 ), R.ptr );
         pmatvec( A, M, Rc( j ), Rc( j + 1 ), tmp.ptr, s );
         hipLaunchKernelGGL( k_axpy, vg( n ), dim3( 256 ), 0, s, n, alpha, Uc( 0 ), x );
      }
      // MGS part
      for( int j = 1; j <= ell; j++ ) {
         for( int i = 1; i < j; i++ ) {
            const int ij = ( i - 1 ) + ( j - 1 ) * ell;
            T[ ij ] = D.dot1( Rc( i ), Rc( j ) ) / sigma[ i ];
            hipLaunchKernelGGL( k_axmy2, vg( n ), dim3( 256 ), 0, s, n, T[ ij ], Rc( i ), Rc( j ), nullptr,
                                nullptr );
         }
         double rr, r0r;
         D.dot2( Rc( j ), Rc( j ), Rc( 0 ), Rc( j ), rr, r0r );
         sigma[ j ] = rr;
         g1[ j ] = r0r / sigma[ j ];
      }
      omega = g1[ ell ];
      for( int j = ell; j >= 1; j-- ) {  // g0 = T^-1 g1
         g0[ j ] = g1[ j ];
         for( int i = j + 1; i <= ell; i++ )
            g0[ j ] -= T[ ( j - 1 ) + ( i - 1 ) * ell ] * g0[ i ];
      }
      for( int j = 1; j < ell; j++ ) {  // g2 = T S g0
         g2[ j ] = g0[ j + 1 ];
         for( int i = j + 1; i < ell; i++ )
            g2[ j ] += T[ ( j - 1 ) + ( i - 1 ) * ell ] * g0[ i + 1 ];
      }
      g2[ 0 ] = g0[ 1 ];
      Coef cx{}, cr{}, cu{};
      for( int c = 0; c < ell; c++ ) {
         cx.c[ c ] = g2[ c ];
         cr.c[ c ] = g1[ c + 1 ];
         cu.c[ c ] = g0[ c + 1 ];
      }
      hipLaunchKernelGGL( k_gemv, vg( n ), dim3( 256 ), 0, s, n, ell, ld, R.ptr, 1.0, cx, x, x );
      hipLaunchKernelGGL( k_gemv, vg( n ), dim3( 256 ), 0, s, n, ell, ld, Rc( 1 ), -1.0, cr, Rc( 0 ), Rc( 0 ) );
      hipLaunchKernelGGL( k_gemv, vg( n ), dim3( 256 ), 0, s, n, ell, ld, Uc( 1 ), -1.0, cu, Uc( 0 ), Uc( 0 ) );
      if( exact ) {
         presidue( A, M, x, b, res.ptr, tmp.ptr, s );
         sigma[ 0 ] = D.norm( res.ptr );
      }
      else
         sigma[ 0 ] = D.norm( Rc( 0 ) );
      S.res = sigma[ 0 ] / b_norm;
   }
   finish( st, S, S.converged() );
}

// ------------------------------------------------------------------ IDR(s)
void idrs_solve( const Operator& A, const tnlmi_preconditioner* M, int sdim, bool smoothing, const double* b,
                 double* x, tnlmi_solver_state* st, hipStream_t s )
{
   const int64_t n = A.rows;
   const int64_t ld = std::max< int64_t >( 1, n );
   DeviceArray< double > P, G, U, r, v, t, tmp, x_s, r_s;
   for( auto* a : { &P, &G, &U } )
      a->allocate( sdim * ld );
   for( auto* a : { &r, &v, &t, &tmp } )
      a->allocate( ld );
   if( smoothing ) {
      x_s.allocate( ld );
      r_s.allocate( ld );
   }
   auto col = [ & ]( DeviceArray< double >& V, int j ) { return V.ptr + j * ld; };
   Dots D( n, s );
   Stop S{ st };
   const bool pre = M && precond_kind( M ) != TNLMI_PRECOND_NONE;
   // the shadow space, as IDRs.hpp:354-385 draws and orthonormalises it
   {
      std::vector< double > h( sdim * n );
      std::mt19937 rng;
      rng.seed( 0 );
      std::normal_distribution< double > dist( 0, 1 );
      for( auto& p : h )
         p = dist( rng );
      auto hdot = [ & ]( int i, int j ) {
         double acc = 0;
         for( int64_t q = 0; q < n; q++ )
            acc += h[ i * n + q ] * h[ j * n + q ];
         return acc;
      };
      for( int i = 0; i < sdim; i++ ) {
         for( int j = 0; j < i; j++ ) {
            const double d = hdot( i, j );
            for( int64_t q = 0; q < n; q++ )
               h[ i * n + q ] = h[ i * n + q ] - d;
         }
         const double nr = std::sqrt( hdot( i, i ) );
         for( int64_t q = 0; q < n; q++ )
            h[ i * n + q ] /= nr;
      }
      if( n > 0 )
         TNLMI_HIP( hipMemcpyAsync( P.ptr, h.data(), sizeof( double ) * h.size(), hipMemcpyHostToDevice, s ) );
      TNLMI_HIP( hipStreamSynchronize( s ) );
   }
   auto psolve = [ & ]( const double* src, double* dst ) {  // dst may be src
      if( pre ) {
         precond_apply( M, n, src, tmp.ptr, s );
         copy( n, tmp.ptr, dst, s );
      }
      else if( src != dst )
         copy( n, src, dst, s );
   };
   const double b_norm = D.norm( b );
   if( b_norm == 0 ) {
      fill( n, 0.0, x, s );
      finish( st, S, true );
      return;
   }
   A.apply( x, r.ptr, s );
   rsub( n, b, r.ptr, s );
   double r_norm = D.norm( r.ptr );
   if( std::isnan( r_norm ) )
      fail( TNLMI_RUNTIME_ERROR, "IDR(s): initial residue is NAN" );
   const double angle = 0.7;
   double omega = 1.0;
   fill( sdim * ld, 0.0, G.ptr, s );
   fill( sdim * ld, 0.0, U.ptr, s );
   std::vector< double > Mm( sdim * sdim, 0.0 ), f( sdim ), aux( sdim );
   auto Mat = [ & ]( int i, int j ) -> double& { return Mm[ i * sdim + j ]; };
   for( int i = 0; i < sdim; i++ )
      Mat( i, i ) = 1;
   if( smoothing ) {
      copy( n, x, x_s.ptr, s );
      copy( n, r.ptr, r_s.ptr, s );
   }
   auto smooth = [ & ]() {
      hipLaunchKernelGGL( k_lin, vg( n ), dim3( 256 ), 0, s, n, 1.0, r_s.ptr, -1.0, r.ptr, t.ptr );
      double tr_s, tt;
      D.dot2( t.ptr, r_s.ptr, t.ptr, t.ptr, tr_s, tt );
      hipLaunchKernelGGL( k_smooth, vg( n ), dim3( 256 ), 0, s, n, tr_s / tt, t.ptr, x, r_s.ptr, x_s.ptr );
      r_norm = D.norm( r_s.ptr );
   };
   S.res = r_norm / b_norm;
   while( S.check() ) {
      D.multi( sdim, P.ptr, ld, r.ptr, f.data() );  // f = P' r
      for( int k = 0; k < sdim; k++ ) {
         // c = M[k:s, k:s]^-1 f[k:s] (LUsequential.h, no pivoting)
         const int m = sdim - k;
         std::vector< double > L( m * m );
         Coef c{};
         for( int i = 0; i < m; i++ )
            for( int j = 0; j < m; j++ )
               L[ i * m + j ] = Mat( i + k, j + k );
         for( int q = 0; q < m; q++ )
            for( int j = q + 1; j < m; j++ ) {
               const double factor = L[ j * m + q ] / L[ q * m + q ];
               L[ j * m + q ] = factor;
               for( int i = q + 1; i < m; i++ )
                  L[ j * m + i ] = L[ j * m + i ] - factor * L[ q * m + i ];
            }
         for( int i = 0; i < m; i++ )
            c.c[ i ] = f[ k + i ];
         for( int q = 1; q < m; q++ )
            for( int j = 0; j < q; j++ )
               c.c[ q ] -= L[ q * m + j ] * c.c[ j ];
         for( int q = m - 1; q >= 0; q-- ) {
            for( int j = m - 1; j > q; j-- )
               c.c[ q ] -= L[ q * m + j ] * c.c[ j ];
            c.c[ q ] /= L[ q * m + q ];
         }
         // v = r - G[:, k:s] c; preconditioning
         hipLaunchKernelGGL( k_gemv, vg( n ), dim3( 256 ), 0, s, n, m, ld, col( G, k ), -1.0, c, r.ptr, v.ptr );
         psolve( v.ptr, v.ptr );
         // U_k = U[:, k:s] c + omega v; G_k = A U_k
         double* Uk = col( U, k );
         double* Gk = col( G, k );
         hipLaunchKernelGGL( k_gemv, vg( n ), dim3( 256 ), 0, s, n, m, ld, Uk, 1.0, c, (const double*) nullptr, Uk );
         hipLaunchKernelGGL( k_axpy, vg( n ), dim3( 256 ), 0, s, n, omega, v.ptr, Uk );
         A.apply( Uk, Gk, s );
         // bi-orthogonalise against the earlier columns
         for( int i = 0; i < k; i++ ) {
            const double al = D.dot1( col( P, i ), Gk ) / Mat( i, i );
            hipLaunchKernelGGL( k_axmy2, vg( n ), dim3( 256 ), 0, s, n, al, col( U, i ), Uk, col( G, i ), Gk );
         }
         // M[k:s, k] = P[:, k:s]' G_k
         D.multi( m, col( P, k ), ld, Gk, aux.data() );
         for( int i = k; i < sdim; i++ )
            Mat( i, k ) = aux[ i - k ];
         if( Mat( k, k ) == 0 ) {  // breakdown
            finish( st, S, false );
            return;
         }
         const double beta = f[ k ] / Mat( k, k );
         hipLaunchKernelGGL( k_axpy, vg( n ), dim3( 256 ), 0, s, n, beta, Uk, x );
         hipLaunchKernelGGL( k_axmy2, vg( n ), dim3( 256 ), 0, s, n, beta, Gk, r.ptr, nullptr, nullptr );
         r_norm = D.norm( r.ptr );
         if( smoothing )
            smooth();
         S.res = r_norm / b_norm;
         if( ! S.next() ) {
            if( smoothing )
               copy( n, x_s.ptr, x, s );
            finish( st, S, S.converged() );
            return;
         }
         for( int i = k + 1; i < sdim; i++ )
            f[ i ] = f[ i ] - beta * Mat( i, k );
      }
      // the step into G_{j+1}
      psolve( r.ptr, v.ptr );
      A.apply( v.ptr, t.ptr, s );
      const double t_norm = D.norm( t.ptr );
      const double tr = D.dot1( t.ptr, r.ptr );
      const double rho = std::abs( tr / ( t_norm * r_norm ) );
      omega = tr / ( t_norm * t_norm );
      if( rho < angle )
         omega = omega * angle / rho;
      if( omega == 0 ) {  // breakdown
         finish( st, S, false );
         return;
      }
      hipLaunchKernelGGL( k_axpy, vg( n ), dim3( 256 ), 0, s, n, omega, v.ptr, x );
      hipLaunchKernelGGL( k_axmy2, vg( n ), dim3( 256 ), 0, s, n, omega, t.ptr, r.ptr, nullptr, nullptr );
      r_norm = D.norm( r.ptr );
      if( smoothing )
         smooth();
      S.res = r_norm / b_norm;
      S.next();
   }
   if( smoothing )
      copy( n, x_s.ptr, x, s );
   finish( st, S, S.converged() );
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

tnlmi_status tnlmi_bicgstabl_solve( const tnlmi_matrix* A, const tnlmi_spmv_plan* plan,
                                    const tnlmi_preconditioner* M, int32_t ell, int32_t exact_residue,
                                    const double* d_b, double* d_x, tnlmi_solver_state* st, tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( ! st )
         fail( TNLMI_INVALID_ARGUMENT, "solver: state is NULL" );
      if( ell < 1 || ell > kMaxCols )
         fail( TNLMI_NOT_IMPLEMENTED, "BiCGstab(ell): ell must lie in [1, 32]" );
      const Operator op = make_operator( A, plan );
      bicgstabl_solve( op, M, ell, exact_residue != 0, d_b, d_x, st, as_stream( stream ) );
      TNLMI_HIP( hipGetLastError() );
   } );
}

tnlmi_status tnlmi_idrs_solve( const tnlmi_matrix* A, const tnlmi_spmv_plan* plan, const tnlmi_preconditioner* M,
                               int32_t s, int32_t smoothing, const double* d_b, double* d_x, tnlmi_solver_state* st,
                               tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( ! st )
         fail( TNLMI_INVALID_ARGUMENT, "solver: state is NULL" );
      if( s < 1 || s > kMaxCols )
         fail( TNLMI_NOT_IMPLEMENTED, "IDR(s): s must lie in [1, 32]" );
      const Operator op = make_operator( A, plan );
      idrs_solve( op, M, s, smoothing != 0, d_b, d_x, st, as_stream( stream ) );
      TNLMI_HIP( hipGetLastError() );
   } );
}

}  // extern "C"
