// krylov.hpp -- internal interfaces of the Krylov layer.
#pragma once

#include <cmath>

#include "common.hpp"

struct tnlmi_spmv_plan;
struct tnlmi_preconditioner;

namespace tnlmi {

// The matrix operator of a solver: y = A x with FP64 vectors and
// accumulation, through the matrix's own format kernels or its split plan.
struct Operator
{
   const tnlmi_matrix* A;
   const tnlmi_spmv_plan* plan;
   int64_t rows;
   void apply( const double* x, double* y, hipStream_t s ) const;
};

// x = M^{-1} b (identity when P is NULL)
void precond_apply( const tnlmi_preconditioner* P, int64_t n, const double* b, double* x, hipStream_t s );

// The preconditioner's kind and, for the Diagonal one, its diagonal d (x = b / d).
tnlmi_precond_kind precond_kind( const tnlmi_preconditioner* P );
const double* precond_diagonal( const tnlmi_preconditioner* P );

// ---------------------------------------------------------- solver state

// IterativeSolver.hpp:98-143
struct Stop
{
   tnlmi_solver_state* st;
   int64_t it = 0;
   double res = 0;
   bool check() const
   {
      return ! ( std::isnan( res ) || it > st->max_iterations
                 || ( res > st->divergence_residue && it >= st->min_iterations )
                 || ( res < st->convergence_residue && it >= st->min_iterations ) );
   }
   bool next()
   {
      const bool r = check();
      it++;
      return r;
   }
   bool converged() const
   {
      if( std::isnan( res ) )
         return false;
      if( res > st->divergence_residue && it > st->min_iterations )
         return false;
      if( it >= st->max_iterations )
         return false;
      if( res > st->convergence_residue )
         return false;
      return true;
   }
};

// pinned host mirror of a few device scalars
struct HostScalars
{
   double* h = nullptr;
   HostScalars() { TNLMI_HIP( hipHostMalloc( (void**) &h, 64 * sizeof( double ), hipHostMallocDefault ) ); }
   ~HostScalars()
   {
      if( h )
         (void) hipHostFree( h );
   }
   // synchronous read of k device scalars starting at d
   const double* read( const double* d, int k, hipStream_t s )
   {
      TNLMI_HIP( hipMemcpyAsync( h, d, k * sizeof( double ), hipMemcpyDeviceToHost, s ) );
      TNLMI_HIP( hipStreamSynchronize( s ) );
      return h;
   }
};

// CG with fused kernels for no or Diagonal preconditioning (fused_cg.hip).
// residue_kind 0: TNL's residue sqrt((r, z)) / ||M^-1 b||; 1: the recurrence
// residual ||r|| / ||b|| (the true-residual stopping rule of the mixed solvers).
// column_codes: 8-bit column codes for a column-major Ellpack / SlicedEllpack
// matrix of few distinct column offsets (as the mixed solver's inner matrix).
void fused_cg_solve( const Operator& A, const tnlmi_preconditioner* M, const double* b, double* x,
                     tnlmi_solver_state* st, hipStream_t s, int residue_kind = 0, bool column_codes = false );

}  // namespace tnlmi
