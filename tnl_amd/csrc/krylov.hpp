// krylov.hpp -- internal interfaces of the Krylov layer.
#pragma once

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

}  // namespace tnlmi
