// spmv.hpp -- internal entry of the SpMV dispatch (used by the solvers).
#pragma once

#include "common.hpp"

namespace tnlmi {

// A symmetric matrix's product (lower-triangle storage, atomic adds into y;
// asynchronous on s).  ct must equal xt.
void spmv_symmetric_dispatch( const tnlmi_matrix& A, tnlmi_dtype xt, const void* x, void* y, tnlmi_dtype ct,
                              double alpha, double beta, int64_t begin, int64_t end, hipStream_t s );

// y = alpha*A*x + beta*y over rows [begin, end) on stream s (asynchronous);
// a symmetric matrix goes to spmv_symmetric_dispatch.
void spmv_dispatch( const tnlmi_matrix& A, tnlmi_dtype xt, const void* x, void* y, tnlmi_dtype ct, double alpha,
                    double beta, int64_t begin, int64_t end, const tnlmi_launch& cfg, hipStream_t s );

}  // namespace tnlmi

struct tnlmi_spmv_plan;

namespace tnlmi {

// y = alpha*A*x + beta*y with the column-split plan (asynchronous on s).
void plan_execute( const tnlmi_spmv_plan& P, tnlmi_dtype xt, const void* x, void* y, tnlmi_dtype ct, double alpha,
                   double beta, int64_t begin, int64_t end, hipStream_t s );

}  // namespace tnlmi
