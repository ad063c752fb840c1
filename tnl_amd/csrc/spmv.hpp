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

// The plan tnlmi_spmv runs a product of A through under the automatic launch
// configuration, or NULL for A's own kernels: the windowed plan for a general
// CSR matrix whose x is wider than an L2 (2^20 columns or more, 2^22 stored
// elements or more) and whose columns mostly lie far from the diagonal
// (sampled rows), with values, vectors and compute type all FP64 or all FP32.
// Built once per (values, column indexes) array pair and kept in the
// segments object.
const tnlmi_spmv_plan* auto_plan( const tnlmi_matrix& A, tnlmi_dtype xt, tnlmi_dtype ct, hipStream_t s );

// Drops the automatic plans of the matrices on `seg` whose values or column
// indexes are the given arrays (all of them when both are NULL).
void drop_auto_plans( const tnlmi_segments& seg, const void* values, const int32_t* cols );

}  // namespace tnlmi
