// blas1.hpp -- device vector operations of the Krylov layer: deterministic
// two-pass reductions into device scalars and fused update kernels.
// Replaces the expression-template BLAS-1 of Containers/Expressions
// ((a,b), lpNorm, x += a*p, p = z + b*p) and Algorithms/detail/Reduction.hpp,
// but keeps every scalar on the device (no host read-back per reduction).
#pragma once

#include "common.hpp"

namespace tnlmi {

constexpr int kRedBlocks = 1024;  // fixed grid of the first pass (deterministic order)

// Workspace of the reductions: kRedBlocks partials per slot.
struct RedWork
{
   DeviceArray< double > partials;  // slots x kRedBlocks
   int slots = 0;
   void ensure( int s )
   {
      if( s > slots ) {
         partials.allocate( (int64_t) s * kRedBlocks );
         slots = s;
      }
   }
   double* slot( int s ) { return partials.ptr + (int64_t) s * kRedBlocks; }
};

// out[0] = sum_i a[i]*b[i] (a, b double), written to device memory.
void dot( int64_t n, const double* a, const double* b, double* out, RedWork& w, hipStream_t s, int slot = 0 );
// out[0] = sqrt(sum a[i]^2)
void nrm2( int64_t n, const double* a, double* out, RedWork& w, hipStream_t s, int slot = 0 );
// out[k] = sum_i V[i + k*ld] * w[i] for k < cols (Reduction2D of GMRES CGS / CWY)
void multi_dot( int64_t n, int cols, const double* V, int64_t ld, const double* w, double* out, RedWork& rw,
                hipStream_t s );
// y = alpha*x + beta*y with alpha, beta host scalars
void axpby( int64_t n, double alpha, const double* x, double beta, double* y, hipStream_t s );
// y[i] = x[i] * scale where scale = num[0] / den[0] (device scalars; den == 0 -> 0) -- e.g. v = w / ||w||
void scale_by_ratio( int64_t n, const double* x, const double* num, const double* den, double* y, hipStream_t s );
// y -= sum_k V[:,k] * h[k] for k < cols (gemv with device coefficients h)
void gemv_sub( int64_t n, int cols, const double* V, int64_t ld, const double* h, double* y, hipStream_t s );
// y += sum_k V[:,k] * c[k]
void gemv_add( int64_t n, int cols, const double* V, int64_t ld, const double* c, double* y, hipStream_t s );
// y = x / d
void divide( int64_t n, const double* x, const double* d, double* y, hipStream_t s );
// y = x - y
void rsub( int64_t n, const double* x, double* y, hipStream_t s );
void copy( int64_t n, const double* x, double* y, hipStream_t s );
void fill( int64_t n, double v, double* y, hipStream_t s );

}  // namespace tnlmi
