/*
 * oracle.h -- CPU restatement of TNL's sparse matrix-vector product, segment
 * converters and Krylov solvers.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (tnl_amd/) links, loads
 * or calls this library.  It is used by tests/, by __graft_entry__.smoke() and
 * by bench.py's cpu_baseline leg, always as the checker and never as the thing
 * that is measured or shipped.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the TNL source tree, src/TNL/...).  Pinning status of each function (which
 * golden vectors / closed forms / reference builds check it) is listed next to
 * its declaration.
 *
 * Conventions: int32 indices (the index type of every configuration), padding
 * column index -1 (Matrices/MatrixBase.h:27), all arrays caller-allocated.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- segments */

/* CSR offsets = exclusive scan of sizes (Algorithms/Segments/CSR.hpp:86-102).
 * offsets has n+1 entries.  Pinned: closed form + oracle/_ref. */
void orc_csr_offsets( const int32_t* sizes, int64_t n, int32_t* offsets );

/* Ellpack dimensions (Algorithms/Segments/Ellpack.hpp:81-97).  Returns 0, or 3
 * when alignedSize*segmentSize overflows int32 (the reference throws
 * std::overflow_error).  Pinned: closed form + oracle/_ref. */
int orc_ellpack_dims( const int32_t* sizes, int64_t n, int col_major, int alignment,
                      int32_t* segment_size, int32_t* aligned_size );

/* SlicedEllpack slice tables (Algorithms/Segments/SlicedEllpack.hpp:116-152).
 * slice_offsets: slices+1 entries, slice_seg_sizes: slices entries, where
 * slices = ceil(n / slice_size).  Pinned: closed form + oracle/_ref. */
void orc_sliced_ellpack( const int32_t* sizes, int64_t n, int slice_size,
                         int32_t* slice_offsets, int32_t* slice_seg_sizes );

/* BiEllpack permutation and group pointers (Algorithms/Segments/BiEllpack.hpp:
 * 104-211, initSegmentsPermutation :245-298, initGroupPointers :304-340).
 * perm: n entries; group_ptrs: strips*(log2(warp)+1)+1 entries.
 * Pinned: brute-force loop-for-loop restatement (orc_biellpack_bruteforce)
 * + oracle/_ref. */
void orc_biellpack( const int32_t* sizes, int64_t n, int warp, int32_t* perm, int32_t* group_ptrs );
void orc_biellpack_bruteforce( const int32_t* sizes, int64_t n, int warp, int32_t* perm, int32_t* group_ptrs );

/* ChunkedEllpack tables (Algorithms/Segments/ChunkedEllpack.hpp:156-216,
 * resolveSliceSizes :281-310, setSlice :313-380).
 * slices: 4 int32 per slice {size, chunkSize, firstSegment, pointer}, capacity n;
 * seg_to_chunk, seg_to_slice: n; seg_ptrs: n+1;
 * chunks_to_seg: capacity chunks_cap (must be >= slices*chunks_in_slice).
 * Returns the number of slices (or -1 when chunks_cap is too small); the
 * storage size goes to *storage.  Pinned: oracle/_ref (double-precision ratio
 * rounding). */
int64_t orc_chunked_ellpack( const int32_t* sizes, int64_t n, int chunks_in_slice, int desired_chunk_size,
                             int32_t* slices, int32_t* seg_to_chunk, int32_t* seg_to_slice, int32_t* seg_ptrs,
                             int32_t* chunks_to_seg, int64_t chunks_cap, int32_t* storage );

/* AdaptiveCSR block descriptors for one value size (Algorithms/Segments/
 * AdaptiveCSR.hpp:165-231, CSRAdaptiveKernelParameters.h).  Descriptor = 6
 * int32 {type, firstSegment, blockSize, segmentsInBlock, warpIdx, warpsCount};
 * type LONG=0, STREAM=1, VECTOR=2; fields the reference leaves uninitialised
 * are written as 0.  host_warp is the warp size seen by the compile-time
 * getWarpSize() on the host pass (32), device_warp the runtime device warp
 * size (64 on gfx950).  Returns the descriptor count, or -1 when cap is too
 * small.  Pinned: oracle/_ref at device_warp=32 (the only warp size a host
 * build sees) + hand-checked small cases. */
int64_t orc_adaptive_csr_blocks( const int32_t* offsets, int64_t n, int value_size, int host_warp,
                                 int device_warp, int32_t* blocks, int64_t cap );

/* SortedSegments (Algorithms/Segments/SortedSegments.hpp:129-200): sort by
 * (size desc, index asc) in windows of sigma (-1 = whole array).
 * Pinned: total-order comparator => unique result; + oracle/_ref. */
void orc_sorted_segments( const int32_t* sizes, int64_t n, int64_t sigma, int32_t* sorted_sizes,
                          int32_t* inverse_perm, int32_t* perm );

/* Global-index maps: for every (row r, local k < sizes[r]) in CSR order write
 * the storage position getGlobalIndex(r, k) into gidx[offsets[r]+k].
 * Ellpack: EllpackBase.hpp:108-117; SlicedEllpack: SlicedEllpackBase.hpp:119-146;
 * BiEllpack: detail/BiEllpack.h getGlobalIndex; ChunkedEllpack:
 * detail/ChunkedEllpack.h getGlobalIndex.  Pinned: oracle/_ref (the value and
 * column arrays that TNL's SparseMatrix builds for each format). */
void orc_gidx_ellpack( const int32_t* sizes, int64_t n, int32_t width, int32_t aligned, int col_major, int32_t* gidx );
void orc_gidx_sliced( const int32_t* sizes, int64_t n, int slice_size, const int32_t* slice_offsets,
                      const int32_t* slice_seg_sizes, int col_major, int32_t* gidx );
void orc_gidx_biellpack( const int32_t* sizes, int64_t n, int warp, const int32_t* perm, const int32_t* group_ptrs,
                         int col_major, int32_t* gidx );
void orc_gidx_chunked( const int32_t* sizes, int64_t n, int chunks_in_slice, const int32_t* slices,
                       const int32_t* seg_to_chunk, const int32_t* seg_to_slice, int col_major, int32_t* gidx );

/* -------------------------------------------------------------------- SpMV */

/* y = alpha*A*x + beta*y over rows [begin,end) of a matrix stored as
 * row_ptr/cols/vals (any format flattened by row: entries of row r are
 * vals[row_ptr[r] .. row_ptr[r+1])), padding column -1 contributes 0.
 * Matrices/SparseMatrixBase.hpp:331-389 (fetch + 4 keep variants),
 * Algorithms/Segments/detail/ReducingOperations_CSR.h:46-74 (sequential row
 * sum from 0, left to right).  end == 0 means all rows (:298).
 * Pinned: 8 KATs of SparseMatrixVectorProductTest.hpp, closed-form Laplacian,
 * dense brute force, oracle/_ref bit-exact. */
void orc_spmv_f64( int64_t rows, const int32_t* row_ptr, const int32_t* cols, const double* vals,
                   const double* x, double* y, double alpha, double beta, int64_t begin, int64_t end );
void orc_spmv_f32( int64_t rows, const int32_t* row_ptr, const int32_t* cols, const float* vals,
                   const float* x, float* y, float alpha, float beta, int64_t begin, int64_t end );
/* OpenMP version of orc_spmv_f64 with TNL's schedule(dynamic,100); the same
 * per-row arithmetic, so the result is bit-identical. */
void orc_spmv_f64_omp( int64_t rows, const int32_t* row_ptr, const int32_t* cols, const double* vals,
                       const double* x, double* y, int threads );
/* SparseMatrixBase.hpp:298-329: the product of a SymmetricMatrix whose CSR
 * arrays hold the lower triangle with the diagonal. */
void orc_spmv_symmetric( int64_t rows, const int32_t* row_ptr, const int32_t* cols, const double* vals,
                         const double* x, double* y, double alpha, double beta, int64_t begin, int64_t end );
/* SparseMatrixBase.hpp:541-588: transposedVectorProduct of a general matrix
 * (x has rows entries, y columns; [begin, end) ranges over columns). */
/* SparseMatrixBase::addElement / setElement (SparseMatrixBase.hpp:152-231) on
 * CSR-ordered slots (row r's slots [off[r], off[r+1]) in getGlobalIndex order),
 * ops (row, column) in op_pos and (value, multiplicator) in op_val; symmetric
 * stores (column, row) for row < column.  Returns the index of the first op
 * whose row has no room (the reference throws there), -1 if none. */
int64_t orc_add_elements( int64_t n, const int32_t* off, int32_t* cols, double* vals, int symmetric,
                          const int64_t* op_pos, const double* op_val, int64_t nops );
void orc_spmv_transposed( int64_t rows, int64_t columns, const int32_t* row_ptr, const int32_t* cols,
                          const double* vals, const double* x, double* y, double alpha, double beta, int64_t begin,
                          int64_t end );

/* ----------------------------------------------------------------- solvers */

/* Solver status block shared by the solver restatements. */
typedef struct {
   int64_t max_iterations;      /* IterativeSolver.h:226 default 1e9 */
   int64_t min_iterations;      /* default 0 */
   double convergence_residue;  /* default 1e-6 */
   double divergence_residue;   /* default DBL_MAX */
   int64_t iterations;          /* out */
   double residue;              /* out (solver's own residue) */
   int converged;               /* out: checkConvergence() */
} orc_solver_state;

/* preconditioner: 0 none, 1 Diagonal (Preconditioners/Diagonal.hpp:12-43),
 * 2 ILU0 (Preconditioners/ILU0.hpp:12-139 + Utils/TriangularSolve.h). */
/* CG: Solvers/Linear/CG.hpp:10-108 with IterativeSolver.hpp:98-143. */
void orc_cg( int64_t n, const int32_t* row_ptr, const int32_t* cols, const double* vals, const double* b,
             double* x, int precond, orc_solver_state* st );
/* GMRES: Solvers/Linear/GMRES.hpp:71-806.  variant: 0 CGS, 1 CGSR, 2 MGS,
 * 3 MGSR, 4 CWY. */
void orc_gmres( int64_t n, const int32_t* row_ptr, const int32_t* cols, const double* vals, const double* b,
                double* x, int precond, int variant, int restart_min, int restart_max, int step_min,
                int step_max, orc_solver_state* st );
/* BICGStab.hpp:33-127 (exact_residue: the bicgstab-exact-residue option) */
void orc_bicgstab( int64_t n, const int32_t* row_ptr, const int32_t* cols, const double* vals, const double* b,
                   double* x, int precond, int exact_residue, orc_solver_state* st );
/* TFQMR.hpp:12-121 */
void orc_tfqmr( int64_t n, const int32_t* row_ptr, const int32_t* cols, const double* vals, const double* b,
                double* x, int precond, orc_solver_state* st );
/* BICGStabL.hpp:37-257: BiCGStab(ell); an iteration per Bi-CG step. */
void orc_bicgstabl( int64_t n, const int32_t* row_ptr, const int32_t* cols, const double* vals, const double* b,
                    double* x, int precond, int ell, int exact_residue, orc_solver_state* st );
/* IDRs.hpp:47-386: IDR(s) with the reference's shadow space (mt19937 seed 0,
 * normal_distribution) and optional residual smoothing. */
void orc_idrs( int64_t n, const int32_t* row_ptr, const int32_t* cols, const double* vals, const double* b, double* x,
               int precond, int s, int smoothing, orc_solver_state* st );
/* Jacobi: Solvers/Linear/Jacobi.hpp:65-125. */
void orc_jacobi( int64_t n, const int32_t* row_ptr, const int32_t* cols, const double* vals, const double* b,
                 double* x, double omega, int64_t residue_period, orc_solver_state* st );
/* ILU(0) factors (Preconditioners/ILU0.hpp:12-96): L (strict lower, unit
 * diagonal implied) and U stored with reversed rows (U row N-1-i holds row i,
 * diagonal first).  Row pointers have n+1 entries; column/value arrays need
 * nnz(A) capacity each.  Pinned: tridiagonal => exact LU (Thomas algorithm);
 * + oracle/_ref. */
void orc_ilu0_factor( int64_t n, const int32_t* row_ptr, const int32_t* cols, const double* vals,
                      int32_t* l_ptr, int32_t* l_cols, double* l_vals, int32_t* u_ptr, int32_t* u_cols,
                      double* u_vals );
/* x = U^{-1} L^{-1} b (ILU0.hpp:99-113, Utils/TriangularSolve.h). */
void orc_ilu0_solve( int64_t n, const int32_t* l_ptr, const int32_t* l_cols, const double* l_vals,
                     const int32_t* u_ptr, const int32_t* u_cols, const double* u_vals, const double* b,
                     double* x );
/* ||A x - b||_2 / ||b||_2 (Solvers/Linear/Utils/LinearResidueGetter.hpp:11-23). */
double orc_true_residual( int64_t n, const int32_t* row_ptr, const int32_t* cols, const double* vals,
                          const double* x, const double* b );

#ifdef __cplusplus
}
#endif
