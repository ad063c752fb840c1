// fused_cg.hip -- Jacobi-preconditioned CG with the dot products fused into
// the kernels that produce their operands, and the reliable-update
// mixed-precision CG built from the same kernels.
//
// One CG iteration (CG.hpp:60-105) is three launches:
//   spmv_dot   Ap = A p  and the partial sums of (Ap, p)            (SpMV epilogue)
//   cg_upd     alpha = s1/s2;  x += alpha p;  r -= alpha Ap;  z = r / d (not stored);
//              partial sums of (r, z) [and (r, r)]
//   cg_dir     beta, the residue, TNL's stopping rule (IterativeSolver.hpp:98-121)
//              and the iteration counter;  p = r / d + beta p
// Every block of a kernel reduces the previous kernels' partial sums (one
// per block: 1024 in the FP64 CG, 2048 in the mixed solves' FP32 iterations)
// in the same fixed order, so the scalars are identical in every block and
// from run to run.  The iteration state lives on the device, double-buffered
// by iteration parity (iteration k reads state k & 1, block 0 of its last
// kernel writes state (k + 1) & 1); every kernel returns at once when the
// state says the solver has stopped, so the host enqueues iterations in
// batches and reads the state back once per batch.
//
// Mixed precision (tnlmi_mixed_cg_solve): the same iteration in FP32 on the
// FP32- or BF16-valued matrix, FP32 vectors and accumulation, FP64 dot
// products; x is accumulated in FP64.  Reliable updates (residual
// replacement, Clark et al., "Solving lattice QCD systems of equations using
// mixed precision solvers on GPUs", 2010): when the FP32 residual has dropped
// by `delta` since the last update (or its estimate reaches the tolerance),
// x += correction and r = b - A x are recomputed in FP64 and the search
// direction p is kept, so the Krylov process is not restarted.  The stopping
// test is the FP64 true residual ||b - A x|| / ||b||
// (LinearResidueGetter.hpp:11-23).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "blas1.hpp"
#include "krylov.hpp"
#include "segments.hpp"
#include "spmv.hpp"
#include "spmv_tile.hpp"

namespace tnlmi {
namespace {

// Blocks of a fused vector kernel = its number of partial sums: 1024 for the
// FP64 CG, 2048 for the FP32 iterations of the mixed solves (their FP32
// vectors move half the bytes per row, and twice the waves keep enough loads
// in flight: reliable BF16 CG on Poisson 256^3 0.281 against 0.320 s; the FP64
// CG measured 2 % slower with 2048, its blocks then summing twice the partials
// first).  kFB: the length of every partial-sum array.
constexpr int kFB = 2048;
template< typename T >
constexpr int blocks_of()
{
   return sizeof( T ) == 8 ? 1024 : 2048;
}
constexpr int kFB32 = blocks_of< float >();  // the mixed solve's kernels
constexpr int kFT = 256;
constexpr int kBatch = 8;  // iterations enqueued between two host reads of the state
static_assert( blocks_of< double >() % kXcds == 0 && blocks_of< float >() % kXcds == 0 && blocks_of< float >() <= kFB,
               "grids split evenly over the XCDs, partial sums fit their arrays" );

struct CgDev
{
   double s1, normb, res, beta;
   double rr, rr_max;  // mixed: (r, r) now and its maximum since the last update
   double s1_prev;     // mixed: (r, z) of the previous iteration (the beta denominator)
   double proj;        // mixed: (r, p) / (r, z) after a residual replacement
   int64_t it;         // TNL's currentIteration
   int32_t active;     // the iteration being run was admitted by nextIteration()
   int32_t s2zero;     // s2 == 0 ended the solve (CG.hpp:66-70)
   int32_t want_update;
   int32_t pad;
};

struct StopRule
{
   int64_t maxit, minit;
   double tol, div;
   int32_t kind = 0;  // residue: 0 TNL's sqrt((r, z)) / ||M^-1 b||, 1 ||r|| / ||b|| (the recurrence residual)
};

__device__ __forceinline__ bool check_next( const StopRule& R, double res, int64_t it )
{
   return ! ( isnan( res ) || it > R.maxit || ( res > R.div && it >= R.minit ) || ( res < R.tol && it >= R.minit ) );
}

// Sum of np partial sums in a fixed order; the same value in every block.
__device__ __forceinline__ double sum_parts( const double* __restrict__ part, double* sm, int np )
{
   double a = 0;
   for( int i = threadIdx.x; i < np; i += kFT )
      a += part[ i ];
   return block_sum< kFT >( a, sm );
}

// N partial-sum arrays at once: their loads in flight together and one pair
// of barriers, each total summed in the same order as sum_parts (so equal to
// it bit for bit).  sm holds N * kFT / 64 doubles.
template< int N >
__device__ __forceinline__ void sum_parts_n( const double* const ( &part )[ N ], double* sm, double ( &out )[ N ],
                                             int np )
{
   double a[ N ];
#pragma unroll
   for( int j = 0; j < N; j++ )
      a[ j ] = 0;
   for( int i = threadIdx.x; i < np; i += kFT )
#pragma unroll
      for( int j = 0; j < N; j++ )
         a[ j ] += part[ j ][ i ];
   const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
   for( int j = 0; j < N; j++ )
      a[ j ] = group_sum< 64 >( a[ j ] );
   __syncthreads();
   if( lane == 0 )
#pragma unroll
      for( int j = 0; j < N; j++ )
         sm[ j * ( kFT / 64 ) + w ] = a[ j ];
   __syncthreads();
#pragma unroll
   for( int j = 0; j < N; j++ ) {
      double r = 0;
#pragma unroll
      for( int i = 0; i < kFT / 64; i++ )
         r += sm[ j * ( kFT / 64 ) + i ];
      out[ j ] = wave_uniform( r );
   }
}

__device__ __forceinline__ void put_part( double v, double* sm, double* part )
{
   v = block_sum< kFT >( v, sm );
   if( threadIdx.x < 64 )
      part[ blockIdx.x ] = v;
}

// ---- fused SpMV + (Ap, p) partial sums (column-major Ellpack and
// SlicedEllpack; every other format runs its own SpMV kernel, then dot_part)
// Loops stride over the grid kU rows / elements at a time, so that every
// thread has kU independent loads in flight.
constexpr int kU = 4;

// Ellpack / SlicedEllpack, column-major: each thread R adjacent rows, whose
// slots are adjacent in every slot column (vector loads of R column indexes
// and R values), a full grid of row blocks placed XCD-contiguously (the rows
// that share x entries -- a stencil's neighbouring rows and planes -- are
// read by one XCD); one partial sum per block.  R divides the Ellpack row
// alignment (32) and the slice size, so a group never straddles a slice and
// its loads stay inside the padded storage (padding slots hold column -1).
template< typename E, int R >
struct Vec
{
   typedef E type __attribute__( ( ext_vector_type( R ) ) );
};
template< typename V, int R >
struct ValVec
{
   using type = typename Vec< V, R >::type;
   static __device__ __forceinline__ type load( const V* p ) { return __builtin_nontemporal_load( (const type*) p ); }
   template< typename T >
   static __device__ __forceinline__ T get( const type& v, int j )
   {
      return (T) v[ j ];
   }
};
template< int R >
struct ValVec< bf16, R >
{
   using type = typename Vec< uint16_t, R >::type;
   static __device__ __forceinline__ type load( const bf16* p )
   {
      return __builtin_nontemporal_load( (const type*) p );
   }
   template< typename T >
   static __device__ __forceinline__ T get( const type& v, int j )
   {
      return (T) bf16_to_float( v[ j ] );
   }
};

template< int R, typename V, typename T >
__global__ void __launch_bounds__( kFT ) spmv_dot_ell( SegView v, const int32_t* __restrict__ cols,
                                                       const V* __restrict__ vals, const T* __restrict__ p,
                                                       T* __restrict__ Ap, double* __restrict__ part,
                                                       const CgDev* __restrict__ st )
{
   using CV = typename Vec< int32_t, R >::type;
   using VV = ValVec< V, R >;
   __shared__ double sm[ kFT / 64 ];
   if( ! st->active )
      return;
   const int64_t blk = xcd_contiguous( blockIdx.x, gridDim.x );
   const int64_t r0 = ( blk * kFT + threadIdx.x ) * R;
   double acc = 0;
   if( r0 < v.n ) {
      int64_t g, stride;
      int32_t w;
      if( v.format == TNLMI_ELLPACK ) {
         g = r0;
         stride = v.ell_aligned;
         w = v.ell_width;
      }
      else {
         const int64_t sl = r0 / v.slice_size;
         g = v.slice_offsets[ sl ] + ( r0 - sl * v.slice_size );
         stride = v.slice_size;
         w = v.slice_seg_sizes[ sl ];
      }
      T y[ R ] = {};
      for( int32_t k = 0; k < w; k++, g += stride ) {
         const CV c = __builtin_nontemporal_load( (const CV*) ( cols + g ) );
         const typename VV::type a = VV::load( vals + g );
#pragma unroll
         for( int j = 0; j < R; j++ )
            if( c[ j ] >= 0 )
               y[ j ] += VV::template get< T >( a, j ) * p[ c[ j ] ];
      }
#pragma unroll
      for( int j = 0; j < R; j++ )
         if( r0 + j < v.n ) {
            Ap[ r0 + j ] = y[ j ];
            acc += (double) y[ j ] * (double) p[ r0 + j ];
         }
   }
   const double bs = block_sum< kFT >( acc, sm );
   if( threadIdx.x < 64 )
      part[ blk ] = bs;
}

// CSR with no row longer than kWholeRowMax (stencil matrices): the tiled CSR
// product in its one-launch form (each wavefront streams its TILE elements,
// every row is summed whole by its owner lane, left to right from 0) with the
// (Ap, p) partial sums in its epilogue: each lane adds v * p[r] for the rows
// it stores, one partial per workgroup.
template< typename V, typename T, int TILE >
__global__ void __launch_bounds__( kFT ) spmv_dot_csr( const int32_t* __restrict__ off,
                                                       const int32_t* __restrict__ tile_rows, int64_t tiles,
                                                       int64_t nnz, int64_t n, const int32_t* __restrict__ cols,
                                                       const V* __restrict__ vals, const T* __restrict__ p,
                                                       T* __restrict__ Ap, double* __restrict__ part,
                                                       const CgDev* __restrict__ st, int32_t tail_max )
{
   static_assert( kFT == 256, "four tile wavefronts per workgroup" );
   __shared__ T lds[ 4 ][ TILE + kWholeRowMax ];
   __shared__ double sm[ kFT / 64 ];
   if( ! st->active )
      return;
   const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
   const int64_t blk = xcd_contiguous( blockIdx.x, gridDim.x );
   const int64_t t = blk * 4 + wave;
   double acc = 0;
   if( t < tiles )
      csr_tile< V, T, T, true, TILE >(
         off, n, cols, vals, p, t, tile_rows[ t ], tile_rows[ t + 1 ], nnz, lds[ wave ], lane,
         [ & ]( int64_t r, T v ) {
            Ap[ r ] = v;
            acc += (double) v * (double) p[ r ];
         },
         []( int64_t, T ) {}, tail_max );
   const double bs = block_sum< kFT >( acc, sm );
   if( threadIdx.x < 64 )
      part[ blk ] = bs;
}

// ---- 8-bit column codes: for a matrix whose entries take few distinct
// offsets column - row (a stencil: 7 for the 3D 7-point Laplacian), every
// slot stores a code into a dictionary of at most kCodes offsets instead of
// its 32-bit column (code kPadCode: padding), so the SpMV reads 1 byte of
// index per entry instead of 4.  The columns decoded are exactly the stored
// ones: results are bit-identical to the int32 kernel above.
constexpr int kCodes = 255;
constexpr uint8_t kPadCode = 255;
constexpr int kHash = 4096;  // open-addressing table of the distinct offsets
constexpr uint64_t kEmpty = ~0ull;

template< int R, typename V, typename T >
__global__ void __launch_bounds__( kFT ) spmv_dot_ell_code( SegView v, const uint8_t* __restrict__ codes,
                                                            const int32_t* __restrict__ dict, int32_t ndict,
                                                            const V* __restrict__ vals, const T* __restrict__ p,
                                                            T* __restrict__ Ap, double* __restrict__ part,
                                                            const CgDev* __restrict__ st )
{
   using CB = typename Vec< uint8_t, R >::type;
   using VV = ValVec< V, R >;
   __shared__ double sm[ kFT / 64 ];
   __shared__ int32_t sd[ kCodes ];
   if( ! st->active )
      return;
   for( int i = threadIdx.x; i < ndict; i += kFT )
      sd[ i ] = dict[ i ];
   __syncthreads();
   const int64_t blk = xcd_contiguous( blockIdx.x, gridDim.x );
   const int64_t r0 = ( blk * kFT + threadIdx.x ) * R;
   double acc = 0;
   if( r0 < v.n ) {
      int64_t g, stride;
      int32_t w;
      if( v.format == TNLMI_ELLPACK ) {
         g = r0;
         stride = v.ell_aligned;
         w = v.ell_width;
      }
      else {
         const int64_t sl = r0 / v.slice_size;
         g = v.slice_offsets[ sl ] + ( r0 - sl * v.slice_size );
         stride = v.slice_size;
         w = v.slice_seg_sizes[ sl ];
      }
      // kKU entries at a time: their codes and values, then their reads of p
      // (one contiguous read when the R rows share a code, as in the interior
      // of a stencil: the columns r0 + j + offset are consecutive), then the
      // products, added in entry order as the int32 kernel adds them
      constexpr int kKU = 8;
      T y[ R ] = {};
      for( int32_t k0 = 0; k0 < w; k0 += kKU ) {
         CB c[ kKU ];
         typename VV::type a[ kKU ];
#pragma unroll
         for( int u = 0; u < kKU; u++ ) {
            if( k0 + u < w ) {
               c[ u ] = __builtin_nontemporal_load( (const CB*) ( codes + g + u * stride ) );
               a[ u ] = VV::load( vals + g + u * stride );
            }
            else
#pragma unroll
               for( int j = 0; j < R; j++ )
                  c[ u ][ j ] = kPadCode;
         }
         T pv[ kKU ][ R ];
#pragma unroll
         for( int u = 0; u < kKU; u++ ) {
            const uint32_t c0 = c[ u ][ 0 ];
            bool same = true;
#pragma unroll
            for( int j = 1; j < R; j++ )
               same = same && c[ u ][ j ] == c0;
            if( same && c0 != kPadCode )
               __builtin_memcpy( pv[ u ], p + r0 + sd[ c0 ], sizeof( pv[ u ] ) );
            else
#pragma unroll
               for( int j = 0; j < R; j++ )
                  pv[ u ][ j ] = c[ u ][ j ] != kPadCode ? p[ r0 + j + sd[ c[ u ][ j ] ] ] : (T) 0;
         }
#pragma unroll
         for( int u = 0; u < kKU; u++ )
#pragma unroll
            for( int j = 0; j < R; j++ )
               if( c[ u ][ j ] != kPadCode )
                  y[ j ] += VV::template get< T >( a[ u ], j ) * pv[ u ][ j ];
         g += kKU * stride;
      }
#pragma unroll
      for( int j = 0; j < R; j++ )
         if( r0 + j < v.n ) {
            Ap[ r0 + j ] = y[ j ];
            acc += (double) y[ j ] * (double) p[ r0 + j ];
         }
   }
   const double bs = block_sum< kFT >( acc, sm );
   if( threadIdx.x < 64 )
      part[ blk ] = bs;
}

__device__ __forceinline__ uint32_t offset_hash( uint32_t key )
{
   return ( key * 2654435761u ) >> 20;  // Knuth's multiplicative hash, 12 bits
}

// the slot of row r, entry k, and the row width (Ellpack / SlicedEllpack, column-major)
__device__ __forceinline__ int64_t ell_slot( const SegView& v, int64_t r, int32_t k, int32_t* w )
{
   if( v.format == TNLMI_ELLPACK ) {
      *w = v.ell_width;
      return r + (int64_t) k * v.ell_aligned;
   }
   const int64_t sl = r / v.slice_size;
   *w = v.slice_seg_sizes[ sl ];
   return v.slice_offsets[ sl ] + ( r - sl * v.slice_size ) + (int64_t) k * v.slice_size;
}

// the distinct offsets column - row into the table (a few recent ones kept in
// registers; a plain read before the atomic); more than kCodes, or a full
// probe sequence, sets *overflow
__global__ void __launch_bounds__( kFT ) code_collect( SegView v, const int32_t* __restrict__ cols,
                                                       unsigned long long* table, int32_t* count, int32_t* overflow )
{
   uint32_t recent[ 4 ] = { 0, 0, 0, 0 };
   bool have[ 4 ] = { false, false, false, false };
   int next = 0;
   for( int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < v.n; r += (int64_t) gridDim.x * blockDim.x ) {
      int32_t w = 0;
      ell_slot( v, r, 0, &w );
      for( int32_t k = 0; k < w; k++ ) {
         const int32_t c = cols[ ell_slot( v, r, k, &w ) ];
         if( c < 0 )
            continue;
         const uint32_t key = (uint32_t) ( c - (int32_t) r );
         bool seen = false;
#pragma unroll
         for( int i = 0; i < 4; i++ )
            seen |= have[ i ] && recent[ i ] == key;
         if( seen )
            continue;
         if( *(volatile int32_t*) overflow )
            return;
         uint32_t h = offset_hash( key );
         bool done = false;
         for( int probe = 0; probe < kHash && ! done; probe++, h = ( h + 1 ) & ( kHash - 1 ) ) {
            const unsigned long long cur = ( (volatile unsigned long long*) table )[ h ];
            if( cur == (unsigned long long) key ) {
               done = true;
               break;
            }
            if( cur != kEmpty )
               continue;
            const unsigned long long prev = atomicCAS( table + h, kEmpty, (unsigned long long) key );
            if( prev == kEmpty ) {
               if( atomicAdd( count, 1 ) >= kCodes )
                  atomicExch( overflow, 1 );
               done = true;
            }
            else if( prev == (unsigned long long) key )
               done = true;
         }
         if( ! done ) {
            atomicExch( overflow, 1 );
            return;
         }
         recent[ next ] = key;
         have[ next ] = true;
         next = ( next + 1 ) & 3;
      }
   }
}

// codes[slot] = code of the slot's offset (slots of padding keep kPadCode)
__global__ void __launch_bounds__( kFT ) code_encode( SegView v, const int32_t* __restrict__ cols,
                                                      const unsigned long long* __restrict__ table,
                                                      const uint8_t* __restrict__ code_of, uint8_t* __restrict__ codes )
{
   for( int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < v.n; r += (int64_t) gridDim.x * blockDim.x ) {
      int32_t w = 0;
      ell_slot( v, r, 0, &w );
      for( int32_t k = 0; k < w; k++ ) {
         const int64_t g = ell_slot( v, r, k, &w );
         const int32_t c = cols[ g ];
         if( c < 0 )
            continue;
         const uint32_t key = (uint32_t) ( c - (int32_t) r );
         uint32_t h = offset_hash( key );
         for( int probe = 0; probe < kHash; probe++, h = ( h + 1 ) & ( kHash - 1 ) )
            if( table[ h ] == (unsigned long long) key ) {
               codes[ g ] = code_of[ h ];
               break;
            }
      }
   }
}

// The SpMV kernels' nin block partial sums folded, in order, into the
// gridDim.x partial sums the following kernels read.
__global__ void __launch_bounds__( kFT ) fold_parts( const double* __restrict__ in, int64_t nin,
                                                     double* __restrict__ out, const CgDev* __restrict__ st )
{
   __shared__ double sm[ kFT / 64 ];
   if( ! st->active )
      return;
   const int64_t chunk = ( nin + gridDim.x - 1 ) / gridDim.x;
   const int64_t b = blockIdx.x * chunk, e = min( nin, b + chunk );
   double a = 0;
   for( int64_t i = b + threadIdx.x; i < e; i += kFT )
      a += in[ i ];
   put_part( a, sm, out );
}

// any other format: the format's own SpMV kernel, then this partial dot
template< typename T >
__global__ void __launch_bounds__( kFT ) dot_part( int64_t n, const T* __restrict__ a, const T* __restrict__ b,
                                                   double* __restrict__ part, const CgDev* __restrict__ st )
{
   __shared__ double sm[ kFT / 64 ];
   if( ! st->active )
      return;
   double acc = 0;
   for( int64_t i0 = (int64_t) blockIdx.x * kFT + threadIdx.x; i0 < n; i0 += kU * (int64_t) gridDim.x * kFT )
#pragma unroll
      for( int u = 0; u < kU; u++ ) {
         const int64_t i = i0 + u * (int64_t) gridDim.x * kFT;
         if( i < n )
            acc += (double) a[ i ] * (double) b[ i ];
      }
   put_part( acc, sm, part );
}

// ---- alpha = s1 / (Ap, p);  x += alpha p;  r -= alpha Ap;  partial sums of
// (r, z) and, for the mixed solver, (r, r);  z = r / d
template< typename T, typename X, typename D >
__global__ void __launch_bounds__( kFT ) cg_upd( int64_t n, const double* __restrict__ part1, T* __restrict__ r,
                                                 const T* __restrict__ p, const T* __restrict__ Ap,
                                                 X* __restrict__ x, const D* __restrict__ d,
                                                 double* __restrict__ part2, double* __restrict__ part3,
                                                 const CgDev* __restrict__ st )
{
   __shared__ double sm[ kFT / 64 ];
   if( ! st->active )
      return;
   const double s2 = sum_parts( part1, sm, (int) gridDim.x );  // the producers' grid is this grid
   if( s2 == 0.0 )
      return;  // cg_dir ends the solve
   const T alpha = (T) ( st->s1 / s2 );
   double a2 = 0, a3 = 0;
   for( int64_t i0 = (int64_t) blockIdx.x * kFT + threadIdx.x; i0 < n; i0 += kU * (int64_t) gridDim.x * kFT )
#pragma unroll
      for( int u = 0; u < kU; u++ ) {
         const int64_t i = i0 + u * (int64_t) gridDim.x * kFT;
         if( i < n ) {
            const T pi = p[ i ];
            x[ i ] += (X) ( alpha * pi );
            const T ri = r[ i ] - alpha * Ap[ i ];
            r[ i ] = ri;
            const T zi = d ? ri / (T) d[ i ] : ri;
            a2 += (double) ri * (double) zi;
            a3 += (double) ri * (double) ri;
         }
      }
   put_part( a2, sm, part2 );
   if( part3 )
      put_part( a3, sm, part3 );
}

// ---- the end of one iteration: every block computes the scalars from the
// partial sums (the same values in every block), block 0 writes the next
// state, every block updates its share of p = z + beta p.
// FP64 CG: residue sqrt(s1) / normb and TNL's stopping rule
// (IterativeSolver.hpp:98-121; nextIteration() counts every check).
template< typename T, typename D >
__global__ void __launch_bounds__( kFT ) cg_dir( int64_t n, const double* __restrict__ part1,
                                                 const double* __restrict__ part2, const double* __restrict__ part3,
                                                 const T* __restrict__ r, const D* __restrict__ d, T* __restrict__ p,
                                                 const CgDev* __restrict__ cur, CgDev* __restrict__ nxt, StopRule R )
{
   __shared__ double sm[ 3 * ( kFT / 64 ) ];
   const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
   if( ! cur->active ) {
      if( lead )
         *nxt = *cur;
      return;
   }
   double t[ 3 ] = { 0, 0, 0 };
   if( R.kind == 1 ) {
      double t3[ 3 ];
      sum_parts_n< 3 >( { part1, part2, part3 }, sm, t3, (int) gridDim.x );
      t[ 0 ] = t3[ 0 ], t[ 1 ] = t3[ 1 ], t[ 2 ] = t3[ 2 ];
   }
   else {
      double t2[ 2 ];
      sum_parts_n< 2 >( { part1, part2 }, sm, t2, (int) gridDim.x );
      t[ 0 ] = t2[ 0 ], t[ 1 ] = t2[ 1 ];
   }
   const double s2 = t[ 0 ];
   const double s1n = t[ 1 ];
   if( s2 == 0.0 ) {
      if( lead ) {
         *nxt = *cur;
         nxt->res = 0.0;
         nxt->s2zero = 1;
         nxt->active = 0;
      }
      return;
   }
   const double s1 = cur->s1;
   const double beta = s1 == 0.0 ? 0.0 : s1n / s1;
   const double res = sqrt( R.kind == 1 ? t[ 2 ] : s1n ) / cur->normb;
   const bool go = check_next( R, res, cur->it );
   if( lead ) {
      *nxt = *cur;
      nxt->s1 = s1n;
      nxt->beta = beta;
      nxt->res = res;
      nxt->active = go ? 1 : 0;
      nxt->it = cur->it + 1;
   }
   if( ! go )
      return;
   const T tb = (T) beta;
   for( int64_t i0 = (int64_t) blockIdx.x * kFT + threadIdx.x; i0 < n; i0 += kU * (int64_t) gridDim.x * kFT )
#pragma unroll
      for( int u = 0; u < kU; u++ ) {
         const int64_t i = i0 + u * (int64_t) gridDim.x * kFT;
         if( i < n ) {
            const T ri = r[ i ];
            p[ i ] = ( d ? ri / (T) d[ i ] : ri ) + tb * p[ i ];
         }
      }
}

// Mixed: the FP32 iteration's end; asks for a reliable update (and leaves p
// to it) when the residual has dropped by delta since the last update, when
// its estimate reaches the tolerance, or when the iteration budget is spent.
__global__ void __launch_bounds__( kFT ) mx_dir( int64_t n, const double* __restrict__ part1,
                                                 const double* __restrict__ part2, const double* __restrict__ part3,
                                                 const float* __restrict__ r, const float* __restrict__ d,
                                                 float* __restrict__ p, const CgDev* __restrict__ cur,
                                                 CgDev* __restrict__ nxt, double delta2, double tol, int64_t maxit )
{
   __shared__ double sm[ 3 * ( kFT / 64 ) ];
   const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
   if( ! cur->active ) {
      if( lead )
         *nxt = *cur;
      return;
   }
   double t[ 3 ];
   sum_parts_n< 3 >( { part1, part2, part3 }, sm, t, (int) gridDim.x );
   const double s2 = t[ 0 ], s1n = t[ 1 ], rr = t[ 2 ];
   const double s1 = cur->s1;
   const int64_t it = cur->it + 1;
   const double beta = s1 == 0.0 ? 0.0 : s1n / s1;
   const double rr_max = fmax( cur->rr_max, rr );
   const bool want = s2 == 0.0 || rr < delta2 * rr_max || sqrt( rr ) < tol * cur->normb || it >= maxit || isnan( rr );
   if( lead ) {
      *nxt = *cur;
      nxt->it = it;
      nxt->s1_prev = s1;
      if( s2 != 0.0 ) {
         nxt->s1 = s1n;
         nxt->beta = beta;
         nxt->rr = rr;
         nxt->rr_max = rr_max;
      }
      if( want ) {
         nxt->want_update = 1;
         nxt->active = 0;
      }
   }
   if( want )
      return;
   const float tb = (float) beta;
   for( int64_t i0 = (int64_t) blockIdx.x * kFT + threadIdx.x; i0 < n; i0 += kU * (int64_t) gridDim.x * kFT )
#pragma unroll
      for( int u = 0; u < kU; u++ ) {
         const int64_t i = i0 + u * (int64_t) gridDim.x * kFT;
         if( i < n )
            p[ i ] = r[ i ] / d[ i ] + tb * p[ i ];
      }
}

// ---- set-up kernels
__global__ void cg_start( CgDev* st, const double* s1, const double* normb, const double* rr, StopRule R )
{
   st->s1 = s1[ 0 ];
   st->normb = normb[ 0 ];
   st->res = sqrt( R.kind == 1 ? rr[ 0 ] : s1[ 0 ] ) / normb[ 0 ];
   st->beta = 0;
   st->s2zero = 0;
   st->want_update = 0;
   st->active = check_next( R, st->res, 0 ) ? 1 : 0;
   st->it = 1;  // nextIteration() counts the admitted (or refused) iteration
}

// r = b - Ap (FP64), partial sums of r.r
__global__ void __launch_bounds__( kFT ) resid_part( int64_t n, const double* __restrict__ b, double* __restrict__ r,
                                                     double* __restrict__ part )
{
   __shared__ double sm[ kFT / 64 ];
   double acc = 0;
   for( int64_t i = (int64_t) blockIdx.x * kFT + threadIdx.x; i < n; i += (int64_t) gridDim.x * kFT ) {
      const double ri = b[ i ] - r[ i ];
      r[ i ] = ri;
      acc += ri * ri;
   }
   put_part( acc, sm, part );
}

__global__ void __launch_bounds__( kFT ) sum_into( const double* __restrict__ part, int np, double* out )
{
   __shared__ double sm[ kFT / 64 ];
   const double v = sum_parts( part, sm, np );
   if( threadIdx.x == 0 )
      out[ 0 ] = v;
}

// Mixed restart after a residual replacement: r = (float) r64, xc = 0, and
// the partial sums of (r, z), (r, r) and (r, p), z = r / d.
__global__ void __launch_bounds__( kFT ) mx_restart( int64_t n, const double* __restrict__ r64,
                                                     float* __restrict__ r, const float* __restrict__ d,
                                                     const float* __restrict__ p, float* __restrict__ xc,
                                                     double* __restrict__ part2, double* __restrict__ part3,
                                                     double* __restrict__ part4 )
{
   __shared__ double sm[ kFT / 64 ];
   double a2 = 0, a3 = 0, a4 = 0;
   for( int64_t i = (int64_t) blockIdx.x * kFT + threadIdx.x; i < n; i += (int64_t) gridDim.x * kFT ) {
      const float ri = (float) r64[ i ];
      const float zi = ri / d[ i ];
      r[ i ] = ri;
      xc[ i ] = 0.0f;
      a2 += (double) ri * (double) zi;
      a3 += (double) ri * (double) ri;
      a4 += (double) ri * (double) p[ i ];
   }
   put_part( a2, sm, part2 );
   put_part( a3, sm, part3 );
   put_part( a4, sm, part4 );
}

// The scalars after a replacement.  first: p = z.  Otherwise the kept
// direction is made orthogonal to the new residual (p -= ((r, p) / (r, z)) z,
// as exact CG has (r_{k+1}, p_k) = 0) and p = z + beta p with
// beta = (r, z) / (r, z)_previous -- the residual replacement of Clark et al.
__global__ void __launch_bounds__( kFT ) mx_state( CgDev* st, const double* __restrict__ part2,
                                                   const double* __restrict__ part3, const double* __restrict__ part4,
                                                   const double* normb, int64_t it, int first, int np )
{
   __shared__ double sm[ kFT / 64 ];
   const double s1 = sum_parts( part2, sm, np );
   const double rr = sum_parts( part3, sm, np );
   const double rp = sum_parts( part4, sm, np );
   if( threadIdx.x != 0 )
      return;
   if( first ) {
      st->beta = 0.0;
      st->proj = 0.0;
   }
   else {
      st->proj = s1 != 0.0 ? rp / s1 : 0.0;
      st->beta = st->s1_prev != 0.0 ? s1 / st->s1_prev : 0.0;
   }
   st->s1 = s1;
   st->rr = rr;
   st->rr_max = rr;
   st->normb = normb[ 0 ];
   st->want_update = 0;
   st->it = it;
   st->active = 1;
}

// p = z + beta (p - proj z), z = r / d (p = z on the first start)
__global__ void __launch_bounds__( kFT ) mx_redir( int64_t n, const float* __restrict__ r, const float* __restrict__ d,
                                                   float* __restrict__ p, const CgDev* __restrict__ st, int first )
{
   const float beta = (float) st->beta, proj = (float) st->proj;
   for( int64_t i = (int64_t) blockIdx.x * kFT + threadIdx.x; i < n; i += (int64_t) gridDim.x * kFT ) {
      const float zi = r[ i ] / d[ i ];
      p[ i ] = first ? zi : zi + beta * ( p[ i ] - proj * zi );
   }
}

// x += xc (FP64 accumulation of the FP32 correction)
__global__ void __launch_bounds__( kFT ) mx_accumulate( int64_t n, const float* __restrict__ xc, double* __restrict__ x )
{
   for( int64_t i = (int64_t) blockIdx.x * kFT + threadIdx.x; i < n; i += (int64_t) gridDim.x * kFT )
      x[ i ] += (double) xc[ i ];
}

template< typename D >
__global__ void to_f32( int64_t n, const D* __restrict__ a, float* __restrict__ o )
{
   for( int64_t i = (int64_t) blockIdx.x * kFT + threadIdx.x; i < n; i += (int64_t) gridDim.x * kFT )
      o[ i ] = (float) a[ i ];
}

// ---- which SpMV stage a matrix gets
struct Stage
{
   const tnlmi_matrix* A = nullptr;
   const tnlmi_spmv_plan* plan = nullptr;
   SegView v{};
   int kind = 0;  // 0: the format's own SpMV kernel + dot_part, 1: fused Ellpack/SlicedEllpack column-major,
                  // 2: fused one-launch tiled CSR (no row longer than kWholeRowMax)
   int64_t grid = 0;                      // blocks of the fused SpMV kernel
   int rows_per_thread = 1;               // Ellpack / SlicedEllpack kernel
   std::shared_ptr< DeviceArray< double > > part;  // its per-block partial sums
   // the slots kind 1 reads: the matrix's own, or (a CSR matrix with short
   // rows) the stage's column-major Ellpack copy of them
   const int32_t* cols = nullptr;
   const void* vals = nullptr;
   int64_t storage = 0;
   std::shared_ptr< DeviceArray< int32_t > > ecols;
   std::shared_ptr< DeviceArray< char > > evals;
   // kind 1 with 8-bit column codes (the inner matrix of the mixed solver)
   std::shared_ptr< DeviceArray< uint8_t > > codes;
   std::shared_ptr< DeviceArray< int32_t > > dict;
   int32_t ndict = 0;
};

// The column codes of a column-major Ellpack / SlicedEllpack matrix, or
// nothing (more than kCodes distinct offsets).  Set TNLMI_COLUMN_CODES=0 in
// the environment to keep the int32 columns (identical results, more bytes).
void make_codes( Stage& S, hipStream_t s )
{
   const char* env = std::getenv( "TNLMI_COLUMN_CODES" );
   if( env && std::string( env ) == "0" )
      return;
   const int64_t storage = S.storage;
   DeviceArray< unsigned long long > table;
   DeviceArray< int32_t > flags;  // count, overflow
   table.allocate( kHash );
   flags.allocate( 2 );
   TNLMI_HIP( hipMemsetAsync( table.ptr, 0xff, sizeof( unsigned long long ) * kHash, s ) );
   TNLMI_HIP( hipMemsetAsync( flags.ptr, 0, sizeof( int32_t ) * 2, s ) );
   const unsigned grid = (unsigned) std::min< int64_t >( 4096, std::max< int64_t >( 1, ceil_div( S.v.n, kFT ) ) );
   hipLaunchKernelGGL( code_collect, dim3( grid ), dim3( kFT ), 0, s, S.v, S.cols, table.ptr, flags.ptr,
                       flags.ptr + 1 );
   TNLMI_HIP( hipGetLastError() );
   std::vector< unsigned long long > ht( kHash );
   int32_t hf[ 2 ];
   TNLMI_HIP( hipMemcpyAsync( ht.data(), table.ptr, sizeof( unsigned long long ) * kHash, hipMemcpyDeviceToHost, s ) );
   TNLMI_HIP( hipMemcpyAsync( hf, flags.ptr, sizeof( hf ), hipMemcpyDeviceToHost, s ) );
   TNLMI_HIP( hipStreamSynchronize( s ) );
   if( hf[ 1 ] != 0 || hf[ 0 ] > kCodes )
      return;
   std::vector< int32_t > keys;
   for( const auto k : ht )
      if( k != kEmpty )
         keys.push_back( (int32_t) (uint32_t) k );
   std::sort( keys.begin(), keys.end() );
   std::vector< uint8_t > code_of( kHash, kPadCode );
   for( int h = 0; h < kHash; h++ )
      if( ht[ h ] != kEmpty )
         code_of[ h ] = (uint8_t) ( std::lower_bound( keys.begin(), keys.end(), (int32_t) (uint32_t) ht[ h ] )
                                    - keys.begin() );
   DeviceArray< uint8_t > dcode_of;
   dcode_of.allocate( kHash );
   S.dict = std::make_shared< DeviceArray< int32_t > >();
   S.dict->allocate( std::max< int64_t >( 1, (int64_t) keys.size() ) );
   S.codes = std::make_shared< DeviceArray< uint8_t > >();
   // R bytes are read per row group: room for the last group's overhang
   S.codes->allocate( storage + 16 );
   TNLMI_HIP( hipMemcpyAsync( dcode_of.ptr, code_of.data(), kHash, hipMemcpyHostToDevice, s ) );
   if( ! keys.empty() )
      TNLMI_HIP( hipMemcpyAsync( S.dict->ptr, keys.data(), sizeof( int32_t ) * keys.size(), hipMemcpyHostToDevice,
                                 s ) );
   TNLMI_HIP( hipMemsetAsync( S.codes->ptr, kPadCode, storage + 16, s ) );
   hipLaunchKernelGGL( code_encode, dim3( grid ), dim3( kFT ), 0, s, S.v, S.cols, table.ptr, dcode_of.ptr,
                       S.codes->ptr );
   TNLMI_HIP( hipGetLastError() );
   TNLMI_HIP( hipStreamSynchronize( s ) );
   S.ndict = (int32_t) keys.size();
}

// kind 1 launch shape over S.v (column-major Ellpack / SlicedEllpack)
void ell_shape( Stage& S, hipStream_t s, bool codes )
{
   S.kind = 1;
   // 16-byte loads of the column indexes for 4-byte values, of the values for FP64
   S.rows_per_thread = S.A->value_type == TNLMI_F64 ? 2 : 4;
   if( S.v.format == TNLMI_SLICED_ELLPACK && S.v.slice_size % S.rows_per_thread != 0 )
      S.rows_per_thread = 1;
   S.grid = ceil_div( ceil_div( S.v.n, S.rows_per_thread ), kFT );
   S.part = std::make_shared< DeviceArray< double > >();
   S.part->allocate( S.grid );
   if( codes )
      make_codes( S, s );
}

// CSR rows as column-major Ellpack slots: slot j of row r at j * stride + r
// (padding: column -1, value 0); rows [n, stride) are all padding
template< typename V >
__global__ void csr_to_ell( int64_t n, int64_t stride, int32_t w, const int32_t* __restrict__ off,
                            const int32_t* __restrict__ cols, const V* __restrict__ vals, int32_t* __restrict__ ec,
                            V* __restrict__ ev )
{
   const int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( r >= stride )
      return;
   int32_t b = 0, e = 0;
   if( r < n ) {
      b = off[ r ];
      e = off[ r + 1 ];
   }
   for( int32_t j = 0; j < w; j++ ) {
      const bool in = b + j < e;
      ec[ j * stride + r ] = in ? cols[ b + j ] : -1;
      ev[ j * stride + r ] = in ? vals[ b + j ] : V{};
   }
}

// A CSR matrix whose rows are all short (at most kWholeRowMax entries) and
// about equally long (the Ellpack copy at most 1/8 larger than the CSR
// arrays): the stage keeps a column-major Ellpack copy of its slots, made
// here on every solve, and runs the fused Ellpack kernels on it.  Every row
// keeps its entries in CSR order, so the products and sums are those of the
// Ellpack kernels on the same matrix stored as Ellpack.  TNLMI_CG_CSR_COPY=0
// in the environment keeps the one-launch CSR kernel instead, and so does a
// device without room for the copy.
bool csr_ell_copy( Stage& S, hipStream_t s )
{
   const char* env = std::getenv( "TNLMI_CG_CSR_COPY" );
   if( env && std::string( env ) == "0" )
      return false;
   const tnlmi_matrix& A = *S.A;
   const tnlmi_segments& seg = *A.segments;
   const int64_t n = S.v.n;
   const int32_t w = std::max( 1, seg.max_size );
   const int64_t stride = ceil_div( std::max< int64_t >( n, 1 ), 64 ) * 64;
   if( seg.max_size > kWholeRowMax || n == 0 || stride * w > seg.elements + seg.elements / 8 )
      return false;
   const size_t vs = A.value_type == TNLMI_F64 ? 8 : A.value_type == TNLMI_F32 ? 4 : 2;
   S.ecols = std::make_shared< DeviceArray< int32_t > >();
   S.evals = std::make_shared< DeviceArray< char > >();
   try {
      S.ecols->allocate( stride * w );
      S.evals->allocate( stride * w * (int64_t) vs );
   }
   catch( const Error& e ) {
      // no room for the copy: the one-launch CSR kernel on the matrix itself
      if( e.status != TNLMI_BAD_ALLOC )
         throw;
      (void) hipGetLastError();
      S.ecols.reset();
      S.evals.reset();
      return false;
   }
   const dim3 grid( (unsigned) ceil_div( stride, 256 ) ), block( 256 );
   switch( A.value_type ) {
      case TNLMI_F64:
         hipLaunchKernelGGL( csr_to_ell< double >, grid, block, 0, s, n, stride, w, seg.offsets.ptr, A.column_indexes,
                             (const double*) A.values, S.ecols->ptr, (double*) S.evals->ptr );
         break;
      case TNLMI_F32:
         hipLaunchKernelGGL( csr_to_ell< float >, grid, block, 0, s, n, stride, w, seg.offsets.ptr, A.column_indexes,
                             (const float*) A.values, S.ecols->ptr, (float*) S.evals->ptr );
         break;
      default:
         hipLaunchKernelGGL( csr_to_ell< bf16 >, grid, block, 0, s, n, stride, w, seg.offsets.ptr, A.column_indexes,
                             (const bf16*) A.values, S.ecols->ptr, (bf16*) S.evals->ptr );
         break;
   }
   TNLMI_HIP( hipGetLastError() );
   SegView v{};
   v.format = TNLMI_ELLPACK;
   v.col_major = 1;
   v.n = n;
   v.ell_width = w;
   v.ell_aligned = (int32_t) stride;
   S.v = v;
   S.cols = S.ecols->ptr;
   S.vals = S.evals->ptr;
   S.storage = stride * w;
   return true;
}

Stage make_stage( const tnlmi_matrix* A, const tnlmi_spmv_plan* plan, hipStream_t s, bool codes = false )
{
   Stage S;
   S.A = A;
   S.plan = plan;
   S.v = make_view( *A->segments );
   S.cols = A->column_indexes;
   S.vals = A->values;
   S.storage = A->segments->storage;
   if( plan )
      return S;
   const int fmt = S.v.format;
   if( ( fmt == TNLMI_ELLPACK || fmt == TNLMI_SLICED_ELLPACK ) && S.v.col_major && A->matrix_type != TNLMI_SYMMETRIC ) {
      ell_shape( S, s, codes );
      return S;
   }
   const tnlmi_segments& seg = *A->segments;
   if( fmt == TNLMI_CSR && A->matrix_type != TNLMI_SYMMETRIC && csr_ell_copy( S, s ) ) {
      ell_shape( S, s, codes );
      return S;
   }
   if( fmt == TNLMI_CSR && A->matrix_type != TNLMI_SYMMETRIC && seg.max_size <= kWholeRowMax && seg.tiles > 0
       && ( seg.tile_nnz == kTileNnz || seg.tile_nnz == kTileNnzSmall ) ) {
      S.kind = 2;
      S.grid = ceil_div( seg.tiles, 4 );
      S.part = std::make_shared< DeviceArray< double > >();
      S.part->allocate( S.grid );
   }
   return S;
}

template< typename V, typename T >
void stage_launch_v( const Stage& S, const T* p, T* Ap, double* part, const CgDev* st, hipStream_t s )
{
   const V* vals = (const V*) S.vals;
   double* bp = S.part->ptr;
   if( S.kind == 1 && S.codes ) {
      const uint8_t* c = S.codes->ptr;
      const int32_t* d = S.dict->ptr;
      if( S.rows_per_thread == 4 )
         hipLaunchKernelGGL( ( spmv_dot_ell_code< 4, V, T > ), dim3( (unsigned) S.grid ), dim3( kFT ), 0, s, S.v, c, d,
                             S.ndict, vals, p, Ap, bp, st );
      else if( S.rows_per_thread == 2 )
         hipLaunchKernelGGL( ( spmv_dot_ell_code< 2, V, T > ), dim3( (unsigned) S.grid ), dim3( kFT ), 0, s, S.v, c, d,
                             S.ndict, vals, p, Ap, bp, st );
      else
         hipLaunchKernelGGL( ( spmv_dot_ell_code< 1, V, T > ), dim3( (unsigned) S.grid ), dim3( kFT ), 0, s, S.v, c, d,
                             S.ndict, vals, p, Ap, bp, st );
   }
   else if( S.kind == 2 ) {
      const tnlmi_segments& seg = *S.A->segments;
      const int32_t tail_max = std::max( 0, seg.max_size - 1 );
      if( seg.tile_nnz == kTileNnzSmall )
         hipLaunchKernelGGL( ( spmv_dot_csr< V, T, kTileNnzSmall > ), dim3( (unsigned) S.grid ), dim3( kFT ), 0, s,
                             seg.offsets.ptr, seg.tile_rows.ptr, seg.tiles, seg.elements, S.v.n, S.A->column_indexes,
                             vals, p, Ap, bp, st, tail_max );
      else
         hipLaunchKernelGGL( ( spmv_dot_csr< V, T, kTileNnz > ), dim3( (unsigned) S.grid ), dim3( kFT ), 0, s,
                             seg.offsets.ptr, seg.tile_rows.ptr, seg.tiles, seg.elements, S.v.n, S.A->column_indexes,
                             vals, p, Ap, bp, st, tail_max );
   }
   else if( S.kind == 1 ) {
      if( S.rows_per_thread == 4 )
         hipLaunchKernelGGL( ( spmv_dot_ell< 4, V, T > ), dim3( (unsigned) S.grid ), dim3( kFT ), 0, s, S.v,
                             S.cols, vals, p, Ap, bp, st );
      else if( S.rows_per_thread == 2 )
         hipLaunchKernelGGL( ( spmv_dot_ell< 2, V, T > ), dim3( (unsigned) S.grid ), dim3( kFT ), 0, s, S.v,
                             S.cols, vals, p, Ap, bp, st );
      else
         hipLaunchKernelGGL( ( spmv_dot_ell< 1, V, T > ), dim3( (unsigned) S.grid ), dim3( kFT ), 0, s, S.v,
                             S.cols, vals, p, Ap, bp, st );
   }
   hipLaunchKernelGGL( fold_parts, dim3( blocks_of< T >() ), dim3( kFT ), 0, s, bp, S.grid, part, st );
}

// Ap = A p and the partial sums of (Ap, p)
template< typename T >
void stage_launch( const Stage& S, const T* p, T* Ap, double* part, const CgDev* st, hipStream_t s )
{
   constexpr tnlmi_dtype xt = std::is_same_v< T, double > ? TNLMI_F64 : TNLMI_F32;
   if( S.kind == 0 ) {
      if( S.plan )
         plan_execute( *S.plan, xt, p, Ap, xt, 1.0, 0.0, 0, S.v.n, s );
      else {
         const tnlmi_launch cfg{ TNLMI_MAP_AUTO, 1 };
         spmv_dispatch( *S.A, xt, p, Ap, xt, 1.0, 0.0, 0, S.v.n, cfg, s );
      }
      hipLaunchKernelGGL( dot_part< T >, dim3( blocks_of< T >() ), dim3( kFT ), 0, s, S.v.n, Ap, p, part, st );
      return;
   }
   switch( S.A->value_type ) {
      case TNLMI_F64:
         stage_launch_v< double, T >( S, p, Ap, part, st, s );
         return;
      case TNLMI_F32:
         stage_launch_v< float, T >( S, p, Ap, part, st, s );
         return;
      case TNLMI_BF16:
         stage_launch_v< bf16, T >( S, p, Ap, part, st, s );
         return;
   }
   fail( TNLMI_INVALID_ARGUMENT, "fused CG: unsupported value type" );
}

struct Parts
{
   DeviceArray< double > a, b, c;
   Parts()
   {
      a.allocate( kFB );
      b.allocate( kFB );
      c.allocate( kFB );
   }
};

// pinned host copy of the device state
struct HostState
{
   CgDev* h = nullptr;
   HostState() { TNLMI_HIP( hipHostMalloc( (void**) &h, sizeof( CgDev ), hipHostMallocDefault ) ); }
   ~HostState()
   {
      if( h )
         (void) hipHostFree( h );
   }
   const CgDev& read( const CgDev* d, hipStream_t s )
   {
      TNLMI_HIP( hipMemcpyAsync( h, d, sizeof( CgDev ), hipMemcpyDeviceToHost, s ) );
      TNLMI_HIP( hipStreamSynchronize( s ) );
      return *h;
   }
};

}  // namespace

void fused_cg_solve( const Operator& A, const tnlmi_preconditioner* M, const double* b, double* x,
                     tnlmi_solver_state* st, hipStream_t s, int residue_kind, bool column_codes )
{
   const int64_t n = A.rows;
   const double* d = precond_diagonal( M );
   DeviceArray< double > r, p, Ap, sc;
   DeviceArray< CgDev > dev;
   r.allocate( n );
   p.allocate( n );
   Ap.allocate( n );
   sc.allocate( 4 );
   dev.allocate( 2 );
   Parts parts;
   RedWork w;
   HostState hs;
   const StopRule R{ st->max_iterations, st->min_iterations, st->convergence_residue, st->divergence_residue,
                     residue_kind };
   // normb = ||M^{-1} b|| (CG.hpp:24-29); ||b|| for the recurrence residual
   if( d && residue_kind == 0 ) {
      precond_apply( M, n, b, r.ptr, s );
      nrm2( n, r.ptr, sc.ptr + 1, w, s );
   }
   else
      nrm2( n, b, sc.ptr + 1, w, s );
   double normb = 0;
   TNLMI_HIP( hipMemcpyAsync( &normb, sc.ptr + 1, sizeof( double ), hipMemcpyDeviceToHost, s ) );
   TNLMI_HIP( hipStreamSynchronize( s ) );
   if( normb == 0.0 ) {
      fill( n, 0.0, x, s );
      st->iterations = 0;
      st->residue = 0.0;
      st->converged = 1;
      return;
   }
   // r = b - A x; p = z = M^{-1} r; s1 = (r, z)
   A.apply( x, r.ptr, s );
   rsub( n, b, r.ptr, s );
   precond_apply( M, n, r.ptr, p.ptr, s );
   dot( n, r.ptr, p.ptr, sc.ptr + 0, w, s );
   if( residue_kind == 1 )
      dot( n, r.ptr, r.ptr, sc.ptr + 2, w, s );
   hipLaunchKernelGGL( cg_start, dim3( 1 ), dim3( 1 ), 0, s, dev.ptr, sc.ptr + 0, sc.ptr + 1, sc.ptr + 2, R );
   TNLMI_HIP( hipGetLastError() );
   const Stage S = make_stage( A.A, A.plan, s, column_codes );
   // the state is double-buffered by iteration parity: iteration k reads
   // dev[k & 1] and its last kernel writes dev[(k + 1) & 1]
   int par = 0;
   const CgDev* h = &hs.read( dev.ptr, s );
   while( h->active ) {
      for( int k = 0; k < kBatch; k++, par ^= 1 ) {
         CgDev* cur = dev.ptr + par;
         stage_launch< double >( S, p.ptr, Ap.ptr, parts.a.ptr, cur, s );
         double* part3 = residue_kind == 1 ? parts.c.ptr : nullptr;
         hipLaunchKernelGGL( ( cg_upd< double, double, double > ), dim3( blocks_of< double >() ), dim3( kFT ), 0, s, n,
                             parts.a.ptr,
                             r.ptr, p.ptr, Ap.ptr, x, d, parts.b.ptr, part3, cur );
         hipLaunchKernelGGL( ( cg_dir< double, double > ), dim3( blocks_of< double >() ), dim3( kFT ), 0, s, n,
                             parts.a.ptr,
                             parts.b.ptr, part3, r.ptr, d, p.ptr, cur, dev.ptr + ( par ^ 1 ), R );
      }
      TNLMI_HIP( hipGetLastError() );
      h = &hs.read( dev.ptr + par, s );
   }
   Stop S2{ st };
   S2.it = h->it;
   S2.res = h->res;
   st->iterations = h->it;
   st->residue = h->res;
   st->converged = S2.converged();
}

}  // namespace tnlmi

using namespace tnlmi;

extern "C" {

tnlmi_status tnlmi_mixed_cg_solve( const tnlmi_matrix* A, const tnlmi_matrix* A_lo, const tnlmi_spmv_plan* plan,
                                   double delta, const double* d_b, double* d_x, tnlmi_solver_state* st,
                                   int64_t* updates, tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( ! A || ! A_lo || ! A->segments || ! A_lo->segments || ! d_b || ! d_x || ! st )
         fail( TNLMI_INVALID_ARGUMENT, "mixed_cg_solve: NULL argument" );
      if( A->value_type != TNLMI_F64 )
         fail( TNLMI_INVALID_ARGUMENT, "mixed_cg_solve: the outer matrix must hold FP64 values" );
      if( A_lo->value_type != TNLMI_F32 && A_lo->value_type != TNLMI_BF16 )
         fail( TNLMI_INVALID_ARGUMENT, "mixed_cg_solve: the inner matrix must hold FP32 or BF16 values" );
      if( A_lo->rows != A->rows || A_lo->columns != A->columns || A->rows != A->columns )
         fail( TNLMI_INVALID_ARGUMENT, "mixed_cg_solve: the matrices must be square and of one size" );
      if( ! ( delta > 0.0 && delta < 1.0 ) )
         fail( TNLMI_INVALID_ARGUMENT, "mixed_cg_solve: delta must lie in (0, 1)" );
      hipStream_t s = as_stream( stream );
      const int64_t n = A->rows;
      const Operator op{ A, plan, n };
      DeviceArray< double > r64, d64, sc;
      DeviceArray< float > r, p, Ap, xc, d32;
      DeviceArray< CgDev > dev;
      for( auto* v : { &r64, &d64 } )
         v->allocate( n );
      for( auto* v : { &r, &p, &Ap, &xc, &d32 } )
         v->allocate( n );
      sc.allocate( 4 );
      dev.allocate( 2 );
      TNLMI_HIP( hipMemsetAsync( p.ptr, 0, sizeof( float ) * n, s ) );
      Parts parts;
      RedWork w;
      HostState hs;
      // Jacobi: d = diag(A) (Diagonal.hpp:27-30); d32 = (float) d
      {
         tnlmi_preconditioner* M = nullptr;
         const tnlmi_status rc = tnlmi_preconditioner_create( TNLMI_PRECOND_DIAGONAL, A, stream, &M );
         if( rc != TNLMI_OK )
            fail( rc, tnlmi_last_error() );
         hipLaunchKernelGGL( to_f32< double >, dim3( kFB32 ), dim3( kFT ), 0, s, n, precond_diagonal( M ), d32.ptr );
         TNLMI_HIP( hipStreamSynchronize( s ) );
         tnlmi_preconditioner_destroy( M );
      }
      nrm2( n, d_b, sc.ptr + 0, w, s );
      double normb = 0;
      TNLMI_HIP( hipMemcpyAsync( &normb, sc.ptr, sizeof( double ), hipMemcpyDeviceToHost, s ) );
      TNLMI_HIP( hipStreamSynchronize( s ) );
      int64_t nupd = 0;
      if( normb == 0.0 ) {
         fill( n, 0.0, d_x, s );
         st->iterations = 0;
         st->residue = 0.0;
         st->converged = 1;
         if( updates )
            *updates = 0;
         return;
      }
      const Stage S = make_stage( A_lo, nullptr, s, true );
      const StopRule R{ st->max_iterations, st->min_iterations, st->convergence_residue, st->divergence_residue };
      int64_t it = 0;
      double res = 0;
      bool first = true;
      int par = 0;  // the state is double-buffered by iteration parity
      while( true ) {
         // reliable update: r = b - A x in FP64 and the true residual
         op.apply( d_x, r64.ptr, s );
         hipLaunchKernelGGL( resid_part, dim3( kFB32 ), dim3( kFT ), 0, s, n, d_b, r64.ptr, parts.c.ptr );
         hipLaunchKernelGGL( sum_into, dim3( 1 ), dim3( kFT ), 0, s, parts.c.ptr, kFB32, sc.ptr + 1 );
         double rr = 0;
         TNLMI_HIP( hipMemcpyAsync( &rr, sc.ptr + 1, sizeof( double ), hipMemcpyDeviceToHost, s ) );
         TNLMI_HIP( hipStreamSynchronize( s ) );
         res = std::sqrt( rr ) / normb;
         nupd++;
         if( std::isnan( res ) || it > R.maxit || ( res > R.div && it >= R.minit ) || ( res < R.tol && it >= R.minit ) )
            break;
         // replace the FP32 residual by the FP64 one and form the next direction
         CgDev* cur = dev.ptr + par;
         hipLaunchKernelGGL( mx_restart, dim3( kFB32 ), dim3( kFT ), 0, s, n, r64.ptr, r.ptr, d32.ptr, p.ptr, xc.ptr,
                             parts.a.ptr, parts.b.ptr, parts.c.ptr );
         hipLaunchKernelGGL( mx_state, dim3( 1 ), dim3( kFT ), 0, s, cur, parts.a.ptr, parts.b.ptr, parts.c.ptr,
                             sc.ptr + 0, it, first ? 1 : 0, kFB32 );
         hipLaunchKernelGGL( mx_redir, dim3( kFB32 ), dim3( kFT ), 0, s, n, r.ptr, d32.ptr, p.ptr, cur,
                             first ? 1 : 0 );
         first = false;
         const CgDev* h = nullptr;
         do {
            for( int k = 0; k < kBatch; k++, par ^= 1 ) {
               cur = dev.ptr + par;
               stage_launch< float >( S, p.ptr, Ap.ptr, parts.a.ptr, cur, s );
               hipLaunchKernelGGL( ( cg_upd< float, float, float > ), dim3( kFB32 ), dim3( kFT ), 0, s, n, parts.a.ptr,
                                   r.ptr, p.ptr, Ap.ptr, xc.ptr, d32.ptr, parts.b.ptr, parts.c.ptr, cur );
               hipLaunchKernelGGL( mx_dir, dim3( kFB32 ), dim3( kFT ), 0, s, n, parts.a.ptr, parts.b.ptr, parts.c.ptr,
                                   r.ptr, d32.ptr, p.ptr, cur, dev.ptr + ( par ^ 1 ), delta * delta, R.tol, R.maxit );
            }
            TNLMI_HIP( hipGetLastError() );
            h = &hs.read( dev.ptr + par, s );
         } while( h->active );
         it = h->it;
         hipLaunchKernelGGL( mx_accumulate, dim3( kFB32 ), dim3( kFT ), 0, s, n, xc.ptr, d_x );
      }
      st->iterations = it;
      st->residue = res;
      st->converged = ! std::isnan( res ) && res <= R.tol && it <= R.maxit ? 1 : 0;
      if( updates )
         *updates = nupd;
   } );
}

}  // extern "C"
