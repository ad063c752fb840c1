// krylov.hip -- Krylov solvers and preconditioners on the GPU, FP64.
//
// Solvers/Linear/CG.hpp, GMRES.hpp, Jacobi.hpp, Preconditioners/Diagonal.hpp,
// ILU0.hpp, Utils/LinearResidueGetter.hpp and the stopping rules of
// Solvers/IterativeSolver.hpp:98-143.  The iteration structure and the
// stopping semantics follow the reference line by line; the arithmetic runs
// in this library's kernels with every Krylov scalar (dot products, norms,
// alpha/beta, the Hessenberg matrix, Givens rotations and the least-squares
// back-substitution of GMRES) resident on the GPU.  The host reads one
// scalar (the solver's residue) per iteration to evaluate the stopping rule.
#include <algorithm>
#include <cmath>
#include <vector>

#include "blas1.hpp"
#include "krylov.hpp"
#include "segments.hpp"
#include "spmv.hpp"

using namespace tnlmi;

// ------------------------------------------------------- preconditioners
struct tnlmi_preconditioner
{
   tnlmi_precond_kind kind = TNLMI_PRECOND_NONE;
   int64_t n = 0;
   DeviceArray< double > diag;
   // ILU(0): L (strictly lower, unit diagonal implied) and U with reversed rows
   // (U row N-1-i holds row i, diagonal first), CSR on the device; rows grouped
   // by dependency level for the triangular solves.
   DeviceArray< int32_t > lp, lc, up, uc, lrows, urows;
   DeviceArray< double > lv, uv;
   std::vector< int64_t > llev, ulev;  // level boundaries into lrows / urows
   DeviceArray< int64_t > dllev, dulev;  // the same on the device (read by the one-workgroup solves)
   // launch plan of each solve: runs [l0, l1) of consecutive levels; narrow
   // runs (every level at most kLvThreads * kLvRows rows) in one one-workgroup
   // launch, a wide level alone in a launch of its own
   struct Run
   {
      int64_t l0, l1;
      bool narrow;
      bool ring;  // every dependency of the run lies within kRing positions of its level's end
      bool flow;  // every dependency lies within kFlowDist positions of its row (the sync-free solve)
   };
   std::vector< Run > lruns, uruns;
   // the factors in level-order positions (for the LDS-ring solve): per position p
   // (row lrows[p] / urows[p]) its entries [pp[p], pp[p+1]): dependency position
   // and value; the upper factor's pivot per position
   DeviceArray< int32_t > lpp, ldep, upp, udep;
   DeviceArray< double > lval, uval, udiag;
   // the same per position in fixed-size records for the LDS-ring solve (row
   // and its dependency positions, -1: none; their values and the pivot), and
   // the right-hand side of a solve gathered into position order before it
   DeviceArray< int32_t > lridx, uridx;
   DeviceArray< double > lrval, urval, bpos;
   DeviceArray< double > xpos;  // the sync-free solve's results in position order (+ kLvThreads spare slots)
   tnlmi_matrix M{};                   // explicit: x = M b
};

namespace tnlmi {
namespace {

// d_i = A(i, i): linear search of row i (Diagonal.hpp:27-30 via getElement)
template< typename V >
__global__ void diag_kernel( SegView v, const int32_t* __restrict__ cols, const V* __restrict__ vals, double* d )
{
   const int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( r >= v.n )
      return;
   double dv = 0.0;
   const int32_t size = segment_size( v, r );
   for( int32_t k = 0; k < size; k++ ) {
      const int64_t g = global_index( v, r, k );
      if( cols[ g ] == r ) {
         dv = (double) load_val< V, double >( vals, g );
         break;
      }
   }
   d[ r ] = dv;
}

// getSegmentSize of every row (an Ellpack-type row may hold entries past its
// requested capacity)
__global__ void seg_size_kernel( SegView v, int32_t* __restrict__ out )
{
   const int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( r < v.n )
      out[ r ] = segment_size( v, r );
}

// CSR-ordered copy of every slot of any format, row r at eoff[r] (the scan of
// the segment sizes; padding dropped later on the host)
template< typename V >
__global__ void to_csr_kernel( SegView v, const int32_t* __restrict__ cols, const V* __restrict__ vals,
                               const int32_t* __restrict__ eoff, int32_t* __restrict__ oc, double* __restrict__ ov )
{
   const int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( r >= v.n )
      return;
   const int32_t b = eoff[ r ], size = eoff[ r + 1 ] - b;
   for( int32_t k = 0; k < size; k++ ) {
      const int64_t g = global_index( v, r, k );
      oc[ b + k ] = cols[ g ];
      ov[ b + k ] = (double) load_val< V, double >( vals, g );
   }
}

template< typename F >
void with_value_type( tnlmi_dtype t, F&& f )
{
   if( t == TNLMI_F64 )
      f( (const double*) nullptr );
   else if( t == TNLMI_F32 )
      f( (const float*) nullptr );
   else
      f( (const bf16*) nullptr );
}

// one dependency level of the lower solve: x_i = b_i - sum_j L_ij x_j
__global__ void lower_level( const int32_t* __restrict__ rows, int64_t nrows, const int32_t* __restrict__ lp,
                             const int32_t* __restrict__ lc, const double* __restrict__ lv,
                             const double* __restrict__ b, double* __restrict__ x )
{
   const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( t >= nrows )
      return;
   const int32_t i = rows[ t ];
   double xi = b[ i ];
   for( int32_t g = lp[ i ]; g < lp[ i + 1 ]; g++ )
      xi -= lv[ g ] * x[ lc[ g ] ];
   x[ i ] = xi;
}

// one level of the upper solve on reversed rows: x_i = (x_i - sum_{j>i} U_ij x_j) / U_ii
__global__ void upper_level( const int32_t* __restrict__ rows, int64_t nrows, int64_t N,
                             const int32_t* __restrict__ up, const int32_t* __restrict__ uc,
                             const double* __restrict__ uv, double* __restrict__ x )
{
   const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( t >= nrows )
      return;
   const int32_t i = rows[ t ];
   const int64_t k = N - 1 - i;
   double xi = x[ i ];
   const double uii = up[ k ] < up[ k + 1 ] ? uv[ up[ k ] ] : 0.0;  // a row without entries has no pivot
   for( int32_t g = up[ k ] + 1; g < up[ k + 1 ]; g++ )
      xi -= uv[ g ] * x[ uc[ g ] ];
   x[ i ] = xi / uii;
}

// Consecutive narrow levels in one launch of one workgroup: the levels in
// order, a barrier after each, so a level's rows read the solution entries of
// earlier levels that other threads of the workgroup wrote.  The barrier's
// workgroup-scope release / acquire makes them visible: the workgroup runs on
// one CU, whose L1 all its waves share, so the x loads and stores are
// workgroup-scope accesses (device-scope ones would go past the XCD's L2 to
// memory: 29.7 ms against 17.4 ms per ILU(0) GMRES iteration on the 1024^2
// convection-diffusion matrix, measured).  While a level
// is being solved, each thread already loads its rows of the next level (row
// index, entry range, the first kLvPre entries, the right-hand side), so after
// a barrier only the loads of the x entries a row needs stand between the
// levels.  Summation order per row as lower_level / upper_level.
constexpr int kLvThreads = 1024;
constexpr int kLvRows = 2;  // rows per thread and level at most
constexpr int kLvPre = 4;   // entries per row loaded ahead of their level

struct LvRow
{
   int32_t i, g0, g1;  // row (-1: none), entries [g0, g1) (off-diagonal)
   int32_t c[ kLvPre ];
   double v[ kLvPre ];
   double bi, dii;  // right-hand side (lower) / current x (upper); pivot (upper)
};

__device__ __forceinline__ double coherent_load( const double* p )
{
   return __hip_atomic_load( p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP );
}
__device__ __forceinline__ void coherent_store( double* p, double v )
{
   __hip_atomic_store( p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP );
}

template< bool UPPER >
__device__ __forceinline__ void lv_load( LvRow ( &R )[ kLvRows ], const int32_t* __restrict__ rows, int64_t p0,
                                         int64_t p1, int64_t N, const int32_t* __restrict__ ptr,
                                         const int32_t* __restrict__ col, const double* __restrict__ val,
                                         const double* b, const double* x )
{
#pragma unroll
   for( int r = 0; r < kLvRows; r++ ) {
      const int64_t p = p0 + threadIdx.x + (int64_t) r * kLvThreads;
      LvRow& w = R[ r ];
      w.i = -1;
      if( p < p1 ) {
         w.i = rows[ p ];
         const int64_t k = UPPER ? N - 1 - w.i : w.i;
         const int32_t a = ptr[ k ], e = ptr[ k + 1 ];
         if( UPPER ) {
            w.dii = a < e ? val[ a ] : 0.0;  // a row without entries has no pivot
            w.g0 = a + 1;
            w.bi = coherent_load( x + w.i );
         }
         else {
            w.dii = 1.0;
            w.g0 = a;
            w.bi = b[ w.i ];
         }
         w.g1 = e;
#pragma unroll
         for( int q = 0; q < kLvPre; q++ )
            if( w.g0 + q < w.g1 ) {
               w.c[ q ] = col[ w.g0 + q ];
               w.v[ q ] = val[ w.g0 + q ];
            }
      }
   }
}

// The LDS-ring form of tri_levels: the solution entries of the latest kRing
// positions (in level order) stay in LDS -- those of the positions before
// the run, solved by earlier launches, are copied in first -- so a row reads
// the entries it depends on from LDS and the levels synchronise through LDS
// alone: the barrier waits for no global store.  Used for a run in which
// every dependency lies within kRing positions of its level's end (a
// stencil's: within two levels), checked when the factors are built.
constexpr int kRing = 8192;

constexpr int kRingPre = 3;  // entries per row loaded ahead (a 7-point stencil's triangle has 3)

// per-position records of the ring solve (16 and 32 bytes, naturally
// aligned); a ring run's rows have at most kRingPre entries besides the pivot
struct alignas( 16 ) RingIdx
{
   int32_t i, q[ kRingPre ];  // row; positions of its entries' dependencies (-1: none)
};
struct alignas( 16 ) RingVal
{
   double v[ kRingPre ], dii;  // the entries' values; the upper factor's pivot
};

struct RingRow
{
   int32_t p;  // position (-1: none)
   RingIdx r;
   RingVal v;
   double bi;
};

// the row of this thread in the level [p0, p1), read at a clamped position:
// every load is issued whatever the level's size and none depends on another,
// so the wait for a stage counts exactly the loads issued after it (a
// conditional or dependent load would make the wave wait for all of them:
// loads complete in order)
__device__ __forceinline__ void ring_load( RingRow& w, int64_t p0, int64_t p1, const RingIdx* __restrict__ ridx,
                                           const RingVal* __restrict__ rval, const double* __restrict__ bpos )
{
   const int64_t p = p0 + threadIdx.x;
   const int64_t pc = p < p1 ? p : p1 - 1;
   w.p = p < p1 ? (int32_t) p : -1;
   w.r = ridx[ pc ];
   w.v = rval[ pc ];
   w.bi = bpos[ pc ];
}

// A level's rows are requested kRingDepth levels ahead (one row per thread:
// ring runs have levels of at most kLvThreads rows), so their loads have
// kRingDepth levels' time to arrive; the stages rotate through registers by
// unrolling the level loop kRingDepth times, and a stage is reloaded after
// its level is solved (no copy of it waits for its loads).
constexpr int kRingDepth = 6;
// levels of a run at most (their bounds are staged in LDS by the ring solve)
constexpr int kRunLevels = 16384;

// bpos[p] = src[rows[p]]: a solve's right-hand side in position order
__global__ void gather_pos( int64_t n, const int32_t* __restrict__ rows, const double* __restrict__ src,
                            double* __restrict__ bpos )
{
   const int64_t p = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( p < n )
      bpos[ p ] = src[ rows[ p ] ];
}

template< bool UPPER >
__global__ void __launch_bounds__( kLvThreads )
   tri_ring( const int32_t* __restrict__ rows, const int64_t* __restrict__ bounds, int64_t l0, int64_t l1,
             const RingIdx* __restrict__ ridx, const RingVal* __restrict__ rval, const double* __restrict__ bpos,
             const double* __restrict__ x, double* __restrict__ xpos, int64_t N )
{
   __shared__ double ring[ kRing ];
   __shared__ int32_t lb[ kRunLevels + 1 ];  // the run's level bounds (a stage's addresses wait for no scalar load)
   const int64_t P0 = bounds[ l0 ];  // the run's first position: dependencies before it are in x
   for( int64_t j = threadIdx.x; j <= l1 - l0; j += kLvThreads )
      lb[ j ] = (int32_t) bounds[ l0 + j ];
   __syncthreads();
   RingRow st[ kRingDepth ];
   // levels past the run load the run's last level again (p = -1: nothing to solve)
   auto load = [ & ]( RingRow& R, int64_t l ) {
      const int32_t j = (int32_t) ( ( l < l1 ? l : l1 - 1 ) - l0 );
      ring_load( R, lb[ j ], lb[ j + 1 ], ridx, rval, bpos );
      if( l >= l1 )
         R.p = -1;
   };
#pragma unroll
   for( int d = 0; d < kRingDepth; d++ )
      load( st[ d ], l0 + d );
   // the kRing positions before the run into the ring: their rows were
   // solved by earlier launches, and every dependency of the run lies within
   // kRing positions of its level's end (checked when the factors are built),
   // so every dependency is read from LDS (a read that may go to x instead
   // would be a flat load, whose wait is a wait for every load in flight)
   for( int64_t q = P0 - kRing + threadIdx.x; q < P0; q += kLvThreads )
      if( q >= 0 )
         ring[ q & ( kRing - 1 ) ] = x[ rows[ q ] ];
   __syncthreads();
   auto xdep = [ & ]( int32_t q ) -> double { return ring[ q & ( kRing - 1 ) ]; };
   // every thread runs every stage and meets every barrier (the last group of
   // stages may run past the run with nothing to solve)
   for( int64_t l = l0; l < l1; l += kRingDepth ) {
#pragma unroll
      for( int d = 0; d < kRingDepth; d++ ) {
         const RingRow& w = st[ d ];
         double xi = 0.0;
         if( w.p >= 0 ) {
            xi = w.bi;
#pragma unroll
            for( int q = 0; q < kRingPre; q++ )
               if( w.r.q[ q ] >= 0 )
                  xi -= w.v.v[ q ] * xdep( w.r.q[ q ] );
            if( UPPER )
               xi = xi / w.v.dii;
            ring[ w.p & ( kRing - 1 ) ] = xi;
         }
         // the value into xpos, in position order: a coalesced store that
         // every lane issues (a lane without a row stores to a spare slot past
         // N), where a store into x would scatter
         xpos[ w.p >= 0 ? (int64_t) w.p : N + threadIdx.x ] = xi;
         load( st[ d ], l + d + kRingDepth );  // the stage's next level, kRingDepth ahead
         // LDS alone orders the levels: the stores of xpos are read back by no
         // later level of this launch
         __builtin_amdgcn_fence( __ATOMIC_RELEASE, "workgroup", "local" );
         __builtin_amdgcn_s_barrier();
         __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "workgroup", "local" );
      }
   }
}

// x[rows[p]] = xpos[p] for the positions [P0, P1) of a sync-free run
__global__ void scatter_pos( int64_t P0, int64_t P1, const int32_t* __restrict__ rows,
                             const double* __restrict__ xpos, double* __restrict__ x )
{
   const int64_t p = P0 + (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( p < P1 )
      x[ rows[ p ] ] = xpos[ p ];
}

// The sync-free form of a run, without a barrier per level: thread t solves
// the run's positions P0 + t, P0 + t + kLvThreads, ... in turn (chunk j =
// the kLvThreads positions P0 + j kLvThreads ...), and a row waits only for
// the positions it depends on: LDS holds a ring of solved values with a stamp
// per slot (the slot of position q holds q's value once its stamp reads q).
// The slot of q is reused by q + kRing only after every position of chunk
// chunk(q + kRing) - kFlowLag is solved (a per-chunk count of the waves done
// with it); a position's readers lie at most kFlowDist positions after it
// (checked when the factors are built), so by then they have read it.
// Every wait is bounded: a row whose wait runs past spin_cap polls is
// published as NaN and the whole launch then stops waiting (its dependents,
// and the solve, carry NaN), so a fault in the schedule cannot hang the
// workgroup.  A row's records are requested kFlowDepth chunks ahead, and every
// global load and store of an iteration is unconditional (the waits between
// them count them exactly).
constexpr int kFlowLag = 4;                                // chunks
constexpr int kFlowDist = kRing - kFlowLag * kLvThreads;  // 4096 positions
constexpr int kFlowSlots = 16;                            // chunk counters in use at once (> kFlowLag + 1)
constexpr int kFlowDepth = 4;
constexpr int64_t kFlowSpinCap = int64_t( 1 ) << 20;

template< bool UPPER >
__global__ void __launch_bounds__( kLvThreads )
   tri_flow( const int32_t* __restrict__ rows, int64_t P0, int64_t P1, const RingIdx* __restrict__ ridx,
             const RingVal* __restrict__ rval, const double* __restrict__ bpos, const double* __restrict__ x,
             double* __restrict__ xpos, int64_t N, int64_t spin_cap )
{
   __shared__ double ring[ kRing ];
   __shared__ int32_t stamp[ kRing ];
   __shared__ int32_t done[ kFlowSlots ];
   __shared__ int32_t failed;
   constexpr int32_t kMask = kRing - 1;
   const int t = threadIdx.x, lane = t & 63;
   // the kRing positions before the run (solved by earlier launches); a slot
   // of a position before 0 gets the stamp -1, which no position matches
   for( int k = t; k < kRing; k += kLvThreads ) {
      const int64_t q = P0 - kRing + k;
      const int32_t slot = (int32_t) ( q & kMask );
      if( q >= 0 ) {
         ring[ slot ] = x[ rows[ q ] ];
         stamp[ slot ] = (int32_t) q;
      }
      else
         stamp[ slot ] = -1;
   }
   if( t < kFlowSlots )
      done[ t ] = 0;
   if( t == 0 )
      failed = 0;
   __syncthreads();
   const int64_t J = ( P1 - P0 + kLvThreads - 1 ) / kLvThreads;  // chunks
   RingRow st[ kFlowDepth ];
   auto load = [ & ]( RingRow& R, int64_t j ) {
      const int64_t p = P0 + j * kLvThreads + t;
      ring_load( R, p - t, P1, ridx, rval, bpos );  // position p at a clamped index; -1 past the run
   };
#pragma unroll
   for( int d = 0; d < kFlowDepth; d++ )
      load( st[ d ], d );
   auto stamped = [ & ]( int32_t q ) {
      return __hip_atomic_load( &stamp[ q & kMask ], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP ) == q;
   };
   for( int64_t j0 = 0; j0 < J; j0 += kFlowDepth ) {
#pragma unroll
      for( int d = 0; d < kFlowDepth; d++ ) {
         const int64_t j = j0 + d;
         const RingRow& w = st[ d ];
         // the slots of chunk j were last used kRing positions ago: wait until
         // chunk j - kFlowLag is solved (the same value for every lane)
         if( j >= kFlowLag && j < J ) {
            const int64_t c = j - kFlowLag;
            const int32_t target = (int32_t) ( ( c / kFlowSlots + 1 ) * kLvThreads );
            for( int64_t n = 0; n < spin_cap; n++ ) {
               if( __hip_atomic_load( &done[ c % kFlowSlots ], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP )
                      >= target
                   || __hip_atomic_load( &failed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP ) )
                  break;
               __builtin_amdgcn_s_sleep( 1 );
            }
         }
         // this lane's row: solved as soon as its dependencies are stamped;
         // the wave passes until none of its lanes waits
         bool pending = w.p >= 0;
         double xi = 0.0;
         for( int64_t n = 0;; n++ ) {
            if( pending ) {
               bool ready = true;
#pragma unroll
               for( int q = 0; q < kRingPre; q++ )
                  if( w.r.q[ q ] >= 0 )
                     ready = ready && stamped( w.r.q[ q ] );
               const bool give_up =
                  ! ready
                  && ( n >= spin_cap || __hip_atomic_load( &failed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP ) );
               if( ready || give_up ) {
                  __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "workgroup", "local" );
                  xi = w.bi;
#pragma unroll
                  for( int q = 0; q < kRingPre; q++ )
                     if( w.r.q[ q ] >= 0 )
                        xi -= w.v.v[ q ] * ring[ w.r.q[ q ] & kMask ];
                  if( UPPER )
                     xi = xi / w.v.dii;
                  if( give_up ) {
                     xi = __builtin_nan( "" );
                     __hip_atomic_store( &failed, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP );
                  }
                  ring[ w.p & kMask ] = xi;
                  __builtin_amdgcn_fence( __ATOMIC_RELEASE, "workgroup", "local" );
                  __hip_atomic_store( &stamp[ w.p & kMask ], w.p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP );
                  pending = false;
               }
            }
            if( __ballot( pending ) == 0 )
               break;
            __builtin_amdgcn_s_sleep( 1 );
         }
         // the value into xpos, in position order (coalesced; a lane without a
         // row stores to a spare slot past N: every lane stores)
         xpos[ w.p >= 0 ? (int64_t) w.p : N + t ] = xi;
         // this wave is done with chunk j: its reads of the ring are complete
         if( j < J && lane == 0 )
            __hip_atomic_fetch_add( &done[ j % kFlowSlots ], 64, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP );
         load( st[ d ], j + kFlowDepth );  // the stage's next chunk, kFlowDepth ahead
      }
   }
}

template< bool UPPER >
__global__ void __launch_bounds__( kLvThreads )
   tri_levels( const int32_t* __restrict__ rows, const int64_t* __restrict__ bounds, int64_t l0, int64_t l1,
               int64_t N, const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
               const double* __restrict__ val, const double* b, double* x )
{
   LvRow cur[ kLvRows ], nxt[ kLvRows ];
   lv_load< UPPER >( cur, rows, bounds[ l0 ], bounds[ l0 + 1 ], N, ptr, col, val, b, x );
   for( int64_t l = l0; l < l1; l++ ) {  // uniform bounds: every thread meets every barrier
      if( l + 1 < l1 )
         lv_load< UPPER >( nxt, rows, bounds[ l + 1 ], bounds[ l + 2 ], N, ptr, col, val, b, x );
#pragma unroll
      for( int r = 0; r < kLvRows; r++ ) {
         const LvRow& w = cur[ r ];
         if( w.i < 0 )
            continue;
         double xi = w.bi;
#pragma unroll
         for( int q = 0; q < kLvPre; q++ )
            if( w.g0 + q < w.g1 )
               xi -= w.v[ q ] * coherent_load( x + w.c[ q ] );
         for( int32_t g = w.g0 + kLvPre; g < w.g1; g++ )
            xi -= val[ g ] * coherent_load( x + col[ g ] );
         coherent_store( x + w.i, UPPER ? xi / w.dii : xi );
      }
      __syncthreads();
#pragma unroll
      for( int r = 0; r < kLvRows; r++ )
         cur[ r ] = nxt[ r ];
   }
}

// rows grouped by level; level(i) = 1 + max level of the rows it depends on
void levels( int64_t N, const std::vector< int32_t >& ptr, const std::vector< int32_t >& col, bool upper,
             std::vector< int32_t >& order, std::vector< int64_t >& bounds )
{
   std::vector< int32_t > lev( N, 0 );
   int32_t maxl = 0;
   if( ! upper )
      for( int64_t i = 0; i < N; i++ ) {
         int32_t l = 0;
         for( int32_t g = ptr[ i ]; g < ptr[ i + 1 ]; g++ )
            l = std::max( l, lev[ col[ g ] ] + 1 );
         lev[ i ] = l;
         maxl = std::max( maxl, l );
      }
   else
      for( int64_t i = N - 1; i >= 0; i-- ) {
         const int64_t k = N - 1 - i;
         int32_t l = 0;
         for( int32_t g = ptr[ k ] + 1; g < ptr[ k + 1 ]; g++ )
            l = std::max( l, lev[ col[ g ] ] + 1 );
         lev[ i ] = l;
         maxl = std::max( maxl, l );
      }
   std::vector< int64_t > cnt( maxl + 2, 0 );
   for( int64_t i = 0; i < N; i++ )
      cnt[ lev[ i ] + 1 ]++;
   for( int32_t l = 0; l <= maxl; l++ )
      cnt[ l + 1 ] += cnt[ l ];
   bounds.assign( cnt.begin(), cnt.end() );
   order.assign( N, 0 );
   std::vector< int64_t > pos( cnt.begin(), cnt.end() - 1 );
   for( int64_t i = 0; i < N; i++ )
      order[ pos[ lev[ i ] ]++ ] = (int32_t) i;
}

template< typename T >
void upload_vec( DeviceArray< T >& d, const std::vector< T >& h, hipStream_t s )
{
   d.allocate( std::max< int64_t >( 1, (int64_t) h.size() ) );
   if( ! h.empty() )
      TNLMI_HIP( hipMemcpyAsync( d.ptr, h.data(), sizeof( T ) * h.size(), hipMemcpyHostToDevice, s ) );
}

// ILU0.hpp:12-96 (IKJ on A's pattern) on the host, as the reference does;
// the triangular solves run on the GPU by levels.
void ilu0_build( tnlmi_preconditioner& P, const tnlmi_matrix& A, hipStream_t s )
{
   const tnlmi_segments& seg = *A.segments;
   const int64_t N = A.rows;
   DeviceArray< int32_t > dsz, doff;
   dsz.allocate( std::max< int64_t >( 1, N ) );
   doff.allocate( N + 1 );
   if( N )
      hipLaunchKernelGGL( seg_size_kernel, dim3( (unsigned) ceil_div( N, 256 ) ), dim3( 256 ), 0, s, make_view( seg ),
                          dsz.ptr );
   const int64_t nel = exclusive_scan_i32( dsz.ptr, doff.ptr, N, s );
   if( nel > INT32_MAX )
      fail( TNLMI_OVERFLOW_ERROR, "ILU0: the matrix has more than 2^31 - 1 slots" );
   DeviceArray< int32_t > dc;
   DeviceArray< double > dv;
   dc.allocate( std::max< int64_t >( 1, nel ) );
   dv.allocate( std::max< int64_t >( 1, nel ) );
   with_value_type( A.value_type, [ & ]( auto tag ) {
      using V = std::remove_const_t< std::remove_pointer_t< decltype( tag ) > >;
      if( N )
         hipLaunchKernelGGL( to_csr_kernel< V >, dim3( (unsigned) ceil_div( N, 256 ) ), dim3( 256 ), 0, s,
                             make_view( seg ), A.column_indexes, (const V*) A.values, doff.ptr, dc.ptr, dv.ptr );
   } );
   std::vector< int32_t > off( N + 1 ), col( nel );
   std::vector< double > val( nel );
   TNLMI_HIP( hipMemcpyAsync( off.data(), doff.ptr, 4 * ( N + 1 ), hipMemcpyDeviceToHost, s ) );
   if( nel ) {
      TNLMI_HIP( hipMemcpyAsync( col.data(), dc.ptr, 4 * nel, hipMemcpyDeviceToHost, s ) );
      TNLMI_HIP( hipMemcpyAsync( val.data(), dv.ptr, 8 * nel, hipMemcpyDeviceToHost, s ) );
   }
   TNLMI_HIP( hipStreamSynchronize( s ) );
   // compress each row (drop padding), keep the column order of the row
   std::vector< int32_t > rp( N + 1, 0 ), rc;
   std::vector< double > rv;
   rc.reserve( nel );
   rv.reserve( nel );
   for( int64_t i = 0; i < N; i++ ) {
      for( int32_t g = off[ i ]; g < off[ i + 1 ]; g++ )
         if( col[ g ] >= 0 ) {
            rc.push_back( col[ g ] );
            rv.push_back( val[ g ] );
         }
      rp[ i + 1 ] = (int32_t) rc.size();
   }
   std::vector< int32_t > Lr( N ), Ur( N );
   for( int64_t i = 0; i < N; i++ ) {
      int32_t le = 0, ue = 0;
      for( int32_t g = rp[ i ]; g < rp[ i + 1 ]; g++ ) {
         if( rc[ g ] < i )
            le++;
         else if( rc[ g ] < N )
            ue++;
         else
            break;
      }
      Lr[ i ] = le;
      Ur[ N - 1 - i ] = ue;
   }
   std::vector< int32_t > lp( N + 1, 0 ), up( N + 1, 0 );
   for( int64_t i = 0; i < N; i++ ) {
      lp[ i + 1 ] = lp[ i ] + Lr[ i ];
      up[ i + 1 ] = up[ i ] + Ur[ i ];
   }
   std::vector< int32_t > lc( lp[ N ] ), uc( up[ N ] );
   std::vector< double > lv( lp[ N ] ), uv( up[ N ] );
   auto uget = [ & ]( int64_t urow, int32_t c ) -> double {
      for( int32_t g = up[ urow ]; g < up[ urow + 1 ]; g++ )
         if( uc[ g ] == c )
            return uv[ g ];
      return 0.0;
   };
   for( int64_t i = 0; i < N; i++ ) {
      const int32_t le = Lr[ i ], ue = Ur[ N - 1 - i ];
      const int64_t ub = up[ N - 1 - i ], lb = lp[ i ], rb = rp[ i ];
      for( int32_t j = 0; j < ue; j++ ) {
         uc[ ub + j ] = rc[ rb + le + j ];
         uv[ ub + j ] = rv[ rb + le + j ];
      }
      for( int32_t j = 0; j < le; j++ ) {
         lc[ lb + j ] = rc[ rb + j ];
         lv[ lb + j ] = rv[ rb + j ];
      }
      for( int32_t ck = 0; ck < le; ck++ ) {
         const int32_t k = lc[ lb + ck ];
         const double lik = lv[ lb + ck ] / uget( N - 1 - k, k );
         lv[ lb + ck ] = lik;
         for( int32_t cj = ck + 1; cj < le; cj++ )
            lv[ lb + cj ] = lv[ lb + cj ] - lik * uget( N - 1 - k, lc[ lb + cj ] );
         for( int32_t cj = 0; cj < ue; cj++ )
            uv[ ub + cj ] = uv[ ub + cj ] - lik * uget( N - 1 - k, uc[ ub + cj ] );
      }
   }
   std::vector< int32_t > lo, uo;
   levels( N, lp, lc, false, lo, P.llev );
   levels( N, up, uc, true, uo, P.ulev );
   auto plan_runs = []( const std::vector< int64_t >& bounds, std::vector< tnlmi_preconditioner::Run >& runs ) {
      runs.clear();
      const int64_t L = (int64_t) bounds.size() - 1;
      for( int64_t l = 0; l < L; ) {
         const bool narrow = bounds[ l + 1 ] - bounds[ l ] <= (int64_t) kLvThreads * kLvRows;
         int64_t e = l + 1;
         if( narrow )
            while( e < L && e - l < kRunLevels && bounds[ e + 1 ] - bounds[ e ] <= (int64_t) kLvThreads * kLvRows )
               e++;
         runs.push_back( { l, e, narrow, false, false } );
         l = e;
      }
   };
   plan_runs( P.llev, P.lruns );
   plan_runs( P.ulev, P.uruns );
   // the factors in level-order positions, and which narrow runs may keep
   // their recent solution entries in the LDS ring
   auto positions = [ & ]( bool upper, const std::vector< int32_t >& order, const std::vector< int64_t >& bounds,
                           std::vector< tnlmi_preconditioner::Run >& runs, DeviceArray< int32_t >& dpp,
                           DeviceArray< int32_t >& ddep, DeviceArray< double >& dval,
                           DeviceArray< double >* ddiag, DeviceArray< int32_t >& dridx,
                           DeviceArray< double >& drval ) {
      const std::vector< int32_t >& ptr = upper ? up : lp;
      const std::vector< int32_t >& cc = upper ? uc : lc;
      const std::vector< double >& vv = upper ? uv : lv;
      std::vector< int32_t > pos( N );
      for( int64_t p = 0; p < N; p++ )
         pos[ order[ p ] ] = (int32_t) p;
      std::vector< int32_t > pp( N + 1, 0 ), dep;
      std::vector< double > val, diag;
      for( int64_t p = 0; p < N; p++ ) {
         const int64_t i = order[ p ];
         const int64_t k = upper ? N - 1 - i : i;
         int32_t g = ptr[ k ];
         if( upper ) {
            diag.push_back( g < ptr[ k + 1 ] ? vv[ g ] : 0.0 );  // a row without entries has no pivot
            g++;
         }
         for( ; g < ptr[ k + 1 ]; g++ ) {
            dep.push_back( pos[ cc[ g ] ] );
            val.push_back( vv[ g ] );
         }
         pp[ p + 1 ] = (int32_t) dep.size();
      }
      // the sync-free solve: rows of at most kRingPre entries, every
      // dependency at most kFlowDist positions before its row
      const char* fenv = std::getenv( "TNLMI_ILU_FLOW" );
      const bool flow_on = ! ( fenv && std::string( fenv ) == "0" );
      for( auto& R : runs ) {
         R.flow = flow_on;
         for( int64_t p = bounds[ R.l0 ]; p < bounds[ R.l1 ] && R.flow; p++ ) {
            if( pp[ p + 1 ] - pp[ p ] > kRingPre )
               R.flow = false;
            for( int32_t g = pp[ p ]; g < pp[ p + 1 ]; g++ )
               if( p - dep[ g ] > kFlowDist ) {
                  R.flow = false;
                  break;
               }
         }
      }
      for( auto& R : runs ) {
         R.ring = R.narrow;
         if( ! R.narrow )
            continue;
         for( int64_t l = R.l0; l < R.l1 && R.ring; l++ )
            R.ring = bounds[ l + 1 ] - bounds[ l ] <= kLvThreads;  // one row per thread
         for( int64_t l = R.l0; l < R.l1 && R.ring; l++ )
            for( int64_t p = bounds[ l ]; p < bounds[ l + 1 ] && R.ring; p++ ) {
               if( pp[ p + 1 ] - pp[ p ] > kRingPre )
                  R.ring = false;
               for( int32_t g = pp[ p ]; g < pp[ p + 1 ]; g++ )
                  if( bounds[ l + 1 ] - dep[ g ] > kRing ) {
                     R.ring = false;
                     break;
                  }
            }
      }
      // the ring solve's records: row, the entries (dependency position,
      // value; -1 / 0 past the row's end) and the pivot (1 for the lower factor)
      std::vector< int32_t > ridx( 4 * std::max< int64_t >( N, 1 ), -1 );
      std::vector< double > rval( 4 * std::max< int64_t >( N, 1 ), 0.0 );
      for( int64_t p = 0; p < N; p++ ) {
         ridx[ 4 * p ] = order[ p ];
         for( int q = 0; q < kRingPre; q++ )
            if( pp[ p ] + q < pp[ p + 1 ] ) {
               ridx[ 4 * p + 1 + q ] = dep[ pp[ p ] + q ];
               rval[ 4 * p + q ] = val[ pp[ p ] + q ];
            }
         rval[ 4 * p + kRingPre ] = upper ? diag[ p ] : 1.0;
      }
      upload_vec( dpp, pp, s );
      upload_vec( ddep, dep, s );
      upload_vec( dval, val, s );
      if( ddiag )
         upload_vec( *ddiag, diag, s );
      upload_vec( dridx, ridx, s );
      upload_vec( drval, rval, s );
   };
   positions( false, lo, P.llev, P.lruns, P.lpp, P.ldep, P.lval, nullptr, P.lridx, P.lrval );
   positions( true, uo, P.ulev, P.uruns, P.upp, P.udep, P.uval, &P.udiag, P.uridx, P.urval );
   P.bpos.allocate( std::max< int64_t >( N, 1 ) );
   P.xpos.allocate( std::max< int64_t >( N, 1 ) + kLvThreads );
   upload_vec( P.dllev, P.llev, s );
   upload_vec( P.dulev, P.ulev, s );
   upload_vec( P.lp, lp, s );
   upload_vec( P.lc, lc, s );
   upload_vec( P.lv, lv, s );
   upload_vec( P.up, up, s );
   upload_vec( P.uc, uc, s );
   upload_vec( P.uv, uv, s );
   upload_vec( P.lrows, lo, s );
   upload_vec( P.urows, uo, s );
   TNLMI_HIP( hipStreamSynchronize( s ) );
}

}  // namespace

void precond_apply( const tnlmi_preconditioner* P, int64_t n, const double* b, double* x, hipStream_t s )
{
   if( ! P || P->kind == TNLMI_PRECOND_NONE ) {
      copy( n, b, x, s );
      return;
   }
   if( P->kind == TNLMI_PRECOND_DIAGONAL ) {
      divide( n, b, P->diag.ptr, x, s );  // Diagonal.hpp:38-43: x = b / diagonal
      return;
   }
   if( P->kind == TNLMI_PRECOND_EXPLICIT ) {
      const tnlmi_launch cfg{ TNLMI_MAP_AUTO, 1 };
      spmv_dispatch( P->M, TNLMI_F64, b, x, TNLMI_F64, 1.0, 0.0, 0, n, cfg, s );
      return;
   }
   // ILU0.hpp:99-113: triangularSolveLower then triangularSolveUpper<reversed rows>
   auto any_ring = []( const std::vector< tnlmi_preconditioner::Run >& runs ) {
      return std::any_of( runs.begin(), runs.end(),
                          []( const tnlmi_preconditioner::Run& R ) { return R.ring || R.flow; } );
   };
   int64_t spin_cap = kFlowSpinCap;
   if( const char* e = std::getenv( "TNLMI_ILU_SPIN_CAP" ) )
      spin_cap = std::max< int64_t >( 1, std::atoll( e ) );
   const dim3 gpos( (unsigned) std::max< int64_t >( 1, ceil_div( n, 256 ) ) );
   // the ring and sync-free solves leave their results in position order in xpos
   auto scatter_run = [ & ]( const int32_t* rows, int64_t p0, int64_t p1 ) {
      if( p1 > p0 )
         hipLaunchKernelGGL( scatter_pos, dim3( (unsigned) ceil_div( p1 - p0, 256 ) ), dim3( 256 ), 0, s, p0, p1,
                             rows, (const double*) P->xpos.ptr, x );
   };
   if( any_ring( P->lruns ) )
      hipLaunchKernelGGL( gather_pos, gpos, dim3( 256 ), 0, s, n, P->lrows.ptr, b, P->bpos.ptr );
   for( const auto& R : P->lruns ) {
      if( R.flow ) {
         const int64_t p0 = P->llev[ R.l0 ], p1 = P->llev[ R.l1 ];
         hipLaunchKernelGGL( tri_flow< false >, dim3( 1 ), dim3( kLvThreads ), 0, s, P->lrows.ptr, p0, p1,
                             (const RingIdx*) P->lridx.ptr, (const RingVal*) P->lrval.ptr, P->bpos.ptr,
                             (const double*) x, P->xpos.ptr, n, spin_cap );
         scatter_run( P->lrows.ptr, p0, p1 );
         continue;
      }
      if( R.ring ) {
         hipLaunchKernelGGL( tri_ring< false >, dim3( 1 ), dim3( kLvThreads ), 0, s, P->lrows.ptr, P->dllev.ptr, R.l0,
                             R.l1, (const RingIdx*) P->lridx.ptr, (const RingVal*) P->lrval.ptr, P->bpos.ptr,
                             (const double*) x, P->xpos.ptr, n );
         scatter_run( P->lrows.ptr, P->llev[ R.l0 ], P->llev[ R.l1 ] );
         continue;
      }
      if( R.narrow ) {
         hipLaunchKernelGGL( tri_levels< false >, dim3( 1 ), dim3( kLvThreads ), 0, s, P->lrows.ptr, P->dllev.ptr,
                             R.l0, R.l1, n, P->lp.ptr, P->lc.ptr, P->lv.ptr, b, x );
         continue;
      }
      const int64_t cnt = P->llev[ R.l0 + 1 ] - P->llev[ R.l0 ];
      hipLaunchKernelGGL( lower_level, dim3( (unsigned) ceil_div( cnt, 256 ) ), dim3( 256 ), 0, s,
                          P->lrows.ptr + P->llev[ R.l0 ], cnt, P->lp.ptr, P->lc.ptr, P->lv.ptr, b, x );
   }
   // the upper solve's right-hand side is the lower solve's result
   if( any_ring( P->uruns ) )
      hipLaunchKernelGGL( gather_pos, gpos, dim3( 256 ), 0, s, n, P->urows.ptr, (const double*) x, P->bpos.ptr );
   for( const auto& R : P->uruns ) {
      if( R.flow ) {
         const int64_t p0 = P->ulev[ R.l0 ], p1 = P->ulev[ R.l1 ];
         hipLaunchKernelGGL( tri_flow< true >, dim3( 1 ), dim3( kLvThreads ), 0, s, P->urows.ptr, p0, p1,
                             (const RingIdx*) P->uridx.ptr, (const RingVal*) P->urval.ptr, P->bpos.ptr,
                             (const double*) x, P->xpos.ptr, n, spin_cap );
         scatter_run( P->urows.ptr, p0, p1 );
         continue;
      }
      if( R.ring ) {
         hipLaunchKernelGGL( tri_ring< true >, dim3( 1 ), dim3( kLvThreads ), 0, s, P->urows.ptr, P->dulev.ptr, R.l0,
                             R.l1, (const RingIdx*) P->uridx.ptr, (const RingVal*) P->urval.ptr, P->bpos.ptr,
                             (const double*) x, P->xpos.ptr, n );
         scatter_run( P->urows.ptr, P->ulev[ R.l0 ], P->ulev[ R.l1 ] );
         continue;
      }
      if( R.narrow ) {
         hipLaunchKernelGGL( tri_levels< true >, dim3( 1 ), dim3( kLvThreads ), 0, s, P->urows.ptr, P->dulev.ptr,
                             R.l0, R.l1, n, P->up.ptr, P->uc.ptr, P->uv.ptr, (const double*) x, x );
         continue;
      }
      const int64_t cnt = P->ulev[ R.l0 + 1 ] - P->ulev[ R.l0 ];
      hipLaunchKernelGGL( upper_level, dim3( (unsigned) ceil_div( cnt, 256 ) ), dim3( 256 ), 0, s,
                          P->urows.ptr + P->ulev[ R.l0 ], cnt, n, P->up.ptr, P->uc.ptr, P->uv.ptr, x );
   }
}

tnlmi_precond_kind precond_kind( const tnlmi_preconditioner* P )
{
   return P ? P->kind : TNLMI_PRECOND_NONE;
}

const double* precond_diagonal( const tnlmi_preconditioner* P )
{
   return P && P->kind == TNLMI_PRECOND_DIAGONAL ? P->diag.ptr : nullptr;
}

// y = A x (FP64 vectors and accumulation) with the matrix or its plan
void Operator::apply( const double* x, double* y, hipStream_t s ) const
{
   if( plan )
      plan_execute( *plan, TNLMI_F64, x, y, TNLMI_F64, 1.0, 0.0, 0, rows, s );
   else {
      const tnlmi_launch cfg{ TNLMI_MAP_AUTO, 1 };
      spmv_dispatch( *A, TNLMI_F64, x, y, TNLMI_F64, 1.0, 0.0, 0, rows, cfg, s );
   }
}

namespace {

// ---- CG kernels (CG.hpp:60-105), scalars in sc[]:
// 0 s1, 1 s2, 2 normb, 3 residue, 4 alpha, 5 beta, 6 s1_new
__global__ void cg_alpha( double* sc )
{
   // alpha = s1 / s2 (s2 == 0 is caught by the host before the update)
   sc[ 4 ] = sc[ 1 ] != 0.0 ? sc[ 0 ] / sc[ 1 ] : 0.0;
}

__global__ void cg_update( int64_t n, const double* __restrict__ p, const double* __restrict__ Ap,
                           double* __restrict__ x, double* __restrict__ r, const double* __restrict__ sc )
{
   const double a = sc[ 4 ];
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x ) {
      x[ i ] += a * p[ i ];
      r[ i ] -= a * Ap[ i ];
   }
}

__global__ void cg_beta( double* sc )
{
   // s2 = s1; s1 = (r, z); beta = s1 / s2; residue = sqrt(s1) / normb
   const double s2 = sc[ 0 ], s1 = sc[ 6 ];
   sc[ 1 ] = s2;
   sc[ 0 ] = s1;
   sc[ 5 ] = s2 == 0.0 ? 0.0 : s1 / s2;
   sc[ 3 ] = sqrt( s1 ) / sc[ 2 ];
}

__global__ void cg_dir( int64_t n, const double* __restrict__ z, double* __restrict__ p, const double* __restrict__ sc )
{
   const double b = sc[ 5 ];
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x )
      p[ i ] = z[ i ] + b * p[ i ];
}

__global__ void cg_init( double* sc )
{
   sc[ 3 ] = sqrt( sc[ 0 ] ) / sc[ 2 ];
}

inline dim3 vgrid( int64_t n )
{
   return dim3( (unsigned) std::max< int64_t >( 1, std::min< int64_t >( 8192, ceil_div( n, 256 ) ) ) );
}

void cg_solve( const Operator& A, const tnlmi_preconditioner* M, const double* b, double* x, tnlmi_solver_state* st,
               hipStream_t s )
{
   if( precond_kind( M ) == TNLMI_PRECOND_NONE || precond_kind( M ) == TNLMI_PRECOND_DIAGONAL ) {
      fused_cg_solve( A, M, b, x, st, s );
      return;
   }
   const int64_t n = A.rows;
   DeviceArray< double > r, p, Ap, z, sc;
   r.allocate( n );
   p.allocate( n );
   Ap.allocate( n );
   z.allocate( n );
   sc.allocate( 8 );
   RedWork w;
   HostScalars hs;
   Stop S{ st };
   const bool pre = M && M->kind != TNLMI_PRECOND_NONE;
   // normb = ||M^{-1} b|| or ||b||  (CG.hpp:24-29)
   if( pre ) {
      precond_apply( M, n, b, r.ptr, s );
      nrm2( n, r.ptr, sc.ptr + 2, w, s );
   }
   else
      nrm2( n, b, sc.ptr + 2, w, s );
   if( hs.read( sc.ptr + 2, 1, s )[ 0 ] == 0.0 ) {
      fill( n, 0.0, x, s );
      st->iterations = 0;
      st->residue = 0.0;
      st->converged = 1;
      return;
   }
   // r = b - A x; z = M^{-1} r; p = z; s1 = (r, z)
   A.apply( x, r.ptr, s );
   rsub( n, b, r.ptr, s );
   const double* zz = r.ptr;
   if( pre ) {
      precond_apply( M, n, r.ptr, z.ptr, s );
      zz = z.ptr;
   }
   copy( n, zz, p.ptr, s );
   dot( n, r.ptr, zz, sc.ptr + 0, w, s );
   hipLaunchKernelGGL( cg_init, dim3( 1 ), dim3( 1 ), 0, s, sc.ptr );
   S.res = hs.read( sc.ptr + 3, 1, s )[ 0 ];
   while( S.next() ) {
      A.apply( p.ptr, Ap.ptr, s );
      dot( n, Ap.ptr, p.ptr, sc.ptr + 1, w, s );
      // s2 == 0 => p = 0 => r = 0: the solution is found (CG.hpp:66-70)
      if( hs.read( sc.ptr + 1, 1, s )[ 0 ] == 0.0 ) {
         S.res = 0.0;
         break;
      }
      hipLaunchKernelGGL( cg_alpha, dim3( 1 ), dim3( 1 ), 0, s, sc.ptr );
      hipLaunchKernelGGL( cg_update, vgrid( n ), dim3( 256 ), 0, s, n, p.ptr, Ap.ptr, x, r.ptr, sc.ptr );
      if( pre )
         precond_apply( M, n, r.ptr, z.ptr, s );
      dot( n, r.ptr, zz, sc.ptr + 6, w, s );
      hipLaunchKernelGGL( cg_beta, dim3( 1 ), dim3( 1 ), 0, s, sc.ptr );
      hipLaunchKernelGGL( cg_dir, vgrid( n ), dim3( 256 ), 0, s, n, zz, p.ptr, sc.ptr );
      S.res = hs.read( sc.ptr + 3, 1, s )[ 0 ];
   }
   st->iterations = S.it;
   st->residue = S.res;
   st->converged = S.converged();
}

// ---- Jacobi (Jacobi.hpp:65-125)
__global__ void jacobi_step( int64_t n, const double* __restrict__ Ain, const double* __restrict__ in,
                             const double* __restrict__ b, const double* __restrict__ d, double omega,
                             double* __restrict__ out )
{
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x )
      out[ i ] = in[ i ] + omega / d[ i ] * ( b[ i ] - Ain[ i ] );
}

__global__ void diff_k( int64_t n, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ o )
{
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x )
      o[ i ] = a[ i ] - b[ i ];
}

__global__ void ratio_k( const double* a, const double* b, double* o )
{
   o[ 0 ] = a[ 0 ] / b[ 0 ];
}

// ||A x - b|| / bnorm into d_out[0] (LinearResidueGetter.hpp:11-23)
void residue_dev( const Operator& A, const double* x, const double* b, const double* d_bnorm, double* tmp,
                  double* d_out, RedWork& w, hipStream_t s )
{
   const int64_t n = A.rows;
   A.apply( x, tmp, s );
   hipLaunchKernelGGL( diff_k, vgrid( n ), dim3( 256 ), 0, s, n, tmp, b, tmp );
   nrm2( n, tmp, d_out, w, s, 1 );
   hipLaunchKernelGGL( ratio_k, dim3( 1 ), dim3( 1 ), 0, s, d_out, d_bnorm, d_out );
}

void jacobi_solve( const Operator& A, const tnlmi_segments& seg, const tnlmi_matrix& M, const double* b, double* x,
                   double omega, int64_t period, tnlmi_solver_state* st, hipStream_t s )
{
   const int64_t n = A.rows;
   DeviceArray< double > d, aux, tmp, sc;
   d.allocate( n );
   aux.allocate( n );
   tmp.allocate( n );
   sc.allocate( 4 );
   RedWork w;
   HostScalars hs;
   with_value_type( M.value_type, [ & ]( auto tag ) {
      using V = std::remove_const_t< std::remove_pointer_t< decltype( tag ) > >;
      hipLaunchKernelGGL( diag_kernel< V >, dim3( (unsigned) ceil_div( n, 256 ) ), dim3( 256 ), 0, s,
                          make_view( seg ), M.column_indexes, (const V*) M.values, d.ptr );
   } );
   Stop S{ st };
   S.res = st->convergence_residue + 1.0;
   nrm2( n, b, sc.ptr + 0, w, s );
   copy( n, x, aux.ptr, s );
   auto step = [ & ]( const double* in, double* out ) {
      A.apply( in, tmp.ptr, s );
      hipLaunchKernelGGL( jacobi_step, vgrid( n ), dim3( 256 ), 0, s, n, tmp.ptr, in, b, d.ptr, omega, out );
   };
   auto residue = [ & ]() {
      residue_dev( A, x, b, sc.ptr + 0, tmp.ptr, sc.ptr + 1, w, s );
      return hs.read( sc.ptr + 1, 1, s )[ 0 ];
   };
   while( S.next() ) {
      step( x, aux.ptr );
      if( S.it % period == 0 )
         S.res = residue();
      S.it++;
      step( aux.ptr, x );
      if( S.it % period == 0 )
         S.res = residue();
   }
   S.res = residue();
   st->iterations = S.it;
   st->residue = S.res;
   st->converged = S.converged();
}

// ---- GMRES (GMRES.hpp:71-431, 645-770) with CGS/CGSR/MGS/MGSR orthogonalisation.
// Device layout of the small FP64 state g[]:
//   H: (m+1)*m at 0, s: m+1, cs: m+1, sn: m+1, y: m+1, tmp: 32, scalars: 8
struct GLayout
{
   int m;
   int64_t H, s, cs, sn, y, tmp, sc, total;
   explicit GLayout( int m_ ) : m( m_ )
   {
      H = 0;
      s = H + ( m + 1 ) * m;
      cs = s + m + 1;
      sn = cs + m + 1;
      y = sn + m + 1;
      tmp = y + m + 1;
      sc = tmp + 32;
      total = sc + 8;  // sc: 0 normb, 1 beta, 2 normw, 3 residue
   }
};

__global__ void gm_start( double* g, GLayout L, int m )
{
   // H = 0, s = 0, s[0] = beta
   for( int64_t i = threadIdx.x; i < ( m + 1 ) * m; i += blockDim.x )
      g[ L.H + i ] = 0.0;
   for( int i = threadIdx.x; i <= m; i += blockDim.x )
      g[ L.s + i ] = 0.0;
   __syncthreads();
   if( threadIdx.x == 0 )
      g[ L.s ] = g[ L.sc + 1 ];
}

// H[k, i] += h[k] for k <= i (one reorthogonalisation pass of CGS)
__global__ void gm_hadd( double* g, GLayout L, int i, int m, const double* h )
{
   const int k = threadIdx.x;
   if( k <= i )
      g[ L.H + k + i * ( m + 1 ) ] += h[ k ];
}

// MGS step: h = (w, v_k) already in tmp[0]: H[k, i] += h
__global__ void gm_hadd1( double* g, GLayout L, int k, int i, int m )
{
   g[ L.H + k + i * ( m + 1 ) ] += g[ L.tmp ];
}

// w -= h * v_k with h = tmp[0]
__global__ void gm_axpy( int64_t n, const double* __restrict__ vk, double* __restrict__ w, const double* __restrict__ g,
                         GLayout L )
{
   const double h = g[ L.tmp ];
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x )
      w[ i ] = w[ i ] - h * vk[ i ];
}

__device__ void gen_rot( double dx, double dy, double& c, double& sn )
{
   // generatePlaneRotation, GMRES.hpp:729-746
   if( dy == 0.0 ) {
      c = 1.0;
      sn = 0.0;
   }
   else if( fabs( dy ) > fabs( dx ) ) {
      const double t = dx / dy;
      sn = 1.0 / sqrt( 1.0 + t * t );
      c = t * sn;
   }
   else {
      const double t = dy / dx;
      c = 1.0 / sqrt( 1.0 + t * t );
      sn = t * c;
   }
}
__device__ void app_rot( double& dx, double& dy, double c, double sn )
{
   const double t = c * dx + sn * dy;
   dy = c * dy - sn * dx;
   dx = t;
}

// H[i+1, i] = ||w|| (already in sc[2]); Givens rotations on column i;
// residue = |s[i+1]| / normb   (GMRES.hpp:327-341, 758-770)
__global__ void gm_givens( double* g, GLayout L, int i, int m )
{
   double* H = g + L.H;
   double* s = g + L.s;
   double* cs = g + L.cs;
   double* sn = g + L.sn;
   H[ i + 1 + i * ( m + 1 ) ] = g[ L.sc + 2 ];
   for( int k = 0; k < i; k++ )
      app_rot( H[ k + i * ( m + 1 ) ], H[ k + 1 + i * ( m + 1 ) ], cs[ k ], sn[ k ] );
   if( H[ i + 1 + i * ( m + 1 ) ] != 0.0 ) {
      gen_rot( H[ i + i * ( m + 1 ) ], H[ i + 1 + i * ( m + 1 ) ], cs[ i ], sn[ i ] );
      app_rot( H[ i + i * ( m + 1 ) ], H[ i + 1 + i * ( m + 1 ) ], cs[ i ], sn[ i ] );
      app_rot( s[ i ], s[ i + 1 ], cs[ i ], sn[ i ] );
   }
   g[ L.sc + 3 ] = fabs( s[ i + 1 ] ) / g[ L.sc + 0 ];
}

// y = H^{-1} s on the leading (k+1) x (k+1) block (update, GMRES.hpp:645-662);
// flag sc[4] = 1 when a zero pivot is met (the reference throws)
__global__ void gm_backsub( double* g, GLayout L, int k, int m )
{
   const double* H = g + L.H;
   double* y = g + L.y;
   for( int i = 0; i <= m; i++ )
      y[ i ] = g[ L.s + i ];
   g[ L.sc + 4 ] = 0.0;
   for( int i = k; i >= 0; i-- ) {
      if( H[ i + i * ( m + 1 ) ] == 0.0 ) {
         g[ L.sc + 4 ] = 1.0;
         return;
      }
      y[ i ] /= H[ i + i * ( m + 1 ) ];
      for( int j = i - 1; j >= 0; j-- )
         y[ j ] -= H[ j + i * ( m + 1 ) ] * y[ i ];
   }
}

__global__ void gm_set_scalar( double* dst, const double* src )
{
   dst[ 0 ] = src[ 0 ];
}

__global__ void gm_residue_beta( double* g, GLayout L )
{
   g[ L.sc + 3 ] = g[ L.sc + 1 ] / g[ L.sc + 0 ];
}

// v = w / ||w||: scale by 1 / sc[2] (the reference multiplies by (1.0 / normw))
__global__ void gm_normalize( int64_t n, const double* __restrict__ w, double* __restrict__ v,
                              const double* __restrict__ g, int64_t idx )
{
   const double inv = 1.0 / g[ idx ];
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x )
      v[ i ] = inv * w[ i ];
}

// ---- CWY: compact-WY Householder orthogonalisation (GMRES.hpp:345-431, 467-643)
// Y: the Householder vectors (columns of length ld), T: the upper-triangular
// compact-WY factor (R x R, R = rmax + 1), YL: the first R entries of each
// Householder vector; the small arrays sit in one device buffer cw.
struct CwyLayout
{
   int R;
   int64_t T, YL, aux, total;
   explicit CwyLayout( int rmax ) : R( rmax + 1 )
   {
      T = 0;
      YL = T + (int64_t) R * R;
      aux = YL + (int64_t) R * R;
      total = aux + 64;
   }
};

// y_i = z with the entries above i zeroed (hauseholder_generate, :467-478)
__global__ void cwy_mask( int64_t n, const double* __restrict__ z, double* __restrict__ y, int i )
{
   for( int64_t j = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t) gridDim.x * blockDim.x )
      y[ j ] = j < i ? 0.0 : z[ j ];
}

// y_i[i] += sign(y_ii) ||y_i||;  t_i = 2 / ||y_i||^2 (new y_i);  T[i, i] = t_i  (:480-492)
__global__ void cwy_gen( double* __restrict__ y, int i, const double* __restrict__ normz, double* __restrict__ cw,
                         CwyLayout C )
{
   const double nz = normz[ 0 ];
   const double yii = y[ i ];
   y[ i ] = yii > 0.0 ? yii + nz : yii - nz;
   const double nys = 2.0 * ( nz * nz + fabs( yii ) * nz );
   cw[ C.T + i + (int64_t) i * C.R ] = 2.0 / nys;
}

// T[k, i] = -sum_{j=k}^{i-1} T[k, j] t_i aux[j], aux[j] = (y_j, y_i)  (:494-510)
__global__ void cwy_tcol( double* __restrict__ cw, CwyLayout C, int i, const double* __restrict__ aux )
{
   const int k = threadIdx.x;
   if( k >= i )
      return;
   const double ti = cw[ C.T + i + (int64_t) i * C.R ];
   double t = 0.0;
   for( int j = k; j < i; j++ )
      t -= cw[ C.T + k + (int64_t) j * C.R ] * ( ti * aux[ j ] );
   cw[ C.T + k + (int64_t) i * C.R ] = t;
}

// YL[i] = the first R entries of y_i;  out[k] = z[k] - y_i[k], k <= i (hauseholder_apply_trunc, :512-538)
__global__ void cwy_trunc( int64_t n, const double* __restrict__ y, int i, const double* __restrict__ z,
                           double* __restrict__ cw, CwyLayout C, double* __restrict__ out )
{
   if( threadIdx.x != 0 )
      return;
   for( int k = 0; k < C.R && k < n; k++ )
      cw[ C.YL + (int64_t) i * C.R + k ] = y[ k ];
   for( int k = 0; k <= i; k++ )
      out[ k ] = z[ k ] - cw[ C.YL + (int64_t) i * C.R + k ];
}

// aux = T (YL[k][i])_k  (upper-triangular product, hauseholder_cwy :540-563)
__global__ void cwy_aux_upper( double* __restrict__ cw, CwyLayout C, int i )
{
   if( threadIdx.x != 0 )
      return;
   double* aux = cw + C.aux;
   for( int k = 0; k <= i; k++ )
      aux[ k ] = cw[ C.YL + i + (int64_t) k * C.R ];
   for( int k = 0; k <= i; k++ ) {
      double a2 = 0.0;
      for( int j = k; j <= i; j++ )
         a2 += cw[ C.T + k + (int64_t) j * C.R ] * aux[ j ];
      aux[ k ] = a2;
   }
}

// aux = T^T aux  (lower-triangular product, hauseholder_cwy_transposed :565-592)
__global__ void cwy_aux_lower( double* __restrict__ cw, CwyLayout C, int i )
{
   if( threadIdx.x != 0 )
      return;
   double* aux = cw + C.aux;
   for( int k = i; k >= 0; k-- ) {
      double a2 = 0.0;
      for( int j = 0; j <= k; j++ )
         a2 += cw[ C.T + j + (int64_t) k * C.R ] * aux[ j ];
      aux[ k ] = a2;
   }
}

// v[i] += 1
__global__ void cwy_unit( double* v, int i )
{
   v[ i ] += 1.0;
}

// apply_givens_rotations to column i of H (filled by cwy_trunc) and s;
// residue = |s[i + 1]| / normb  (:758-770)
__global__ void cwy_givens( double* g, GLayout L, int i, int m )
{
   double* H = g + L.H;
   double* s = g + L.s;
   double* cs = g + L.cs;
   double* sn = g + L.sn;
   for( int k = 0; k < i; k++ )
      app_rot( H[ k + i * ( m + 1 ) ], H[ k + 1 + i * ( m + 1 ) ], cs[ k ], sn[ k ] );
   if( H[ i + 1 + i * ( m + 1 ) ] != 0.0 ) {
      gen_rot( H[ i + i * ( m + 1 ) ], H[ i + 1 + i * ( m + 1 ) ], cs[ i ], sn[ i ] );
      app_rot( H[ i + i * ( m + 1 ) ], H[ i + 1 + i * ( m + 1 ) ], cs[ i ], sn[ i ] );
      app_rot( s[ i ], s[ i + 1 ], cs[ i ], sn[ i ] );
   }
   g[ L.sc + 3 ] = fabs( s[ i + 1 ] ) / g[ L.sc + 0 ];
}

// residue = |s[0]| / normb (the first Householder step)
__global__ void cwy_res0( double* g, GLayout L )
{
   g[ L.sc + 3 ] = fabs( g[ L.s ] ) / g[ L.sc + 0 ];
}

// update (:684-727): aux[i] = sum_{j=i}^{k} y[j] YL[i][j], aux = T aux (upper)
__global__ void cwy_update_aux( double* __restrict__ cw, CwyLayout C, const double* __restrict__ g, GLayout L, int k )
{
   if( threadIdx.x != 0 )
      return;
   const double* y = g + L.y;
   double* aux = cw + C.aux;
   for( int i = 0; i <= k; i++ )
      aux[ i ] = 0.0;
   for( int j = 0; j <= k; j++ )
      for( int i = 0; i <= j; i++ )
         aux[ i ] += y[ j ] * cw[ C.YL + j + (int64_t) i * C.R ];
   for( int i = 0; i <= k; i++ ) {
      double a2 = 0.0;
      for( int j = i; j <= k; j++ )
         a2 += cw[ C.T + i + (int64_t) j * C.R ] * aux[ j ];
      aux[ i ] = a2;
   }
}

// x[j] += y[j], j <= k
__global__ void cwy_add_y( double* __restrict__ x, const double* __restrict__ g, GLayout L, int k )
{
   const int j = threadIdx.x;
   if( j <= k )
      x[ j ] += g[ L.y + j ];
}

// GMRES.hpp:71-806.  inner == NULL: TNL's GMRES.  inner != NULL (the mixed-
// precision variant): the Arnoldi products use *inner (the matrix with FP32 or
// BF16 values) while every restart residual M^-1 (b - A x) uses A in FP64 --
// iterative refinement with one GMRES cycle per correction; the orthogonal
// basis, the Hessenberg least squares and x stay FP64.  When a cycle's
// estimate meets the stopping rule, the FP64 residue of the updated x decides:
// the solve stops only if that one meets it too.
void gmres_solve( const Operator& A, const tnlmi_preconditioner* M, const double* b, double* x, int variant, int rmin,
                  int rmax, int dmin, int dmax, tnlmi_solver_state* st, hipStream_t s, const Operator* inner )
{
   if( rmin <= 0 || rmax <= 0 || rmin > rmax || dmin < 0 || dmax < 0 || dmin > dmax ) {
      st->converged = 0;
      st->iterations = 0;
      return;  // the reference prints and returns false (GMRES.hpp:76-86)
   }
   if( rmax > 31 )
      fail( TNLMI_NOT_IMPLEMENTED, "GMRES: restarting-max above 31 is not supported by the GPU orthogonalisation" );
   const int64_t n = A.rows;
   const int64_t ld = ( n + 31 ) / 32 * 32;  // columns of V aligned to 256 bytes (GMRES.hpp:775-778)
   const bool pre = M && M->kind != TNLMI_PRECOND_NONE;
   GLayout L( rmax );
   DeviceArray< double > V, r, wv, tmp, g;
   V.allocate( ld * ( rmax + 1 ) );
   r.allocate( n );
   wv.allocate( n );
   tmp.allocate( n );
   g.allocate( L.total );
   RedWork rw;
   HostScalars hs;
   Stop S{ st };
   double* G = g.ptr;
   const Operator& Ak = inner ? *inner : A;  // the Arnoldi operator
   auto matvec_prec = [ & ]( const double* v, double* out ) {  // preconditioned_matvec, :451-465
      if( pre ) {
         Ak.apply( v, tmp.ptr, s );
         precond_apply( M, n, tmp.ptr, out, s );
      }
      else
         Ak.apply( v, out, s );
   };
   auto compute_residue = [ & ]() {  // compute_residue, :433-449
      if( pre ) {
         A.apply( x, tmp.ptr, s );
         rsub( n, b, tmp.ptr, s );
         precond_apply( M, n, tmp.ptr, r.ptr, s );
      }
      else {
         A.apply( x, r.ptr, s );
         rsub( n, b, r.ptr, s );
      }
   };
   // normb
   if( pre ) {
      precond_apply( M, n, b, tmp.ptr, s );
      nrm2( n, tmp.ptr, G + L.sc + 0, rw, s );
   }
   else
      nrm2( n, b, G + L.sc + 0, rw, s );
   if( hs.read( G + L.sc, 1, s )[ 0 ] == 0.0 ) {
      fill( n, 0.0, x, s );
      st->iterations = 0;
      st->residue = 0;
      st->converged = 1;
      return;
   }
   compute_residue();
   nrm2( n, r.ptr, G + L.sc + 1, rw, s );
   hipLaunchKernelGGL( gm_residue_beta, dim3( 1 ), dim3( 1 ), 0, s, G, L );
   S.it = 0;
   const double* h0 = hs.read( G + L.sc, 4, s );
   double beta = h0[ 1 ];
   S.res = h0[ 3 ];
   double beta_ratio = 1;
   int cycles = 0;
   int m = rmax;
   const bool classical = variant <= 1;
   const int passes = ( variant == 1 || variant == 3 ) ? 2 : 1;
   const bool cwy = variant == 4;
   const CwyLayout C( rmax );
   DeviceArray< double > Y, z, cwbuf;
   if( cwy ) {
      Y.allocate( ld * ( rmax + 1 ) );
      z.allocate( n );
      cwbuf.allocate( C.total );
   }
   double* cw = cwbuf.ptr;
   auto update = [ & ]( int k ) {
      hipLaunchKernelGGL( gm_backsub, dim3( 1 ), dim3( 1 ), 0, s, G, L, k, m );
      if( hs.read( G + L.sc + 4, 1, s )[ 0 ] != 0.0 )
         fail( TNLMI_RUNTIME_ERROR, "GMRES: zero on the diagonal of the Hessenberg matrix" );
      if( ! cwy )
         gemv_add( n, k + 1, V.ptr, ld, G + L.y, x, s );
      else {  // x += (I - Y T Y^T) applied to y (:684-727)
         hipLaunchKernelGGL( cwy_update_aux, dim3( 1 ), dim3( 64 ), 0, s, cw, C, G, L, k );
         gemv_sub( n, k + 1, Y.ptr, ld, cw + C.aux, x, s );
         hipLaunchKernelGGL( cwy_add_y, dim3( 1 ), dim3( 64 ), 0, s, x, G, L, k );
      }
   };
   while( S.check() ) {
      if( rmax > rmin && cycles > 0 ) {  // Baker-Jessup-Kolev adaptive restart, :121-146
         if( beta_ratio > 0.99 )
            m = rmax;
         else if( beta_ratio >= 0.175 ) {
            const int step = dmin + ( dmax - dmin ) / static_cast< float >( rmax - rmin ) * ( m - rmin );
            if( m - step >= rmin )
               m -= step;
            else
               m = rmax;
         }
      }
      int o = m;
      if( cwy ) {
         // z = r; H = 0; s = 0; T = 0   (orthogonalize_CWY, :345-360)
         copy( n, r.ptr, z.ptr, s );
         hipLaunchKernelGGL( gm_start, dim3( 1 ), dim3( 256 ), 0, s, G, L, m );
         fill( (int64_t) C.R * C.R, 0.0, cw + C.T, s );
         for( int i = 0; i <= m && S.next(); i++ ) {
            double* yi = Y.ptr + i * ld;
            const unsigned vg = (unsigned) std::min< int64_t >( 8192, ceil_div( n, 256 ) );
            hipLaunchKernelGGL( cwy_mask, dim3( vg ), dim3( 256 ), 0, s, n, z.ptr, yi, i );
            nrm2( n, yi, G + L.sc + 5, rw, s );
            hipLaunchKernelGGL( cwy_gen, dim3( 1 ), dim3( 1 ), 0, s, yi, i, G + L.sc + 5, cw, C );
            if( i > 0 ) {
               multi_dot( n, i, Y.ptr, ld, yi, G + L.tmp, rw, s );
               hipLaunchKernelGGL( cwy_tcol, dim3( 1 ), dim3( 64 ), 0, s, cw, C, i, G + L.tmp );
            }
            hipLaunchKernelGGL( cwy_trunc, dim3( 1 ), dim3( 64 ), 0, s, n, yi, i, z.ptr, cw, C,
                                i == 0 ? G + L.s : G + L.H + (int64_t) ( i - 1 ) * ( m + 1 ) );
            // v = (I - Y T Y^T) e_i
            hipLaunchKernelGGL( cwy_aux_upper, dim3( 1 ), dim3( 64 ), 0, s, cw, C, i );
            fill( n, 0.0, V.ptr, s );
            gemv_sub( n, i + 1, Y.ptr, ld, cw + C.aux, V.ptr, s );
            hipLaunchKernelGGL( cwy_unit, dim3( 1 ), dim3( 1 ), 0, s, V.ptr, i );
            if( i < m ) {
               // z = (I - Y T^T Y^T) M^{-1} A v
               matvec_prec( V.ptr, wv.ptr );
               multi_dot( n, i + 1, Y.ptr, ld, wv.ptr, cw + C.aux, rw, s );
               hipLaunchKernelGGL( cwy_aux_lower, dim3( 1 ), dim3( 64 ), 0, s, cw, C, i );
               copy( n, wv.ptr, z.ptr, s );
               gemv_sub( n, i + 1, Y.ptr, ld, cw + C.aux, z.ptr, s );
            }
            if( i > 0 )
               hipLaunchKernelGGL( cwy_givens, dim3( 1 ), dim3( 1 ), 0, s, G, L, i - 1, m );
            else
               hipLaunchKernelGGL( cwy_res0, dim3( 1 ), dim3( 1 ), 0, s, G, L );
            S.res = hs.read( G + L.sc + 3, 1, s )[ 0 ];
            if( i > 0 && ! S.check() ) {
               o = i - 1;
               break;
            }
         }
      }
      else {
      // v_0 = r / beta; H = 0; s = 0; s[0] = beta   (:190-202, :268-280)
      hipLaunchKernelGGL( gm_normalize, dim3( (unsigned) ceil_div( n, 256 ) ), dim3( 256 ), 0, s, n, r.ptr, V.ptr, G,
                          L.sc + 1 );
      hipLaunchKernelGGL( gm_start, dim3( 1 ), dim3( 256 ), 0, s, G, L, m );
      for( int i = 0; i < m && S.next(); i++ ) {
         const double* vi = V.ptr + i * ld;
         matvec_prec( vi, wv.ptr );
         for( int pass = 0; pass < passes; pass++ ) {
            if( classical ) {
               multi_dot( n, i + 1, V.ptr, ld, wv.ptr, G + L.tmp, rw, s );
               hipLaunchKernelGGL( gm_hadd, dim3( 1 ), dim3( 64 ), 0, s, G, L, i, m, G + L.tmp );
               gemv_sub( n, i + 1, V.ptr, ld, G + L.tmp, wv.ptr, s );
            }
            else
               for( int k = 0; k <= i; k++ ) {
                  dot( n, wv.ptr, V.ptr + k * ld, G + L.tmp, rw, s );
                  hipLaunchKernelGGL( gm_hadd1, dim3( 1 ), dim3( 1 ), 0, s, G, L, k, i, m );
                  hipLaunchKernelGGL( gm_axpy, dim3( (unsigned) std::min< int64_t >( 8192, ceil_div( n, 256 ) ) ),
                                      dim3( 256 ), 0, s, n, V.ptr + k * ld, wv.ptr, G, L );
               }
         }
         nrm2( n, wv.ptr, G + L.sc + 2, rw, s );
         hipLaunchKernelGGL( gm_normalize, dim3( (unsigned) ceil_div( n, 256 ) ), dim3( 256 ), 0, s, n, wv.ptr,
                             V.ptr + ( i + 1 ) * ld, G, L.sc + 2 );
         hipLaunchKernelGGL( gm_givens, dim3( 1 ), dim3( 1 ), 0, s, G, L, i, m );
         S.res = hs.read( G + L.sc + 3, 1, s )[ 0 ];
         if( ! S.check() ) {
            o = i;
            break;
         }
      }
      }
      if( o < m ) {
         update( o );
         if( inner ) {
            // the estimate met the rule: the FP64 residue decides
            compute_residue();
            nrm2( n, r.ptr, G + L.sc + 1, rw, s );
            hipLaunchKernelGGL( gm_residue_beta, dim3( 1 ), dim3( 1 ), 0, s, G, L );
            const double* h = hs.read( G + L.sc, 4, s );
            const double beta_old = beta;
            beta = h[ 1 ];
            S.res = h[ 3 ];
            ++cycles;
            beta_ratio = beta / beta_old;
            if( S.check() )
               continue;
         }
         st->iterations = S.it;
         st->residue = S.res;
         st->converged = S.converged();
         return;
      }
      update( m - 1 );
      compute_residue();
      nrm2( n, r.ptr, G + L.sc + 1, rw, s );
      hipLaunchKernelGGL( gm_residue_beta, dim3( 1 ), dim3( 1 ), 0, s, G, L );
      const double* h = hs.read( G + L.sc, 4, s );
      const double beta_old = beta;
      beta = h[ 1 ];
      S.res = h[ 3 ];
      ++cycles;
      beta_ratio = beta / beta_old;
   }
   st->iterations = S.it;
   st->residue = S.res;
   st->converged = S.converged();
}

}  // namespace
}  // namespace tnlmi

// --------------------------------------------------------------- C ABI
namespace {

Operator make_op( const tnlmi_matrix* A, const tnlmi_spmv_plan* plan )
{
   if( ! A || ! A->segments )
      fail( TNLMI_INVALID_ARGUMENT, "solver: matrix is NULL" );
   if( A->rows != A->columns )
      fail( TNLMI_INVALID_ARGUMENT, "solver: the matrix must be square" );
   return Operator{ A, plan, A->rows };
}

void check_state( tnlmi_solver_state* st )
{
   if( ! st )
      fail( TNLMI_INVALID_ARGUMENT, "solver: state is NULL" );
}

}  // namespace

extern "C" {

tnlmi_status tnlmi_preconditioner_create( tnlmi_precond_kind kind, const tnlmi_matrix* A, tnlmi_stream stream,
                                          tnlmi_preconditioner** out )
{
   return guarded( [ & ] {
      if( ! out || ! A || ! A->segments )
         fail( TNLMI_INVALID_ARGUMENT, "preconditioner: NULL argument" );
      *out = nullptr;
      if( A->rows == 0 )
         fail( TNLMI_INVALID_ARGUMENT, "preconditioner update: the matrix is empty" );
      if( A->rows != A->columns )
         fail( TNLMI_INVALID_ARGUMENT, "preconditioner update: matrix must be square" );
      auto* P = new tnlmi_preconditioner;
      P->kind = kind;
      P->n = A->rows;
      hipStream_t s = as_stream( stream );
      try {
         if( kind == TNLMI_PRECOND_DIAGONAL ) {
            P->diag.allocate( A->rows );
            with_value_type( A->value_type, [ & ]( auto tag ) {
               using V = std::remove_const_t< std::remove_pointer_t< decltype( tag ) > >;
               hipLaunchKernelGGL( diag_kernel< V >, dim3( (unsigned) ceil_div( A->rows, 256 ) ), dim3( 256 ), 0, s,
                                   make_view( *A->segments ), A->column_indexes, (const V*) A->values, P->diag.ptr );
            } );
            TNLMI_HIP( hipGetLastError() );
            TNLMI_HIP( hipStreamSynchronize( s ) );
         }
         else if( kind == TNLMI_PRECOND_ILU0 ) {
            if( A->matrix_type == TNLMI_SYMMETRIC )
               fail( TNLMI_NOT_IMPLEMENTED, "ILU0: a general matrix is needed (the factorization reads both triangles)" );
            ilu0_build( *P, *A, s );
         }
         else if( kind == TNLMI_PRECOND_EXPLICIT )
            P->M = *A;
         else if( kind != TNLMI_PRECOND_NONE )
            fail( TNLMI_INVALID_ARGUMENT, "preconditioner: unknown kind" );
      }
      catch( ... ) {
         delete P;
         throw;
      }
      *out = P;
   } );
}

void tnlmi_preconditioner_destroy( tnlmi_preconditioner* P )
{
   delete P;
}

tnlmi_status tnlmi_preconditioner_solve( const tnlmi_preconditioner* P, const double* d_b, double* d_x,
                                         tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( ! P || ! d_b || ! d_x )
         fail( TNLMI_INVALID_ARGUMENT, "preconditioner solve: NULL argument" );
      precond_apply( P, P->n, d_b, d_x, as_stream( stream ) );
      TNLMI_HIP( hipGetLastError() );
   } );
}

tnlmi_status tnlmi_cg_solve( const tnlmi_matrix* A, const tnlmi_spmv_plan* plan, const tnlmi_preconditioner* M,
                             const double* d_b, double* d_x, tnlmi_solver_state* st, tnlmi_stream stream )
{
   return guarded( [ & ] {
      check_state( st );
      const Operator op = make_op( A, plan );
      cg_solve( op, M, d_b, d_x, st, as_stream( stream ) );
      TNLMI_HIP( hipGetLastError() );
   } );
}

tnlmi_status tnlmi_cg_solve_true_residual( const tnlmi_matrix* A, const tnlmi_spmv_plan* plan,
                                           const tnlmi_preconditioner* M, int32_t column_codes, const double* d_b,
                                           double* d_x, tnlmi_solver_state* st, tnlmi_stream stream )
{
   return guarded( [ & ] {
      check_state( st );
      if( M && precond_kind( M ) != TNLMI_PRECOND_DIAGONAL )
         fail( TNLMI_NOT_IMPLEMENTED, "cg_solve_true_residual: no or the Diagonal preconditioner" );
      const Operator op = make_op( A, plan );
      hipStream_t s = as_stream( stream );
      tnlmi_solver_state total = *st;
      total.iterations = 0;
      while( true ) {
         tnlmi_solver_state part = *st;
         part.max_iterations = st->max_iterations - total.iterations;
         fused_cg_solve( op, M, d_b, d_x, &part, s, 1, column_codes != 0 );
         TNLMI_HIP( hipGetLastError() );
         total.iterations += part.iterations;
         total.residue = part.residue;
         total.converged = part.converged;
         if( ! part.converged )
            break;
         // the recurrence residual met the tolerance: confirm on b - A x in
         // FP64; a drift above it continues from x (r recomputed)
         double res = 0;
         const tnlmi_status rc = tnlmi_true_residual( A, d_x, d_b, &res, stream );
         if( rc != TNLMI_OK )
            fail( rc, tnlmi_last_error() );
         total.residue = res;
         if( res <= st->convergence_residue || total.iterations >= st->max_iterations || part.iterations == 0 ) {
            total.converged = res <= st->convergence_residue ? 1 : 0;
            break;
         }
      }
      *st = total;
   } );
}

tnlmi_status tnlmi_gmres_solve( const tnlmi_matrix* A, const tnlmi_spmv_plan* plan, const tnlmi_preconditioner* M,
                                tnlmi_gmres_variant variant, int32_t restarting_min, int32_t restarting_max,
                                int32_t step_min, int32_t step_max, const double* d_b, double* d_x,
                                tnlmi_solver_state* st, tnlmi_stream stream )
{
   return guarded( [ & ] {
      check_state( st );
      const Operator op = make_op( A, plan );
      gmres_solve( op, M, d_b, d_x, variant, restarting_min, restarting_max, step_min, step_max, st,
                   as_stream( stream ), nullptr );
      TNLMI_HIP( hipGetLastError() );
   } );
}

tnlmi_status tnlmi_gmres_mixed_solve( const tnlmi_matrix* A, const tnlmi_matrix* A_lo, const tnlmi_spmv_plan* plan,
                                      const tnlmi_preconditioner* M, tnlmi_gmres_variant variant,
                                      int32_t restarting_min, int32_t restarting_max, int32_t step_min,
                                      int32_t step_max, const double* d_b, double* d_x, tnlmi_solver_state* st,
                                      tnlmi_stream stream )
{
   return guarded( [ & ] {
      check_state( st );
      if( ! A_lo || ! A_lo->segments || ! A || ! A->segments )
         fail( TNLMI_INVALID_ARGUMENT, "mixed GMRES: matrix is NULL" );
      if( A_lo->rows != A->rows || A_lo->columns != A->columns )
         fail( TNLMI_INVALID_ARGUMENT, "mixed GMRES: the inner matrix has other dimensions" );
      if( A_lo->value_type != TNLMI_F64 && A_lo->value_type != TNLMI_F32 && A_lo->value_type != TNLMI_BF16 )
         fail( TNLMI_INVALID_ARGUMENT, "mixed GMRES: the inner matrix holds FP64, FP32 or BF16 values" );
      if( A->value_type != TNLMI_F64 )
         fail( TNLMI_INVALID_ARGUMENT, "mixed GMRES: the matrix holds FP64 values" );
      const Operator op = make_op( A, plan );
      const Operator lo = make_op( A_lo, nullptr );
      gmres_solve( op, M, d_b, d_x, variant, restarting_min, restarting_max, step_min, step_max, st,
                   as_stream( stream ), &lo );
      TNLMI_HIP( hipGetLastError() );
   } );
}

tnlmi_status tnlmi_jacobi_solve( const tnlmi_matrix* A, const tnlmi_spmv_plan* plan, double omega,
                                 int64_t residue_period, const double* d_b, double* d_x, tnlmi_solver_state* st,
                                 tnlmi_stream stream )
{
   return guarded( [ & ] {
      check_state( st );
      const Operator op = make_op( A, plan );
      if( residue_period <= 0 )
         fail( TNLMI_INVALID_ARGUMENT, "Jacobi: residue period must be positive" );
      jacobi_solve( op, *A->segments, *A, d_b, d_x, omega, residue_period, st, as_stream( stream ) );
      TNLMI_HIP( hipGetLastError() );
   } );
}

tnlmi_status tnlmi_true_residual( const tnlmi_matrix* A, const double* d_x, const double* d_b, double* residual,
                                  tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( ! residual )
         fail( TNLMI_INVALID_ARGUMENT, "true_residual: NULL output" );
      const Operator op = make_op( A, nullptr );
      hipStream_t s = as_stream( stream );
      DeviceArray< double > tmp, sc;
      tmp.allocate( std::max< int64_t >( 1, op.rows ) );
      sc.allocate( 4 );
      RedWork w;
      nrm2( op.rows, d_b, sc.ptr + 0, w, s );
      residue_dev( op, d_x, d_b, sc.ptr + 0, tmp.ptr, sc.ptr + 1, w, s );
      double h[ 2 ];
      TNLMI_HIP( hipMemcpyAsync( h, sc.ptr, 2 * sizeof( double ), hipMemcpyDeviceToHost, s ) );
      TNLMI_HIP( hipStreamSynchronize( s ) );
      *residual = h[ 1 ];
   } );
}

}  // extern "C"
