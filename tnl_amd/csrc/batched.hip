// batched.hip -- thousands of small independent SPD systems solved by one
// persistent kernel (Jacobi-preconditioned CG, CG.hpp:10-108 per system).
//
// The systems are the diagonal blocks of one block-diagonal matrix stored in
// any segment format (row ranges sys_rows[k] .. sys_rows[k+1]).  Workgroups
// take the next system from an atomic counter and run the whole CG for it:
// the rows are spread over the workgroup's threads, dot products are
// workgroup reductions, and the only synchronisation is the workgroup
// barrier -- no grid-wide synchronisation and no host round trip.  Every
// workgroup leaves when the counter passes the last system, and every
// system's loop is bounded by max_iterations.  The launch (counter reset +
// solve) can be captured once in a hipGraph and replayed.
//
// Mixed variant: FP64 iterative refinement per system (true residual
// ||b - A x|| / ||b|| in FP64) around an inner Jacobi-CG in FP32 on the same
// matrix with FP32 values.
//
// No reference counterpart (TNL solves one system per solver call).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>

#include "segments.hpp"

namespace tnlmi {
namespace {

template< typename T, typename V >
__device__ __forceinline__ T row_dot( const SegView& v, const int32_t* __restrict__ cols, const V* __restrict__ vals,
                                      const T* __restrict__ x, int64_t r )
{
   T acc = 0;
   // every getSegmentSize slot: an Ellpack-type row may hold entries past its
   // requested capacity
   const int32_t len = segment_size( v, r );
   if( v.format == TNLMI_CSR ) {
      const int64_t b = v.offsets[ r ];
      for( int32_t k = 0; k < len; k++ ) {
         const int32_t c = cols[ b + k ];
         if( c >= 0 )
            acc += (T) load_val< V, T >( vals, b + k ) * x[ c ];
      }
   }
   else
      for( int32_t k = 0; k < len; k++ ) {
         const int64_t g = global_index( v, r, k );
         const int32_t c = cols[ g ];
         if( c >= 0 )
            acc += (T) load_val< V, T >( vals, g ) * x[ c ];
      }
   return acc;
}

// The same row sum spread over the T lanes of a group (T = 1, 2, 4, 8: the
// rows per wavefront are 64 / T): lane j sums slots j, j + T, ..., then an xor
// butterfly inside the group; every lane returns the row's sum.  All T lanes
// of a group hold the same row, so the group is active or idle as a whole.
template< int T, typename A, typename V >
__device__ __forceinline__ A row_dot_group( const SegView& v, const int32_t* __restrict__ cols,
                                            const V* __restrict__ vals, const A* __restrict__ x, int64_t r, int lane )
{
   if constexpr( T == 1 )
      return row_dot< A, V >( v, cols, vals, x, r );
   else {
      A acc = 0;
      const int32_t len = segment_size( v, r );
      if( v.format == TNLMI_CSR ) {
         const int64_t b = v.offsets[ r ];
         for( int32_t k = lane; k < len; k += T ) {
            const int32_t c = cols[ b + k ];
            if( c >= 0 )
               acc += (A) load_val< V, A >( vals, b + k ) * x[ c ];
         }
      }
      else
         for( int32_t k = lane; k < len; k += T ) {
            const int64_t g = global_index( v, r, k );
            const int32_t c = cols[ g ];
            if( c >= 0 )
               acc += (A) load_val< V, A >( vals, g ) * x[ c ];
         }
      for( int off = T / 2; off > 0; off /= 2 )
         acc += __shfl_xor( acc, off, T );
      return acc;
   }
}

struct BatchArgs
{
   int64_t nsys;
   const int64_t* sys_rows;
   SegView v;
   const int32_t* cols;
   const double* vals;
   const float* vals_lo;  // mixed variant only
   const double* dinv;
   const double* b;
   double* x;
   double* r;
   double* p;
   double* Ap;
   float* r32;  // mixed variant work vectors
   float* p32;
   float* Ap32;
   float* d32;
   double tol;
   int64_t maxit;
   double inner_tol;
   int64_t inner_maxit;
   int* counter;
   int64_t* iters;
   double* res;
};

inline uint64_t double_bits( double v )
{
   uint64_t u;
   std::memcpy( &u, &v, sizeof( u ) );
   return u;
}

__global__ void reset_counter( int* c )
{
   c[ 0 ] = 0;
}

template< typename V >
__global__ void batch_inv_diag( SegView v, const int32_t* __restrict__ cols, const V* __restrict__ vals,
                                double* __restrict__ dinv )
{
   const int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( r >= v.n )
      return;
   double d = 0;
   const int32_t size = segment_size( v, r );
   for( int32_t k = 0; k < size; k++ ) {
      const int64_t g = global_index( v, r, k );
      if( cols[ g ] == r ) {
         d = (double) load_val< V, double >( vals, g );
         break;
      }
   }
   dinv[ r ] = d != 0.0 ? 1.0 / d : 1.0;
}

// Sets *bad when a stored column of row r (padding, -1, aside) lies outside the
// row range of r's system: the persistent kernel solves each system alone and
// would read the vectors of another system that a concurrent workgroup updates.
__global__ void check_blocks( SegView v, const int32_t* __restrict__ cols, const int64_t* __restrict__ sys_rows,
                              int64_t nsys, int32_t* bad )
{
   const int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( r >= v.n )
      return;
   int64_t lo = 0, hi = nsys;  // the system k with sys_rows[k] <= r < sys_rows[k+1]
   while( hi - lo > 1 ) {
      const int64_t mid = ( lo + hi ) / 2;
      if( sys_rows[ mid ] <= r )
         lo = mid;
      else
         hi = mid;
   }
   const int64_t first = sys_rows[ lo ], last = sys_rows[ lo + 1 ];
   const int32_t size = segment_size( v, r );
   for( int32_t k = 0; k < size; k++ ) {
      const int32_t c = cols[ global_index( v, r, k ) ];
      if( c != -1 && ( c < first || c >= last ) )
         *bad = 1;
   }
}

// The next system for the whole workgroup.  Only whole wavefronts branch
// here (wave 0 fetches, the others wait at the barrier): a lane-divergent
// "thread 0 fetches" block next to the loop's back edge lets the compiler
// split the loop so that the other lanes of thread 0's wavefront loop back
// without it and re-read a stale slot forever.  Lane 0 adds one, the other lanes add zero
// (the wave's adds become one atomic); the value is lane 0's.
__device__ __forceinline__ int64_t next_system( int* counter, int* slot )
{
   if( threadIdx.x < 64 ) {
      const int got = atomicAdd( counter, threadIdx.x == 0 ? 1 : 0 );
      *slot = __builtin_amdgcn_readfirstlane( got );
   }
   __syncthreads();
   const int s = __builtin_amdgcn_readfirstlane( *slot );
   __syncthreads();
   return s;
}

// Per-system results, written by wavefront 0 (every lane the same value).
__device__ __forceinline__ void put_result( int64_t* iters, double* res, int64_t sys, int64_t it, double r )
{
   if( threadIdx.x < 64 ) {
      iters[ sys ] = it;
      res[ sys ] = r;
   }
}

// FP64 Jacobi-CG per system, TNL's stopping rule with default min/divergence
template< int B, int T >
__global__ void __launch_bounds__( B ) batched_cg64( BatchArgs a )
{
   constexpr int G = B / T;  // row groups of T lanes
   const int grp = threadIdx.x / T, lane = threadIdx.x % T;
   __shared__ double red[ B / 64 ];
   __shared__ int slot;
   for( int64_t sys = next_system( a.counter, &slot ); sys < a.nsys; sys = next_system( a.counter, &slot ) ) {
      const int64_t r0 = a.sys_rows[ sys ], r1 = a.sys_rows[ sys + 1 ];
      // r = b - A x, p = z = D^{-1} r, s1 = (r, z), normb = ||D^{-1} b||
      double acc1 = 0, acc2 = 0;
      for( int64_t i = r0 + grp; i < r1; i += G ) {
         const double ri = a.b[ i ] - row_dot_group< T, double, double >( a.v, a.cols, a.vals, a.x, i, lane );
         if( lane == 0 ) {
            const double zi = ri * a.dinv[ i ];
            a.r[ i ] = ri;
            a.p[ i ] = zi;
            acc1 += ri * zi;
            const double mb = a.b[ i ] * a.dinv[ i ];
            acc2 += mb * mb;
         }
      }
      double s1 = block_sum< B >( acc1, red );
      const double normb = sqrt( block_sum< B >( acc2, red ) );
      int64_t it = 0;
      double res = 0;
      if( normb == 0.0 ) {
         for( int64_t i = r0 + threadIdx.x; i < r1; i += B )
            a.x[ i ] = 0.0;
      }
      else {
         res = sqrt( s1 ) / normb;
         while( ! ( isnan( res ) || it > a.maxit || res < a.tol ) ) {
            it++;
            double acc = 0;
            for( int64_t i = r0 + grp; i < r1; i += G ) {
               const double api = row_dot_group< T, double, double >( a.v, a.cols, a.vals, a.p, i, lane );
               if( lane == 0 ) {
                  a.Ap[ i ] = api;
                  acc += api * a.p[ i ];
               }
            }
            const double s2 = block_sum< B >( acc, red );
            if( s2 == 0.0 ) {
               res = 0.0;
               break;
            }
            const double alpha = s1 / s2;
            acc = 0;
            for( int64_t i = r0 + threadIdx.x; i < r1; i += B ) {
               a.x[ i ] += alpha * a.p[ i ];
               const double ri = a.r[ i ] - alpha * a.Ap[ i ];
               a.r[ i ] = ri;
               acc += ri * ri * a.dinv[ i ];
            }
            const double s1n = block_sum< B >( acc, red );
            const double beta = s1 == 0.0 ? 0.0 : s1n / s1;
            s1 = s1n;
            for( int64_t i = r0 + threadIdx.x; i < r1; i += B )
               a.p[ i ] = a.r[ i ] * a.dinv[ i ] + beta * a.p[ i ];
            __syncthreads();
            res = sqrt( s1 ) / normb;
         }
      }
      put_result( a.iters, a.res, sys, it, res );
   }
}

// Mixed: FP64 refinement around FP32 Jacobi-CG (FP32 values, vectors, SpMV
// accumulation; workgroup reductions in FP64).  iters[] = inner iterations.
template< int B, int T >
__global__ void __launch_bounds__( B ) batched_cg_mixed( BatchArgs a )
{
   constexpr int G = B / T;
   const int grp = threadIdx.x / T, lane = threadIdx.x % T;
   __shared__ double red[ B / 64 ];
   __shared__ int slot;
   for( int64_t sys = next_system( a.counter, &slot ); sys < a.nsys; sys = next_system( a.counter, &slot ) ) {
      const int64_t r0 = a.sys_rows[ sys ], r1 = a.sys_rows[ sys + 1 ];
      double acc = 0;
      for( int64_t i = r0 + threadIdx.x; i < r1; i += B )
         acc += a.b[ i ] * a.b[ i ];
      const double bnorm = sqrt( block_sum< B >( acc, red ) );
      int64_t inner_total = 0, outer = 0;
      double res = 0;
      while( bnorm > 0.0 ) {
         // r = b - A x (FP64) and the true residual
         acc = 0;
         for( int64_t i = r0 + grp; i < r1; i += G ) {
            const double ri = a.b[ i ] - row_dot_group< T, double, double >( a.v, a.cols, a.vals, a.x, i, lane );
            if( lane == 0 ) {
               a.r[ i ] = ri;
               acc += ri * ri;
            }
         }
         const double rn = sqrt( block_sum< B >( acc, red ) );
         res = rn / bnorm;
         if( isnan( res ) || res < a.tol || outer >= a.maxit )
            break;
         outer++;
         // inner CG in FP32 on A d = r / ||r||, d = 0
         const double inv = 1.0 / rn;
         double s1a = 0;
         for( int64_t i = r0 + threadIdx.x; i < r1; i += B ) {
            const float ri = (float) ( a.r[ i ] * inv );
            const float zi = ri * (float) a.dinv[ i ];
            a.r32[ i ] = ri;
            a.p32[ i ] = zi;
            a.d32[ i ] = 0.0f;
            s1a += (double) ri * (double) zi;
         }
         double s1 = block_sum< B >( s1a, red );
         const double s10 = s1;
         for( int64_t k = 0; k < a.inner_maxit && s1 > a.inner_tol * a.inner_tol * s10; k++ ) {
            inner_total++;
            double q = 0;
            for( int64_t i = r0 + grp; i < r1; i += G ) {
               const float api = row_dot_group< T, float, float >( a.v, a.cols, a.vals_lo, a.p32, i, lane );
               if( lane == 0 ) {
                  a.Ap32[ i ] = api;
                  q += (double) api * (double) a.p32[ i ];
               }
            }
            const double s2 = block_sum< B >( q, red );
            if( s2 == 0.0 )
               break;
            const float alpha = (float) ( s1 / s2 );
            q = 0;
            for( int64_t i = r0 + threadIdx.x; i < r1; i += B ) {
               a.d32[ i ] += alpha * a.p32[ i ];
               const float ri = a.r32[ i ] - alpha * a.Ap32[ i ];
               a.r32[ i ] = ri;
               q += (double) ri * (double) ri * a.dinv[ i ];
            }
            const double s1n = block_sum< B >( q, red );
            const float beta = (float) ( s1 == 0.0 ? 0.0 : s1n / s1 );
            s1 = s1n;
            for( int64_t i = r0 + threadIdx.x; i < r1; i += B )
               a.p32[ i ] = a.r32[ i ] * (float) a.dinv[ i ] + beta * a.p32[ i ];
            __syncthreads();
         }
         for( int64_t i = r0 + threadIdx.x; i < r1; i += B )
            a.x[ i ] += rn * (double) a.d32[ i ];
         __syncthreads();
      }
      put_result( a.iters, a.res, sys, inner_total, res );
   }
}

}  // namespace
}  // namespace tnlmi

using namespace tnlmi;

struct tnlmi_batched_cg
{
   const tnlmi_matrix* A = nullptr;
   const tnlmi_matrix* A_lo = nullptr;
   int64_t nsys = 0, rows = 0;
   int32_t block = 512;
   int32_t lanes = 1;  // lanes per row in the row sums (64 / lanes rows per wavefront)
   int64_t grid = 1;  // persistent workgroups: a few per CU, at most one per system
   DeviceArray< int64_t > sys_rows;
   DeviceArray< double > dinv, r, p, Ap;
   DeviceArray< float > r32, p32, Ap32, d32;
   DeviceArray< int > counter;
   // captured launches keyed by their arguments
   // (the tolerances by their bit patterns: a NaN key still orders)
   std::map< std::tuple< const void*, const void*, const void*, const void*, uint64_t, uint64_t, int64_t, int >,
            hipGraphExec_t >
      graphs;
   hipStream_t capture_stream = nullptr;
   ~tnlmi_batched_cg()
   {
      for( auto& g : graphs )
         (void) hipGraphExecDestroy( g.second );
      if( capture_stream )
         (void) hipStreamDestroy( capture_stream );
   }
};

namespace {

void launch( tnlmi_batched_cg& P, const BatchArgs& a, bool mixed, hipStream_t s )
{
   hipLaunchKernelGGL( reset_counter, dim3( 1 ), dim3( 1 ), 0, s, a.counter );
   const dim3 grid( (unsigned) P.grid );
#define TNLMI_BATCH( BS, T )                                                                                          \
   if( P.block == BS && P.lanes == T ) {                                                                              \
      if( mixed )                                                                                                     \
         hipLaunchKernelGGL( ( batched_cg_mixed< BS, T > ), grid, dim3( BS ), 0, s, a );                              \
      else                                                                                                            \
         hipLaunchKernelGGL( ( batched_cg64< BS, T > ), grid, dim3( BS ), 0, s, a );                                  \
      return;                                                                                                         \
   }
#define TNLMI_BATCH_T( BS ) TNLMI_BATCH( BS, 1 ) TNLMI_BATCH( BS, 2 ) TNLMI_BATCH( BS, 4 ) TNLMI_BATCH( BS, 8 )
   TNLMI_BATCH_T( 128 )
   TNLMI_BATCH_T( 256 )
   TNLMI_BATCH_T( 512 )
   TNLMI_BATCH_T( 1024 )
#undef TNLMI_BATCH_T
#undef TNLMI_BATCH
   fail( TNLMI_INVALID_ARGUMENT, "batched CG: block size must be 128, 256, 512 or 1024 and lanes per row 1, 2, 4 or 8" );
}

}  // namespace

extern "C" {

tnlmi_status tnlmi_batched_cg_create( const tnlmi_matrix* A, const tnlmi_matrix* A_lo, const int64_t* sys_rows,
                                      int64_t nsys, int32_t block_size, int32_t lanes_per_row, tnlmi_stream stream,
                                      tnlmi_batched_cg** out )
{
   return guarded( [ & ] {
      if( ! out || ! A || ! A->segments || ! sys_rows || nsys <= 0 )
         fail( TNLMI_INVALID_ARGUMENT, "batched_cg_create: invalid argument" );
      *out = nullptr;
      if( A->value_type != TNLMI_F64 )
         fail( TNLMI_INVALID_ARGUMENT, "batched_cg_create: the matrix must hold FP64 values" );
      if( A->matrix_type == TNLMI_SYMMETRIC || ( A_lo && A_lo->matrix_type == TNLMI_SYMMETRIC ) )
         fail( TNLMI_NOT_IMPLEMENTED, "batched_cg_create: symmetric storage is not supported by the batched kernel" );
      if( A_lo && ( A_lo->value_type != TNLMI_F32 || A_lo->segments != A->segments ) )
         fail( TNLMI_INVALID_ARGUMENT, "batched_cg_create: the inner matrix must be an FP32 copy on the same segments" );
      if( sys_rows[ 0 ] != 0 || sys_rows[ nsys ] != A->rows )
         fail( TNLMI_INVALID_ARGUMENT, "batched_cg_create: system row ranges must cover the matrix" );
      for( int64_t k = 0; k < nsys; k++ )
         if( sys_rows[ k + 1 ] < sys_rows[ k ] )
            fail( TNLMI_INVALID_ARGUMENT, "batched_cg_create: system row ranges must be ascending" );
      auto* P = new tnlmi_batched_cg;
      try {
         hipStream_t s = as_stream( stream );
         P->A = A;
         P->A_lo = A_lo;
         P->nsys = nsys;
         P->rows = A->rows;
         P->block = block_size ? block_size : 512;
         if( P->block != 128 && P->block != 256 && P->block != 512 && P->block != 1024 )
            fail( TNLMI_INVALID_ARGUMENT, "batched CG: block size must be 128, 256, 512 or 1024" );
         P->lanes = lanes_per_row ? lanes_per_row : 1;
         if( P->lanes != 1 && P->lanes != 2 && P->lanes != 4 && P->lanes != 8 )
            fail( TNLMI_INVALID_ARGUMENT, "batched CG: lanes per row must be 1, 2, 4 or 8" );
         int dev = 0, cus = 256;
         TNLMI_HIP( hipGetDevice( &dev ) );
         TNLMI_HIP( hipDeviceGetAttribute( &cus, hipDeviceAttributeMultiprocessorCount, dev ) );
         const int per_cu = P->block >= 1024 ? 2 : P->block >= 512 ? 4 : 8;
         P->grid = std::max< int64_t >( 1, std::min< int64_t >( (int64_t) cus * per_cu, nsys ) );
         P->sys_rows.allocate( nsys + 1 );
         TNLMI_HIP( hipMemcpyAsync( P->sys_rows.ptr, sys_rows, 8 * ( nsys + 1 ), hipMemcpyHostToDevice, s ) );
         const int64_t n = A->rows;
         for( auto* v : { &P->dinv, &P->r, &P->p, &P->Ap } )
            v->allocate( n );
         if( A_lo )
            for( auto* v : { &P->r32, &P->p32, &P->Ap32, &P->d32 } )
               v->allocate( n );
         P->counter.allocate( 1 );
         {
            DeviceArray< int32_t > bad;
            bad.allocate( 1 );
            TNLMI_HIP( hipMemsetAsync( bad.ptr, 0, sizeof( int32_t ), s ) );
            if( n )
               hipLaunchKernelGGL( check_blocks, dim3( (unsigned) ceil_div( n, 256 ) ), dim3( 256 ), 0, s,
                                   make_view( *A->segments ), A->column_indexes, P->sys_rows.ptr, nsys, bad.ptr );
            TNLMI_HIP( hipGetLastError() );
            int32_t h = 0;
            TNLMI_HIP( hipMemcpyAsync( &h, bad.ptr, sizeof( h ), hipMemcpyDeviceToHost, s ) );
            TNLMI_HIP( hipStreamSynchronize( s ) );
            if( h )
               fail( TNLMI_INVALID_ARGUMENT,
                     "batched_cg_create: a stored column lies outside its system's row range (the systems must be "
                     "the diagonal blocks of the matrix)" );
         }
         hipLaunchKernelGGL( batch_inv_diag< double >, dim3( (unsigned) ceil_div( n, 256 ) ), dim3( 256 ), 0, s,
                             make_view( *A->segments ), A->column_indexes, (const double*) A->values, P->dinv.ptr );
         TNLMI_HIP( hipGetLastError() );
         TNLMI_HIP( hipStreamSynchronize( s ) );
      }
      catch( ... ) {
         delete P;
         throw;
      }
      *out = P;
   } );
}

void tnlmi_batched_cg_destroy( tnlmi_batched_cg* P )
{
   delete P;
}

tnlmi_status tnlmi_batched_cg_solve( tnlmi_batched_cg* P, int32_t mixed, const double* d_b, double* d_x,
                                     double tolerance, int64_t max_iterations, double inner_tolerance,
                                     int64_t* d_iterations, double* d_residues, int32_t use_graph,
                                     tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( ! P || ! d_b || ! d_x || ! d_iterations || ! d_residues )
         fail( TNLMI_INVALID_ARGUMENT, "batched_cg_solve: NULL argument" );
      if( mixed && ! P->A_lo )
         fail( TNLMI_INVALID_ARGUMENT, "batched_cg_solve: the mixed variant needs the FP32 matrix" );
      BatchArgs a{};
      a.nsys = P->nsys;
      a.sys_rows = P->sys_rows.ptr;
      a.v = make_view( *P->A->segments );
      a.cols = P->A->column_indexes;
      a.vals = (const double*) P->A->values;
      a.vals_lo = P->A_lo ? (const float*) P->A_lo->values : nullptr;
      a.dinv = P->dinv.ptr;
      a.b = d_b;
      a.x = d_x;
      a.r = P->r.ptr;
      a.p = P->p.ptr;
      a.Ap = P->Ap.ptr;
      a.r32 = P->r32.ptr;
      a.p32 = P->p32.ptr;
      a.Ap32 = P->Ap32.ptr;
      a.d32 = P->d32.ptr;
      a.tol = tolerance;
      a.maxit = max_iterations;
      a.inner_tol = inner_tolerance;
      a.inner_maxit = 100000;
      a.counter = P->counter.ptr;
      a.iters = d_iterations;
      a.res = d_residues;
      hipStream_t s = as_stream( stream );
      if( ! use_graph ) {
         launch( *P, a, mixed, s );
         TNLMI_HIP( hipGetLastError() );
         return;
      }
      // one captured launch per argument set, replayed on later calls
      const auto key = std::make_tuple( (const void*) d_b, (const void*) d_x, (const void*) d_iterations,
                                        (const void*) d_residues, double_bits( tolerance ),
                                        double_bits( inner_tolerance ), max_iterations, (int) mixed );
      auto it = P->graphs.find( key );
      if( it == P->graphs.end() ) {
         if( ! P->capture_stream )
            TNLMI_HIP( hipStreamCreateWithFlags( &P->capture_stream, hipStreamNonBlocking ) );
         hipGraph_t g;
         TNLMI_HIP( hipStreamBeginCapture( P->capture_stream, hipStreamCaptureModeThreadLocal ) );
         launch( *P, a, mixed, P->capture_stream );
         TNLMI_HIP( hipStreamEndCapture( P->capture_stream, &g ) );
         hipGraphExec_t ge;
         TNLMI_HIP( hipGraphInstantiate( &ge, g, nullptr, nullptr, 0 ) );
         TNLMI_HIP( hipGraphDestroy( g ) );
         it = P->graphs.emplace( key, ge ).first;
      }
      TNLMI_HIP( hipGraphLaunch( it->second, s ) );
   } );
}

}  // extern "C"
