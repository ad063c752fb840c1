// spmv.hip -- y = alpha*A*x + beta*y for every segment format, wave64 kernels
// for gfx950.  Replaces SparseMatrixBase::vectorProduct
// (Matrices/SparseMatrixBase.hpp:267-391) and the per-format
// ReducingOperations<...>::reduceSegments GPU paths
// (Algorithms/Segments/detail/ReducingOperations_*.h).
//
// Semantics kept from the reference: the fetch of element g is
//   column[g] == -1 ? 0 : value[g] * x[column[g]]      (SparseMatrixBase.hpp:331-345)
// accumulated in the compute type from 0, and the row result r is stored by
// one of the four keep variants on (beta == 0, alpha == 1)   (:346-389).
// Only rows in [begin, end) are written.
#include <algorithm>

#include "segments.hpp"
#include "spmv.hpp"
#include "spmv_tile.hpp"

namespace tnlmi {
namespace {

template< typename V, typename X, typename C >
__device__ __forceinline__ C fetch( const int32_t* __restrict__ cols, const V* __restrict__ vals,
                                    const X* __restrict__ x, int64_t g )
{
   const int32_t c = cols[ g ];
   return c >= 0 ? load_val< V, C >( vals, g ) * (C) x[ c ] : C( 0 );
}

template< typename V, typename X, typename C >
__device__ __forceinline__ C fetch_nt( const int32_t* __restrict__ cols, const V* __restrict__ vals,
                                       const X* __restrict__ x, int64_t g )
{
   // the value is read only for a stored entry: a padding slot (column -1)
   // costs its 4-byte column index, and a wavefront deep in a padded format's
   // padding issues no value loads (the gather of x waits for the column
   // anyway, so this adds no memory round trip)
   const int32_t c = nt_load( cols + g );
   if( c < 0 )
      return C( 0 );
   const C v = nt_load_val< V, C >( vals, g );
   return v * (C) x[ c ];
}

// sum of fetch_nt over the slots base + stride * k, k < count, added in k
// order: the column indexes of kSU slots are loaded together, then their
// values and x entries, so a long strided walk (a padded slice or chunk)
// keeps kSU loads in flight instead of one dependent pair per slot
constexpr int kSU = 8;

template< typename V, typename X, typename C >
__device__ __forceinline__ C strided_sum( const int32_t* __restrict__ cols, const V* __restrict__ vals,
                                          const X* __restrict__ x, int64_t base, int64_t stride, int32_t count )
{
   C acc = 0;
   int32_t k = 0;
   for( ; k + kSU <= count; k += kSU ) {
      int32_t c[ kSU ];
#pragma unroll
      for( int u = 0; u < kSU; u++ )
         c[ u ] = nt_load( cols + base + stride * ( k + u ) );
      C t[ kSU ];
#pragma unroll
      for( int u = 0; u < kSU; u++ )
         t[ u ] = c[ u ] >= 0 ? nt_load_val< V, C >( vals, base + stride * ( k + u ) ) * (C) x[ c[ u ] ] : C( 0 );
#pragma unroll
      for( int u = 0; u < kSU; u++ )
         acc += t[ u ];
   }
   for( ; k < count; k++ )
      acc += fetch_nt< V, X, C >( cols, vals, x, base + stride * k );
   return acc;
}

template< typename V, typename X, typename C >
struct Args
{
   const int32_t* __restrict__ cols;
   const V* __restrict__ vals;
   const X* __restrict__ x;
   X* __restrict__ y;
   C alpha, beta;
   int64_t begin, end;
   int64_t n_rows;
};

// ======================================================================= CSR

// Fixed mapping, 1 thread per row (the reference's GPU default,
// ReducingOperations_CSR.h:33-77 via parallelFor).
template< typename V, typename X, typename C >
__global__ void __launch_bounds__( 256 ) csr_scalar( const int32_t* __restrict__ off, Args< V, X, C > a )
{
   const int64_t r = a.begin + (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( r >= a.end )
      return;
   C acc = 0;
   for( int64_t g = off[ r ]; g < off[ r + 1 ]; g++ )
      acc += fetch< V, X, C >( a.cols, a.vals, a.x, g );
   keep( a.y, r, acc, a.alpha, a.beta );
}

// G lanes per row, G in {2..64} (VariableVector / Vector kernels,
// ReducingKernels_CSR.h:18-118): strided partial sums, butterfly reduction.
template< typename V, typename X, typename C, int G >
__global__ void __launch_bounds__( 256 ) csr_vector( const int32_t* __restrict__ off, Args< V, X, C > a )
{
   const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   const int64_t r = a.begin + t / G;
   const int lane = (int) ( t % G );
   if( r >= a.end )
      return;  // uniform over the G-lane group
   C acc = 0;
   const int64_t e = off[ r + 1 ];
   for( int64_t g = off[ r ] + lane; g < e; g += G )
      acc += fetch< V, X, C >( a.cols, a.vals, a.x, g );
   acc = group_sum< G >( acc );
   if( lane == 0 )
      keep( a.y, r, acc, a.alpha, a.beta );
}

// TPS in {128, 256}: several waves per row, per-wave partials combined in LDS
// (LightMultivector kernel, ReducingKernels_CSR.h:120-189).
template< typename V, typename X, typename C, int TPS >
__global__ void __launch_bounds__( 256 ) csr_multivector( const int32_t* __restrict__ off, Args< V, X, C > a )
{
   __shared__ C part[ 4 ];
   constexpr int RPB = 256 / TPS;
   const int local = threadIdx.x % TPS;
   const int64_t r = a.begin + (int64_t) blockIdx.x * RPB + threadIdx.x / TPS;
   C acc = 0;
   if( r < a.end ) {
      const int64_t e = off[ r + 1 ];
      for( int64_t g = off[ r ] + local; g < e; g += TPS )
         acc += fetch< V, X, C >( a.cols, a.vals, a.x, g );
   }
   acc = group_sum< 64 >( acc );
   if( ( threadIdx.x & 63 ) == 0 )
      part[ threadIdx.x >> 6 ] = acc;
   __syncthreads();
   if( local == 0 && r < a.end ) {
      C s = 0;
      const int w0 = ( threadIdx.x / TPS ) * ( TPS / 64 );
      for( int w = 0; w < TPS / 64; w++ )
         s += part[ w0 + w ];
      keep( a.y, r, s, a.alpha, a.beta );
   }
}

// Dynamic grouping (ReducingKernels_CSR.h:191-352): a block owns 256/TPS
// rows; rows longer than the block are reduced by the whole block, rows longer
// than a wavefront by one wavefront, the rest by their own thread.
template< typename V, typename X, typename C >
__global__ void __launch_bounds__( 256 ) csr_dynamic( const int32_t* __restrict__ off, Args< V, X, C > a, int tps )
{
   __shared__ int32_t longq[ 256 ], medq[ 256 ];
   __shared__ int nlong, nmed;
   __shared__ C part[ 4 ];
   const int rpb = 256 / tps;
   const int64_t r = a.begin + (int64_t) blockIdx.x * rpb + threadIdx.x;
   const bool mine = threadIdx.x < rpb && r < a.end;
   if( threadIdx.x == 0 ) {
      nlong = 0;
      nmed = 0;
   }
   __syncthreads();
   int64_t len = mine ? off[ r + 1 ] - off[ r ] : 0;
   if( mine && len > 256 )
      longq[ atomicAdd( &nlong, 1 ) ] = (int32_t) r;
   else if( mine && len > 64 )
      medq[ atomicAdd( &nmed, 1 ) ] = (int32_t) r;
   __syncthreads();
   for( int i = 0; i < nlong; i++ ) {
      const int64_t lr = longq[ i ];
      C acc = 0;
      for( int64_t g = off[ lr ] + threadIdx.x; g < off[ lr + 1 ]; g += 256 )
         acc += fetch< V, X, C >( a.cols, a.vals, a.x, g );
      acc = block_sum< 256 >( acc, part );
      if( threadIdx.x == 0 )
         keep( a.y, lr, acc, a.alpha, a.beta );
      __syncthreads();
   }
   const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
   for( int i = w; i < nmed; i += 4 ) {
      const int64_t mr = medq[ i ];
      C acc = 0;
      for( int64_t g = off[ mr ] + lane; g < off[ mr + 1 ]; g += 64 )
         acc += fetch< V, X, C >( a.cols, a.vals, a.x, g );
      acc = group_sum< 64 >( acc );
      if( lane == 0 )
         keep( a.y, mr, acc, a.alpha, a.beta );
   }
   if( mine && len <= 64 ) {
      C acc = 0;
      for( int64_t g = off[ r ]; g < off[ r + 1 ]; g++ )
         acc += fetch< V, X, C >( a.cols, a.vals, a.x, g );
      keep( a.y, r, acc, a.alpha, a.beta );
   }
}

// ---------------------------------------------------------------- tiled CSR
// The MI355X default ("merge-path" split of the element stream).  Wavefront t
// owns exactly the elements [t*T, (t+1)*T) and the rows whose first element
// lies there (tile_rows, built with the segments).  Phase 1 streams the tile's
// columns/values once with non-temporal, fully coalesced loads (T/64
// independent loads per lane in flight), gathers x and stages the products in
// LDS; phase 2 reduces each owned row's part of the tile from LDS with
// 64/pow2(rows) lanes per row, or one lane per row (left to right from 0,
// the reference's host order) when the tile holds 64 rows or more.  A row
// that continues past the tile end is written by its owner with its partial
// sum; the tiles it continues into leave their partial in carry[] and
// csr_carry_fixup adds them, in tile order, after the main kernel.  Every
// wavefront therefore does the same amount of streaming work, whatever the
// row-length distribution.  Workgroups are renumbered so that each XCD walks
// one contiguous range of tiles (x entries shared by neighbouring rows stay in
// one L2).  WHOLE (no row longer than kWholeRowMax, e.g. stencil matrices):
// the owner also reads the few elements of its last row past the tile end, so
// every row is summed complete by one owner, no carries exist and the product
// is one launch.
template< typename V, typename X, typename C, bool MAPPED, bool WHOLE, int TILE >
__global__ void __launch_bounds__( 256 ) csr_tiled( const int32_t* __restrict__ off,
                                                    const int32_t* __restrict__ tile_rows, int64_t tiles, int64_t nnz,
                                                    const int32_t* __restrict__ row_map, int32_t* __restrict__ carry_row,
                                                    C* __restrict__ carry_val, Args< V, X, C > a, int32_t tail_max )
{
   __shared__ C lds[ 4 ][ TILE + ( WHOLE ? kWholeRowMax : 0 ) ];
   const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
   const int64_t t = xcd_contiguous( blockIdx.x, gridDim.x ) * 4 + wave;
   if( t >= tiles )
      return;
   // is the row at position r (MAPPED: position in the sorted order) written?
   auto in_range = [ & ]( int64_t r ) -> bool {
      const int64_t o = MAPPED ? (int64_t) row_map[ r ] : r;
      return o >= a.begin && o < a.end;
   };
   csr_tile< V, X, C, WHOLE, TILE >(
      off, a.n_rows, a.cols, a.vals, a.x, t, tile_rows[ t ], tile_rows[ t + 1 ], nnz, lds[ wave ], lane,
      [ & ]( int64_t r, C v ) {
         if( in_range( r ) )
            keep( a.y, MAPPED ? (int64_t) row_map[ r ] : r, v, a.alpha, a.beta );
      },
      [ & ]( int64_t rc, C v ) {
         carry_row[ t ] = rc >= 0 && in_range( rc ) ? (int32_t) rc : -1;
         carry_val[ t ] = v;
      },
      tail_max );
}

// Adds the carried partial sums of rows that span several tiles, in tile
// order (deterministic).  One thread per tile; the first tile of each run of
// equal carry rows sums the run.
template< typename X, typename C, bool MAPPED >
__global__ void __launch_bounds__( 256 ) csr_carry_fixup( int64_t tiles, const int32_t* __restrict__ carry_row,
                                                          const C* __restrict__ carry_val,
                                                          const int32_t* __restrict__ row_map, X* __restrict__ y,
                                                          C alpha )
{
   const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( t >= tiles )
      return;
   const int32_t r = carry_row[ t ];
   if( r < 0 || ( t > 0 && carry_row[ t - 1 ] == r ) )
      return;
   C sum = 0;
   for( int64_t u = t; u < tiles && carry_row[ u ] == r; u++ )
      sum = sum + carry_val[ u ];
   const int64_t o = MAPPED ? (int64_t) row_map[ r ] : r;
   y[ o ] = X( (C) y[ o ] + alpha * sum );
}

// ------------------------------------------------------------- AdaptiveCSR
// One wavefront per descriptor (ReducingKernels_AdaptiveCSR.h:18-134):
// STREAM stages the block's products in LDS and reduces each row from there;
// VECTOR is a wavefront per row; LONG spreads one row over the wavefronts of
// a 256-thread workgroup (the descriptors are padded to the workgroup end) and
// combines their partials in LDS.
template< typename V, typename X, typename C >
__global__ void __launch_bounds__( 256 ) csr_adaptive( const int32_t* __restrict__ off,
                                                       const int32_t* __restrict__ blocks, int64_t nblocks,
                                                       Args< V, X, C > a )
{
   constexpr int STREAM_CAP = 24576 / (int) sizeof( C ) / 8;  // host-pass limit per descriptor
   __shared__ C stream[ 4 ][ STREAM_CAP ];
   __shared__ C part[ 4 ];
   const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
   const int64_t bi = (int64_t) blockIdx.x * 4 + wave;
   const bool active = bi < nblocks - 1;  // the last descriptor only terminates the list
   int type = -1;
   if( lane == 0 )
      part[ wave ] = 0;
   __syncthreads();
   int64_t first = 0;
   if( active ) {
      const int32_t* d = blocks + 6 * bi;
      type = d[ 0 ];
      first = d[ 1 ];
      const int64_t b = off[ first ];
      if( type == 1 ) {  // STREAM
         const int32_t size = d[ 2 ], nseg = d[ 3 ];
         C* s = stream[ wave ];
         for( int32_t i = lane; i < size; i += 64 )
            s[ i ] = fetch_nt< V, X, C >( a.cols, a.vals, a.x, b + i );
         wave_sync();
         for( int64_t r = first + lane; r < first + nseg; r += 64 ) {
            C acc = 0;
            for( int64_t i = off[ r ] - b; i < off[ r + 1 ] - b; i++ )
               acc = acc + s[ i ];
            if( r >= a.begin && r < a.end )
               keep( a.y, r, acc, a.alpha, a.beta );
         }
      }
      else if( type == 2 ) {  // VECTOR
         C acc = 0;
         for( int64_t g = b + lane; g < b + d[ 2 ]; g += 64 )
            acc += fetch_nt< V, X, C >( a.cols, a.vals, a.x, g );
         acc = group_sum< 64 >( acc );
         if( lane == 0 && first >= a.begin && first < a.end )
            keep( a.y, first, acc, a.alpha, a.beta );
      }
      else {  // LONG
         const int widx = d[ 4 ], wcnt = d[ 5 ];
         C acc = 0;
         for( int64_t g = b + lane + 64 * widx; g < off[ first + 1 ]; g += 64 * wcnt )
            acc += fetch_nt< V, X, C >( a.cols, a.vals, a.x, g );
         acc = group_sum< 64 >( acc );
         if( lane == 0 )
            part[ wave ] = acc;
      }
   }
   __syncthreads();
   if( type == 0 && lane == 0 && blocks[ 6 * bi + 4 ] == 0 ) {
      const C s = part[ 0 ] + part[ 1 ] + part[ 2 ] + part[ 3 ];
      if( first >= a.begin && first < a.end )
         keep( a.y, first, s, a.alpha, a.beta );
   }
}

// ================================================================= Ellpack
// Column-major (the GPU default organisation, ElementsOrganization.h:22-32):
// one thread per row, slot k of consecutive rows is contiguous.
template< typename V, typename X, typename C >
__global__ void __launch_bounds__( 256 ) ellpack_cm( int32_t width, int32_t aligned, Args< V, X, C > a )
{
   const int64_t r = a.begin + (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( r >= a.end )
      return;
   const C acc = strided_sum< V, X, C >( a.cols, a.vals, a.x, r, aligned, width );
   keep( a.y, r, acc, a.alpha, a.beta );
}

// Row-major: G lanes per row over a contiguous row of `width` slots
// (ReducingKernels_Ellpack.h:12-71).
template< typename V, typename X, typename C, int G >
__global__ void __launch_bounds__( 256 ) ellpack_rm( int32_t width, Args< V, X, C > a )
{
   const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   const int64_t r = a.begin + t / G;
   const int lane = (int) ( t % G );
   if( r >= a.end )
      return;
   C acc = 0;
   for( int32_t k = lane; k < width; k += G )
      acc += fetch_nt< V, X, C >( a.cols, a.vals, a.x, r * (int64_t) width + k );
   acc = group_sum< G >( acc );
   if( lane == 0 )
      keep( a.y, r, acc, a.alpha, a.beta );
}

// =========================================================== SlicedEllpack
// Column-major: thread per row over the slice width (every segment of a slice
// has the slice's size, SlicedEllpackBase.hpp:79-92; padding slots cost only
// their column index, see fetch_nt) (ReducingKernels_SlicedEllpack.h:88-~320).
template< typename V, typename X, typename C >
__global__ void __launch_bounds__( 256 ) sliced_cm( int32_t S, const int32_t* __restrict__ soff,
                                                    const int32_t* __restrict__ ssz, Args< V, X, C > a )
{
   const int64_t r = a.begin + (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( r >= a.end )
      return;
   const int64_t s = r / S;
   const int64_t base = soff[ s ] + r % S;
   const C acc = strided_sum< V, X, C >( a.cols, a.vals, a.x, base, S, ssz[ s ] );
   keep( a.y, r, acc, a.alpha, a.beta );
}

// Row-major: G lanes per row (ReducingKernels_SlicedEllpack.h:14-86).
template< typename V, typename X, typename C, int G >
__global__ void __launch_bounds__( 256 ) sliced_rm( int32_t S, const int32_t* __restrict__ soff,
                                                    const int32_t* __restrict__ ssz, Args< V, X, C > a )
{
   const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   const int64_t r = a.begin + t / G;
   const int lane = (int) ( t % G );
   if( r >= a.end )
      return;
   const int64_t s = r / S;
   const int32_t w = ssz[ s ];
   const int64_t base = soff[ s ] + ( r % S ) * (int64_t) w;
   C acc = 0;
   for( int32_t k = lane; k < w; k += G )
      acc += fetch_nt< V, X, C >( a.cols, a.vals, a.x, base + k );
   acc = group_sum< G >( acc );
   if( lane == 0 )
      keep( a.y, r, acc, a.alpha, a.beta );
}

// =============================================================== BiEllpack
// One 32-lane half-wavefront per strip of 32 rows; lane p ends with the row at
// sorted position p (group g holds the first 2^(5-g) positions), and the
// strip-local inverse permutation (built in LDS from perm) routes the result
// to its row (ReducingKernels_BiEllpack.h:14-226).  Column-major groups are
// swept by the whole half-wave (below); row-major rows by their own lane.
template< typename V, typename X, typename C >
__global__ void __launch_bounds__( 256 ) biellpack( int64_t n, int64_t strips, int col_major,
                                                    const int32_t* __restrict__ perm,
                                                    const int32_t* __restrict__ gptr, Args< V, X, C > a )
{
   constexpr int W = 32, LOGW = 5;
   __shared__ int32_t rowof[ 8 ][ W ];
   const int half = threadIdx.x >> 5, p = threadIdx.x & 31;
   const int64_t strip = (int64_t) blockIdx.x * 8 + half;
   const bool live = strip < strips;
   const int64_t b = strip * W;
   if( live && b + p < n )
      rowof[ half ][ perm[ b + p ] - b ] = p;
   __syncthreads();
   if( ! live )
      return;
   const int64_t cnt = min( (int64_t) W, n - b );
   const int32_t* g = gptr + strip * ( LOGW + 1 );
   C acc = 0;
   if( col_major ) {
      // all 32 lanes sweep each group's height x width slots with stride 32:
      // slot s holds position s % height, so lane p always lands on position
      // p % height (height divides 32) and every load of the half-wave is
      // contiguous; an xor butterfly over the lanes of one position then
      // gives lane p < height its position's sum (a fixed order).  Lanes past
      // the strip's last row only read padding slots (column -1).
      int32_t height = W;
      for( int gi = 0; gi <= LOGW; gi++, height >>= 1 ) {
         const int32_t gs = g[ gi + 1 ] - g[ gi ];
         if( gs == 0 )
            continue;  // uniform over the half-wave
         const int64_t base = g[ gi ];
         C part = gs > p ? strided_sum< V, X, C >( a.cols, a.vals, a.x, base + p, W, ( gs - p + W - 1 ) / W ) : C( 0 );
         for( int o = height; o < W; o <<= 1 )
            part += __shfl_xor( part, o, W );
         if( p < height )
            acc += part;
      }
   }
   else {
      if( p >= cnt )
         return;
      int32_t height = W;
      for( int gi = 0; gi <= LOGW && p < height; gi++, height >>= 1 ) {
         const int32_t gs = g[ gi + 1 ] - g[ gi ];
         if( gs == 0 )
            continue;
         const int32_t w = gs / height;
         acc += strided_sum< V, X, C >( a.cols, a.vals, a.x, g[ gi ] + (int64_t) p * w, 1, w );
      }
   }
   if( p >= cnt )
      return;
   const int64_t r = b + rowof[ half ][ p ];
   if( r >= a.begin && r < a.end )
      keep( a.y, r, acc, a.alpha, a.beta );
}

// ========================================================== ChunkedEllpack
// One workgroup of 256 threads per slice, one thread per chunk; chunk sums go
// to LDS and each row then adds its contiguous run of chunks
// (ReducingKernels_ChunkedEllpack.h:14-~104).
template< typename V, typename X, typename C >
__global__ void __launch_bounds__( 256 ) chunked( int32_t chunks_in_slice, int col_major,
                                                  const int32_t* __restrict__ slices,
                                                  const int32_t* __restrict__ seg_to_chunk, Args< V, X, C > a )
{
   __shared__ C csum[ 256 ];
   const int64_t si = blockIdx.x;
   const int32_t size = slices[ 4 * si ], cs = slices[ 4 * si + 1 ], first = slices[ 4 * si + 2 ],
                 ptr = slices[ 4 * si + 3 ];
   const int q = threadIdx.x;
   C acc = 0;
   if( q < chunks_in_slice ) {
      if( col_major )
         acc = strided_sum< V, X, C >( a.cols, a.vals, a.x, ptr + q, chunks_in_slice, cs );
      else
         acc = strided_sum< V, X, C >( a.cols, a.vals, a.x, ptr + (int64_t) q * cs, 1, cs );
   }
   csum[ q ] = acc;
   __syncthreads();
   if( q < size ) {
      const int64_t r = first + q;
      const int32_t c0 = q ? seg_to_chunk[ r - 1 ] : 0, c1 = seg_to_chunk[ r ];
      C s = 0;
      for( int32_t c = c0; c < c1; c++ )
         s = s + csum[ c ];
      if( r >= a.begin && r < a.end )
         keep( a.y, r, s, a.alpha, a.beta );
   }
}

// ================================================================ dispatch
inline dim3 grid_for( int64_t threads, int block = 256 )
{
   return dim3( (unsigned) std::max< int64_t >( 1, ceil_div( threads, block ) ) );
}

template< typename V, typename X, typename C >
void csr_fixed( int tps, const int32_t* off, const Args< V, X, C >& a, hipStream_t s )
{
   const int64_t rows = a.end - a.begin;
   switch( tps ) {
      case 1:
         hipLaunchKernelGGL( ( csr_scalar< V, X, C > ), grid_for( rows ), dim3( 256 ), 0, s, off, a );
         return;
#define TNLMI_VEC( G )                                                                                              \
   case G:                                                                                                          \
      hipLaunchKernelGGL( ( csr_vector< V, X, C, G > ), grid_for( rows * G ), dim3( 256 ), 0, s, off, a );         \
      return;
         TNLMI_VEC( 2 )
         TNLMI_VEC( 4 )
         TNLMI_VEC( 8 )
         TNLMI_VEC( 16 )
         TNLMI_VEC( 32 )
         TNLMI_VEC( 64 )
#undef TNLMI_VEC
      case 128:
         hipLaunchKernelGGL( ( csr_multivector< V, X, C, 128 > ), grid_for( rows * 128 ), dim3( 256 ), 0, s, off, a );
         return;
      case 256:
         hipLaunchKernelGGL( ( csr_multivector< V, X, C, 256 > ), grid_for( rows * 256 ), dim3( 256 ), 0, s, off, a );
         return;
   }
   // ReducingOperations_CSR.h:294-298
   fail( TNLMI_RUNTIME_ERROR, "vectorProduct: unsupported number of threads per segment " + std::to_string( tps ) );
}

template< typename V, typename X, typename C >
void csr_tiled_launch( const tnlmi_segments& seg, const int32_t* off, const int32_t* row_map, const Args< V, X, C >& a,
                       hipStream_t s )
{
   const dim3 grid( (unsigned) std::max< int64_t >( 1, ceil_div( seg.tiles, 4 ) ) );
   const dim3 fgrid( (unsigned) std::max< int64_t >( 1, ceil_div( seg.tiles, 256 ) ) );
   C* cv = reinterpret_cast< C* >( seg.carry_val.ptr );
   const int32_t tail_max = std::max( 0, seg.max_size - 1 );
   if( seg.tile_nnz != kTileNnz && ! ( seg.tile_nnz == kTileNnzSmall && seg.max_size <= kWholeRowMax ) )
      fail( TNLMI_RUNTIME_ERROR, "vectorProduct: inconsistent CSR tile table" );
   if( seg.max_size <= kWholeRowMax ) {  // one launch, no carries
      if( seg.tile_nnz == kTileNnzSmall ) {
         if( row_map )
            hipLaunchKernelGGL( ( csr_tiled< V, X, C, true, true, kTileNnzSmall > ), grid, dim3( 256 ), 0, s, off,
                                seg.tile_rows.ptr, seg.tiles, seg.elements, row_map, seg.carry_row.ptr, cv, a,
                                tail_max );
         else
            hipLaunchKernelGGL( ( csr_tiled< V, X, C, false, true, kTileNnzSmall > ), grid, dim3( 256 ), 0, s, off,
                                seg.tile_rows.ptr, seg.tiles, seg.elements, row_map, seg.carry_row.ptr, cv, a,
                                tail_max );
      }
      else if( row_map )
         hipLaunchKernelGGL( ( csr_tiled< V, X, C, true, true, kTileNnz > ), grid, dim3( 256 ), 0, s, off,
                             seg.tile_rows.ptr, seg.tiles, seg.elements, row_map, seg.carry_row.ptr, cv, a, tail_max );
      else
         hipLaunchKernelGGL( ( csr_tiled< V, X, C, false, true, kTileNnz > ), grid, dim3( 256 ), 0, s, off,
                             seg.tile_rows.ptr, seg.tiles, seg.elements, row_map, seg.carry_row.ptr, cv, a, tail_max );
   }
   else if( row_map ) {
      hipLaunchKernelGGL( ( csr_tiled< V, X, C, true, false, kTileNnz > ), grid, dim3( 256 ), 0, s, off,
                          seg.tile_rows.ptr, seg.tiles, seg.elements, row_map, seg.carry_row.ptr, cv, a, 0 );
      hipLaunchKernelGGL( ( csr_carry_fixup< X, C, true > ), fgrid, dim3( 256 ), 0, s, seg.tiles, seg.carry_row.ptr,
                          cv, row_map, a.y, a.alpha );
   }
   else {
      hipLaunchKernelGGL( ( csr_tiled< V, X, C, false, false, kTileNnz > ), grid, dim3( 256 ), 0, s, off,
                          seg.tile_rows.ptr, seg.tiles, seg.elements, row_map, seg.carry_row.ptr, cv, a, 0 );
      hipLaunchKernelGGL( ( csr_carry_fixup< X, C, false > ), fgrid, dim3( 256 ), 0, s, seg.tiles, seg.carry_row.ptr,
                          cv, row_map, a.y, a.alpha );
   }
}

template< typename V, typename X, typename C >
void spmv_typed( const tnlmi_matrix& A, const void* x, void* y, double alpha, double beta, int64_t begin, int64_t end,
                 const tnlmi_launch& cfg, hipStream_t s )
{
   const tnlmi_segments& seg = *A.segments;
   Args< V, X, C > a{ A.column_indexes, (const V*) A.values, (const X*) x, (X*) y, (C) alpha, (C) beta, begin, end,
                      A.rows };
   const int64_t rows = end - begin;
   const int tps = cfg.threads_per_segment;
   switch( seg.format ) {
      case TNLMI_CSR: {
         const int32_t* off = seg.offsets.ptr;
         switch( cfg.mapping ) {
            case TNLMI_MAP_AUTO:
               csr_tiled_launch( seg, off, nullptr, a, s );
               break;
            case TNLMI_MAP_FIXED:
               csr_fixed( tps, off, a, s );
               break;
            case TNLMI_MAP_WARP:
               csr_fixed( 64, off, a, s );
               break;
            case TNLMI_MAP_DYNAMIC_GROUPING:
               if( tps < 1 || tps > 256 )
                  fail( TNLMI_RUNTIME_ERROR, "vectorProduct: unsupported number of threads per segment" );
               hipLaunchKernelGGL( ( csr_dynamic< V, X, C > ), grid_for( rows, 256 / tps ), dim3( 256 ), 0, s, off, a,
                                   tps );
               break;
            case TNLMI_MAP_BLOCK:
               csr_fixed( 256, off, a, s );
               break;
            case TNLMI_MAP_BLOCK_MERGED:
               csr_tiled_launch( seg, off, nullptr, a, s );
               break;
            default:
               fail( TNLMI_INVALID_ARGUMENT, "vectorProduct: unknown threads-to-segments mapping" );
         }
         break;
      }
      case TNLMI_ADAPTIVE_CSR: {
         const int l = sizeof( C ) == 8 ? 3 : 2;
         if( seg.nblocks[ l ] > 1 )
            hipLaunchKernelGGL( ( csr_adaptive< V, X, C > ), grid_for( seg.nblocks[ l ] * 64 ), dim3( 256 ), 0, s,
                                seg.offsets.ptr, seg.blocks[ l ].ptr, seg.nblocks[ l ], a );
         else if( rows > 0 )  // no descriptors (all segments empty): plain CSR
            csr_fixed( 1, seg.offsets.ptr, a, s );
         break;
      }
      case TNLMI_SORTED_CSR: {
         Args< V, X, C > b = a;
         csr_tiled_launch( seg, seg.sorted_offsets.ptr, seg.inv_perm.ptr, b, s );
         break;
      }
      case TNLMI_ELLPACK:
         if( seg.org == TNLMI_COLUMN_MAJOR )
            hipLaunchKernelGGL( ( ellpack_cm< V, X, C > ), grid_for( rows ), dim3( 256 ), 0, s, seg.ell_width,
                                seg.ell_aligned, a );
         else {
            const int w = seg.ell_width;
            if( w <= 4 )
               hipLaunchKernelGGL( ( ellpack_rm< V, X, C, 4 > ), grid_for( rows * 4 ), dim3( 256 ), 0, s, w, a );
            else if( w <= 16 )
               hipLaunchKernelGGL( ( ellpack_rm< V, X, C, 16 > ), grid_for( rows * 16 ), dim3( 256 ), 0, s, w, a );
            else
               hipLaunchKernelGGL( ( ellpack_rm< V, X, C, 64 > ), grid_for( rows * 64 ), dim3( 256 ), 0, s, w, a );
         }
         break;
      case TNLMI_SLICED_ELLPACK:
         if( seg.org == TNLMI_COLUMN_MAJOR )
            hipLaunchKernelGGL( ( sliced_cm< V, X, C > ), grid_for( rows ), dim3( 256 ), 0, s, seg.param,
                                seg.slice_offsets.ptr, seg.slice_seg_sizes.ptr, a );
         else
            hipLaunchKernelGGL( ( sliced_rm< V, X, C, 8 > ), grid_for( rows * 8 ), dim3( 256 ), 0, s, seg.param,
                                seg.slice_offsets.ptr, seg.slice_seg_sizes.ptr, a );
         break;
      case TNLMI_BIELLPACK:
         if( seg.param != 32 )
            fail( TNLMI_NOT_IMPLEMENTED, "BiEllpack SpMV is implemented for WarpSize 32" );
         if( seg.strips )
            hipLaunchKernelGGL( ( biellpack< V, X, C > ), dim3( (unsigned) ceil_div( seg.strips, 8 ) ), dim3( 256 ), 0,
                                s, seg.n, seg.strips, seg.org == TNLMI_COLUMN_MAJOR ? 1 : 0, seg.perm.ptr,
                                seg.group_ptrs.ptr, a );
         break;
      case TNLMI_CHUNKED_ELLPACK:
         if( seg.ch_nslices )
            hipLaunchKernelGGL( ( chunked< V, X, C > ), dim3( (unsigned) seg.ch_nslices ), dim3( 256 ), 0, s,
                                seg.chunks_in_slice, seg.org == TNLMI_COLUMN_MAJOR ? 1 : 0, seg.ch_slices.ptr,
                                seg.seg_to_chunk.ptr, a );
         break;
      default:
         fail( TNLMI_INVALID_ARGUMENT, "vectorProduct: unknown segments format" );
   }
   TNLMI_HIP( hipGetLastError() );
}

}  // namespace

void spmv_dispatch( const tnlmi_matrix& A, tnlmi_dtype xt, const void* x, void* y, tnlmi_dtype ct, double alpha,
                    double beta, int64_t begin, int64_t end, const tnlmi_launch& cfg, hipStream_t s )
{
   if( A.matrix_type == TNLMI_SYMMETRIC ) {
      spmv_symmetric_dispatch( A, xt, x, y, ct, alpha, beta, begin, end, s );
      return;
   }
   const tnlmi_dtype vt = A.value_type;
   if( vt == TNLMI_F64 && xt == TNLMI_F64 && ct == TNLMI_F64 )
      spmv_typed< double, double, double >( A, x, y, alpha, beta, begin, end, cfg, s );
   else if( vt == TNLMI_F32 && xt == TNLMI_F32 && ct == TNLMI_F32 )
      spmv_typed< float, float, float >( A, x, y, alpha, beta, begin, end, cfg, s );
   else if( vt == TNLMI_F32 && xt == TNLMI_F64 && ct == TNLMI_F64 )
      spmv_typed< float, double, double >( A, x, y, alpha, beta, begin, end, cfg, s );
   else if( vt == TNLMI_BF16 && xt == TNLMI_F32 && ct == TNLMI_F32 )
      spmv_typed< bf16, float, float >( A, x, y, alpha, beta, begin, end, cfg, s );
   else if( vt == TNLMI_BF16 && xt == TNLMI_F64 && ct == TNLMI_F64 )
      spmv_typed< bf16, double, double >( A, x, y, alpha, beta, begin, end, cfg, s );
   else
      fail( TNLMI_NOT_IMPLEMENTED, "vectorProduct: unsupported (value, vector, compute) type combination" );
}

}  // namespace tnlmi

using namespace tnlmi;

extern "C" tnlmi_status tnlmi_spmv( const tnlmi_matrix* A, tnlmi_dtype x_type, const void* d_x, void* d_y,
                                    tnlmi_dtype compute_type, double alpha, double beta, int64_t begin, int64_t end,
                                    const tnlmi_launch* launch, tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( ! A || ! A->segments )
         fail( TNLMI_INVALID_ARGUMENT, "vectorProduct: matrix or its segments is NULL" );
      if( A->rows != A->segments->n )
         fail( TNLMI_INVALID_ARGUMENT, "vectorProduct: matrix rows do not match the segments count" );
      if( end == 0 )
         end = A->rows;  // SparseMatrixBase.hpp:297-298
      if( begin < 0 || end > A->rows || begin > end )
         fail( TNLMI_INVALID_ARGUMENT, "vectorProduct: row range [begin, end) is out of bounds" );
      if( end == begin && A->matrix_type != TNLMI_SYMMETRIC )
         return;  // (a symmetric product still scales y by beta)
      if( ( A->columns && ! d_x ) || ( A->rows && ! d_y ) )
         fail( TNLMI_INVALID_ARGUMENT, "vectorProduct: vector pointer is NULL" );
      if( A->segments->storage && ( ! A->values || ! A->column_indexes ) )
         fail( TNLMI_INVALID_ARGUMENT, "vectorProduct: matrix storage pointer is NULL" );
      tnlmi_launch cfg{ TNLMI_MAP_AUTO, 1 };
      if( launch )
         cfg = *launch;
      if( cfg.threads_per_segment <= 0 || ( cfg.threads_per_segment & ( cfg.threads_per_segment - 1 ) ) )
         fail( TNLMI_INVALID_ARGUMENT, "Threads per segment count must be a positive power of two." );
      hipStream_t s = as_stream( stream );
      if( cfg.mapping == TNLMI_MAP_AUTO )
         if( const tnlmi_spmv_plan* P = auto_plan( *A, x_type, compute_type, s ) ) {
            plan_execute( *P, x_type, d_x, d_y, compute_type, alpha, beta, begin, end, s );
            return;
         }
      spmv_dispatch( *A, x_type, d_x, d_y, compute_type, alpha, beta, begin, end, cfg, s );
   } );
}
