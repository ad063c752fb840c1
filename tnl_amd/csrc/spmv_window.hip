// spmv_window.hip -- the windowed SpMV plan of a CSR matrix: every x gather
// from LDS, every HBM access a stream.
//
// Random column indexes make every nonzero an 8-byte gather into x; through
// the L2 each one costs a 128-byte line and an L2 request.  The plan splits
// the columns into windows of kWin = 8192 entries (64 KiB of x, one LDS
// allocation) and the rows into blocks of at most 8160 (FP64) / 16320 (FP32)
// nonzeros:
//
//   phase 1  one workgroup per window: x[window] -> LDS, then stream the
//            window's nonzeros (values + 13-bit local columns, sorted by row)
//            and write each product v * x[c] to prod[] (window-major)
//   phase 2  one workgroup per row block: copy the block's products from each
//            window's contiguous segment into LDS at their CSR position
//            (16-bit slot), then sum every row in CSR order
//
// Per nonzero: 8 + 2 bytes read and 8 written in phase 1, 8 + 2 read in
// phase 2 -- about 1.9 GB for the 67 M-nonzero power-law matrix, all of it
// coalesced, and no gather leaves LDS.  Products are rounded to the compute
// type and each row is summed from zero left to right in column order, as
// TNL's host loop does (ReducingOperations_CSR.h:46-74 with
// SparseMatrixBase.hpp:331-389), so every row of at most kLongRow = 256
// nonzeros, summed by one thread, matches it bit for bit; longer rows are
// summed by a wave in a fixed (deterministic) order.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "segments.hpp"
#include "spmv.hpp"
#include "spmv_window.hpp"

namespace tnlmi {
namespace {

constexpr int kWinBits = 13;
constexpr int kWin = 1 << kWinBits;  // x entries per window
constexpr int kP1 = 1024;            // phase-1 threads
constexpr int kP2 = 512;             // phase-2 threads
constexpr int kLongRow = 256;        // longer rows are summed by a whole wave
constexpr int kSegU = 24;            // phase-2 loads in flight per thread (and array padding)

// row-block capacity: the phase-2 LDS buffer within 64 KiB
template< typename C >
constexpr int cap_of()
{
   return sizeof( C ) == 8 ? 8160 : 16320;
}

__global__ void w_erow( int64_t n, const int32_t* __restrict__ off, int32_t* __restrict__ erow )
{
   const int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( r >= n )
      return;
   for( int32_t e = off[ r ]; e < off[ r + 1 ]; e++ )
      erow[ e ] = (int32_t) r;
}

// window of each nonzero (padding slots, column -1, go to an extra window K)
__global__ void w_key( int64_t nnz, const int32_t* __restrict__ cols, int32_t K, int32_t* __restrict__ key,
                       int32_t* __restrict__ ids )
{
   const int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( e >= nnz )
      return;
   const int32_t c = cols[ e ];
   key[ e ] = c >= 0 ? ( c >> kWinBits ) : K;
   ids[ e ] = (int32_t) e;
}

// wstart[k] = first sorted position of window k (k = 0..K+1)
__global__ void w_starts( int64_t nnz, int32_t K, const int32_t* __restrict__ skey, int64_t* __restrict__ wstart )
{
   const int64_t p = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( p > nnz )
      return;
   const int32_t prev = p == 0 ? -1 : skey[ p - 1 ];
   const int32_t cur = p == nnz ? K + 1 : skey[ p ];
   for( int32_t k = prev + 1; k <= cur; k++ )
      wstart[ k ] = p;
}

// the sorted arrays: value, local column, CSR slot within the row block, row
template< typename V >
__global__ void w_fill( int64_t nnz, const int32_t* __restrict__ perm, const int32_t* __restrict__ cols,
                        const V* __restrict__ vals, const int32_t* __restrict__ erow, const int32_t* __restrict__ rblk,
                        const int64_t* __restrict__ eb, V* __restrict__ pv, uint16_t* __restrict__ pc,
                        uint16_t* __restrict__ dest, int32_t* __restrict__ srow )
{
   const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( i >= nnz )
      return;
   const int32_t e = perm[ i ];
   const int32_t r = erow[ e ];
   const int32_t c = cols[ e ];
   pv[ i ] = vals[ e ];
   pc[ i ] = (uint16_t) ( c >= 0 ? ( c & ( kWin - 1 ) ) : 0 );
   dest[ i ] = (uint16_t) ( e - eb[ rblk[ r ] ] );
   srow[ i ] = r;
}

__global__ void w_rblk( int64_t B, const int32_t* __restrict__ rb, int32_t* __restrict__ rblk )
{
   const int64_t b = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( b >= B )
      return;
   for( int32_t r = rb[ b ]; r < rb[ b + 1 ]; r++ )
      rblk[ r ] = (int32_t) b;
}

// seg[b * K + k] = first position of window k whose row is >= rb[b]
__global__ void w_seg( int64_t B, int32_t K, const int32_t* __restrict__ rb, const int64_t* __restrict__ wstart,
                       const int32_t* __restrict__ srow, int64_t* __restrict__ seg )
{
   const int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( q >= ( B + 1 ) * K )
      return;
   const int64_t b = q / K;
   const int32_t k = (int32_t) ( q % K );
   int64_t lo = wstart[ k ], hi = wstart[ k + 1 ];
   const int32_t row = rb[ b ];
   while( lo < hi ) {
      const int64_t mid = ( lo + hi ) / 2;
      if( srow[ mid ] < row )
         lo = mid + 1;
      else
         hi = mid;
   }
   seg[ q ] = lo;
}

// phase 1: prod[i] = v[i] * x[window base + pc[i]] for the window's nonzeros
template< typename V, typename C >
__global__ void __launch_bounds__( kP1 ) w_phase1( int64_t ncols, const int64_t* __restrict__ wstart,
                                                   const V* __restrict__ pv, const uint16_t* __restrict__ pc,
                                                   const C* __restrict__ x, C* __restrict__ prod )
{
   __shared__ C xw[ kWin ];
   const int64_t k = blockIdx.x;
   const int64_t base = k * kWin;
   for( int t = threadIdx.x; t < kWin; t += kP1 )
      xw[ t ] = base + t < ncols ? x[ base + t ] : (C) 0;
   __syncthreads();
   const int64_t a = wstart[ k ], b = wstart[ k + 1 ];
   constexpr int U = 4;
   for( int64_t i0 = a + threadIdx.x; i0 < b; i0 += U * kP1 ) {
      C p[ U ];
#pragma unroll
      for( int u = 0; u < U; u++ ) {
         const int64_t i = i0 + u * kP1;
         p[ u ] = i < b ? (C) __builtin_nontemporal_load( pv + i ) * xw[ __builtin_nontemporal_load( pc + i ) ] : (C) 0;
      }
#pragma unroll
      for( int u = 0; u < U; u++ ) {
         const int64_t i = i0 + u * kP1;
         if( i < b )
            __builtin_nontemporal_store( p[ u ], prod + i );
      }
   }
}

// phase 2: row block b -> LDS in CSR order -> row sums -> y.  Every global
// load whose address is known at entry (the first kSegU terms of the thread's
// first kJ segments, the bounds of its first kJ rows) is issued before the
// LDS buffer is cleared, so the block waits for about one memory round trip.
constexpr int kJ = 1;

template< typename C >
__device__ __forceinline__ C sum_row( const C* buf, int32_t q0, int32_t q1 )
{
   // from zero, left to right; the LDS reads of eight terms ahead of their adds
   C sum = 0;
   int32_t q = q0;
   for( ; q + 8 <= q1; q += 8 ) {
      C t[ 8 ];
#pragma unroll
      for( int u = 0; u < 8; u++ )
         t[ u ] = buf[ q + u ];
#pragma unroll
      for( int u = 0; u < 8; u++ )
         sum += t[ u ];
   }
   for( ; q < q1; q++ )
      sum += buf[ q ];
   return sum;
}

// the stores of one step of a segment copy: terms i0 + u < z
template< typename C >
__device__ __forceinline__ void put_terms( C* buf, const C ( &v )[ kSegU ], const uint32_t ( &d )[ kSegU ],
                                           int64_t i0, int64_t z )
{
#pragma unroll
   for( int u = 0; u < kSegU; u++ )
      if( i0 + u < z )
         buf[ d[ u ] ] = v[ u ];
}

// one step of a segment copy: all kSegU loads are issued unconditionally
// (prod and dest carry kSegU entries of padding, so reading past a segment
// stays in bounds, and what is read there is not stored).  Ordinary cached
// loads: a thread's loads of one segment share cache lines in L1, and
// neighbouring row blocks' segments share lines in L2.
template< typename C >
__device__ __forceinline__ void get_terms( const C* prod, const uint16_t* dest, int64_t i0, C ( &v )[ kSegU ],
                                           uint32_t ( &d )[ kSegU ] )
{
#pragma unroll
   for( int u = 0; u < kSegU; u++ ) {
      v[ u ] = prod[ i0 + u ];
      d[ u ] = dest[ i0 + u ];
   }
}

template< typename C, int CAP >
__global__ void __launch_bounds__( kP2 )
   w_phase2( int32_t K, const int64_t* __restrict__ seg, const C* __restrict__ prod,
             const uint16_t* __restrict__ dest, const int32_t* __restrict__ rb, const int32_t* __restrict__ off,
             const int32_t* __restrict__ lstart, const int32_t* __restrict__ lrows, C* __restrict__ y, C alpha,
             C beta, int64_t begin, int64_t end )
{
   __shared__ C buf[ CAP ];
   // consecutive row blocks on one XCD: their chunks of a window's products
   // (about 120 bytes each, 30 of slot indexes) share cache lines in its L2
   const int64_t b = xcd_contiguous( blockIdx.x, gridDim.x );
   const int32_t r0 = rb[ b ], r1 = rb[ b + 1 ];
   if( r1 <= begin || r0 >= end )
      return;  // uniform: no row of this block is asked for
   const int t = threadIdx.x;
   int64_t a[ kJ ], z[ kJ ];
   C v[ kJ ][ kSegU ];
   uint32_t d[ kJ ][ kSegU ];
#pragma unroll
   for( int j = 0; j < kJ; j++ ) {
      const int k = t + j * kP2;
      a[ j ] = k < K ? seg[ b * K + k ] : 0;
      z[ j ] = k < K ? seg[ ( b + 1 ) * K + k ] : 0;
   }
#pragma unroll
   for( int j = 0; j < kJ; j++ )
      get_terms( prod, dest, a[ j ], v[ j ], d[ j ] );
   int32_t ob[ kJ ], oe[ kJ ];
#pragma unroll
   for( int j = 0; j < kJ; j++ ) {
      const int32_t r = r0 + t + j * kP2;
      ob[ j ] = r < r1 ? off[ r ] : 0;
      oe[ j ] = r < r1 ? off[ r + 1 ] : 0;
   }
   const int32_t e0 = off[ r0 ], e1 = off[ r1 ];
   // padding slots (column -1) have no product: they read as zero
   for( int32_t q = t; q < e1 - e0; q += kP2 )
      buf[ q ] = (C) 0;
   __syncthreads();
#pragma unroll
   for( int j = 0; j < kJ; j++ )
      put_terms( buf, v[ j ], d[ j ], a[ j ], z[ j ] );
   // the rest: terms past the first kSegU of a segment, segments past the first kJ
#pragma unroll
   for( int j = 0; j < kJ; j++ )
      for( int64_t i0 = a[ j ] + kSegU; i0 < z[ j ]; i0 += kSegU ) {
         get_terms( prod, dest, i0, v[ 0 ], d[ 0 ] );
         put_terms( buf, v[ 0 ], d[ 0 ], i0, z[ j ] );
      }
   for( int k = t + kJ * kP2; k < K; k += kP2 ) {
      const int64_t ka = seg[ b * K + k ], kz = seg[ ( b + 1 ) * K + k ];
      for( int64_t i0 = ka; i0 < kz; i0 += kSegU ) {
         get_terms( prod, dest, i0, v[ 0 ], d[ 0 ] );
         put_terms( buf, v[ 0 ], d[ 0 ], i0, kz );
      }
   }
   __syncthreads();
   auto keep = [ & ]( int32_t r, C sum ) {
      if( r >= begin && r < end )
         y[ r ] = beta == (C) 0 ? alpha * sum : alpha * sum + beta * y[ r ];
   };
   // rows of at most kLongRow nonzeros: one thread each, summed from zero
   // left to right
#pragma unroll
   for( int j = 0; j < kJ; j++ ) {
      const int32_t r = r0 + t + j * kP2;
      if( r < r1 && oe[ j ] - ob[ j ] <= kLongRow )
         keep( r, sum_row( buf, ob[ j ] - e0, oe[ j ] - e0 ) );
   }
   for( int32_t r = r0 + t + kJ * kP2; r < r1; r += kP2 ) {
      const int32_t q0 = off[ r ] - e0, q1 = off[ r + 1 ] - e0;
      if( q1 - q0 <= kLongRow )
         keep( r, sum_row( buf, q0, q1 ) );
   }
   // longer rows (listed per block): one wave each, in turn over the waves;
   // lane l sums terms l, l + 64, ... and an xor butterfly adds the 64 parts
   // (a fixed order)
   const int lane = t & 63, wid = wave_uniform( (int32_t) ( t >> 6 ) );
   const int32_t l1 = lstart[ b + 1 ];
   for( int32_t i = lstart[ b ] + wid; i < l1; i += kP2 / 64 ) {
      const int32_t r = lrows[ i ];
      const int32_t q1 = off[ r + 1 ] - e0;
      C part = 0;
      for( int32_t q = off[ r ] - e0 + lane; q < q1; q += 64 )
         part += buf[ q ];
      const C sum = group_sum< 64 >( part );
      if( lane == 0 )
         keep( r, sum );
   }
}

template< typename V >
void build( WindowPlan& W, const tnlmi_matrix& A, hipStream_t s )
{
   const tnlmi_segments& seg = *A.segments;
   const int64_t n = A.rows, nnz = seg.elements;
   const int32_t* off = seg.offsets.ptr;
   const int32_t* cols = A.column_indexes;
   const V* vals = (const V*) A.values;
   W.K = (int32_t) std::max< int64_t >( 1, ceil_div( A.columns, kWin ) );
   // row blocks: consecutive rows with at most cap nonzeros together
   std::vector< int32_t > hoff( n + 1 );
   TNLMI_HIP( hipMemcpyAsync( hoff.data(), off, sizeof( int32_t ) * ( n + 1 ), hipMemcpyDeviceToHost, s ) );
   TNLMI_HIP( hipStreamSynchronize( s ) );
   const int64_t cap = W.cap;
   std::vector< int32_t > rb{ 0 };
   std::vector< int64_t > eb{ 0 };
   for( int64_t r = 0; r < n; ) {
      const int64_t start = hoff[ r ];
      int64_t e = r;
      while( e < n && hoff[ e + 1 ] - start <= cap )
         e++;
      if( e == r )
         fail( TNLMI_NOT_IMPLEMENTED, "windowed SpMV plan: a row is longer than the row-block capacity" );
      rb.push_back( (int32_t) e );
      eb.push_back( hoff[ e ] );
      r = e;
   }
   W.B = (int64_t) rb.size() - 1;
   // rows longer than kLongRow, by row block (summed by a wave in phase 2)
   std::vector< int32_t > lstart{ 0 }, lrows;
   for( int64_t i = 0; i < W.B; i++ ) {
      for( int32_t r = rb[ i ]; r < rb[ i + 1 ]; r++ )
         if( hoff[ r + 1 ] - hoff[ r ] > kLongRow )
            lrows.push_back( r );
      lstart.push_back( (int32_t) lrows.size() );
   }
   W.lstart.allocate( W.B + 1 );
   W.lrows.allocate( std::max< int64_t >( 1, (int64_t) lrows.size() ) );
   TNLMI_HIP( hipMemcpyAsync( W.lstart.ptr, lstart.data(), sizeof( int32_t ) * ( W.B + 1 ), hipMemcpyHostToDevice,
                              s ) );
   if( !lrows.empty() )
      TNLMI_HIP( hipMemcpyAsync( W.lrows.ptr, lrows.data(), sizeof( int32_t ) * lrows.size(), hipMemcpyHostToDevice,
                                 s ) );
   W.rb.allocate( W.B + 1 );
   DeviceArray< int64_t > deb;
   deb.allocate( W.B + 1 );
   TNLMI_HIP( hipMemcpyAsync( W.rb.ptr, rb.data(), sizeof( int32_t ) * ( W.B + 1 ), hipMemcpyHostToDevice, s ) );
   TNLMI_HIP( hipMemcpyAsync( deb.ptr, eb.data(), sizeof( int64_t ) * ( W.B + 1 ), hipMemcpyHostToDevice, s ) );
   // sort the nonzeros by window (stable: rows ascending inside a window)
   DeviceArray< int32_t > erow, key, skey, ids, perm, rblk, srow;
   for( auto* a : { &erow, &key, &skey, &ids, &perm, &srow } )
      a->allocate( std::max< int64_t >( 1, nnz ) );
   rblk.allocate( std::max< int64_t >( 1, n ) );
   const dim3 b256( 256 );
   auto g = [ & ]( int64_t c ) { return dim3( (unsigned) std::max< int64_t >( 1, ceil_div( c, 256 ) ) ); };
   hipLaunchKernelGGL( w_erow, g( n ), b256, 0, s, n, off, erow.ptr );
   hipLaunchKernelGGL( w_key, g( nnz ), b256, 0, s, nnz, cols, W.K, key.ptr, ids.ptr );
   int bits = 1;
   while( ( 1 << bits ) <= W.K + 1 )
      bits++;
   size_t bytes = 0;
   TNLMI_HIP( hipcub::DeviceRadixSort::SortPairs( nullptr, bytes, key.ptr, skey.ptr, ids.ptr, perm.ptr, (int) nnz, 0,
                                                  bits, s ) );
   DeviceArray< char > tmp;
   tmp.allocate( (int64_t) bytes + 1 );
   TNLMI_HIP( hipcub::DeviceRadixSort::SortPairs( tmp.ptr, bytes, key.ptr, skey.ptr, ids.ptr, perm.ptr, (int) nnz, 0,
                                                  bits, s ) );
   W.wstart.allocate( W.K + 2 );
   hipLaunchKernelGGL( w_starts, g( nnz + 1 ), b256, 0, s, nnz, W.K, skey.ptr, W.wstart.ptr );
   hipLaunchKernelGGL( w_rblk, g( W.B ), b256, 0, s, W.B, W.rb.ptr, rblk.ptr );
   W.pv.allocate( std::max< int64_t >( 1, nnz ) * (int64_t) sizeof( V ) );
   W.pc.allocate( std::max< int64_t >( 1, nnz ) );
   W.dest.allocate( nnz + kSegU );
   hipLaunchKernelGGL( ( w_fill< V > ), g( nnz ), b256, 0, s, nnz, perm.ptr, cols, vals, erow.ptr, rblk.ptr, deb.ptr,
                       (V*) W.pv.ptr, W.pc.ptr, W.dest.ptr, srow.ptr );
   W.seg.allocate( ( W.B + 1 ) * W.K );
   hipLaunchKernelGGL( w_seg, g( ( W.B + 1 ) * W.K ), b256, 0, s, W.B, W.K, W.rb.ptr, W.wstart.ptr, srow.ptr,
                       W.seg.ptr );
   W.prod.allocate( ( nnz + kSegU ) * (int64_t) sizeof( V ) );
   TNLMI_HIP( hipMemsetAsync( W.dest.ptr + nnz, 0, sizeof( uint16_t ) * kSegU, s ) );
   TNLMI_HIP( hipMemsetAsync( W.prod.ptr + nnz * (int64_t) sizeof( V ), 0, sizeof( V ) * kSegU, s ) );
   TNLMI_HIP( hipGetLastError() );
   TNLMI_HIP( hipStreamSynchronize( s ) );
}

template< typename C >
void run( const WindowPlan& W, const C* x, C* y, double alpha, double beta, int64_t begin, int64_t end,
          hipStream_t s )
{
   if( W.nnz > 0 )
      hipLaunchKernelGGL( ( w_phase1< C, C > ), dim3( (unsigned) W.K ), dim3( kP1 ), 0, s, W.columns, W.wstart.ptr,
                          (const C*) W.pv.ptr, W.pc.ptr, x, (C*) W.prod.ptr );
   if( W.B > 0 )
      hipLaunchKernelGGL( ( w_phase2< C, cap_of< C >() > ), dim3( (unsigned) W.B ), dim3( kP2 ), 0, s, W.K, W.seg.ptr,
                          (const C*) W.prod.ptr, W.dest.ptr, W.rb.ptr, W.offs.ptr, W.lstart.ptr, W.lrows.ptr, y,
                          (C) alpha, (C) beta, begin, end );
}

}  // namespace

void window_build( WindowPlan& W, const tnlmi_matrix& A, hipStream_t s )
{
   W.rows = A.rows;
   W.columns = A.columns;
   W.nnz = A.segments->elements;
   W.value_type = A.value_type;
   W.offs.allocate( A.rows + 1 );
   TNLMI_HIP( hipMemcpyAsync( W.offs.ptr, A.segments->offsets.ptr, sizeof( int32_t ) * ( A.rows + 1 ),
                              hipMemcpyDeviceToDevice, s ) );
   if( A.value_type == TNLMI_F64 ) {
      W.cap = cap_of< double >();
      build< double >( W, A, s );
   }
   else if( A.value_type == TNLMI_F32 ) {
      W.cap = cap_of< float >();
      build< float >( W, A, s );
   }
   else
      fail( TNLMI_INVALID_ARGUMENT, "windowed SpMV plan: FP64 or FP32 values" );
}

void window_execute( const WindowPlan& W, tnlmi_dtype xt, const void* x, void* y, tnlmi_dtype ct, double alpha,
                     double beta, int64_t begin, int64_t end, hipStream_t s )
{
   if( W.value_type == TNLMI_F64 && xt == TNLMI_F64 && ct == TNLMI_F64 )
      run< double >( W, (const double*) x, (double*) y, alpha, beta, begin, end, s );
   else if( W.value_type == TNLMI_F32 && xt == TNLMI_F32 && ct == TNLMI_F32 )
      run< float >( W, (const float*) x, (float*) y, alpha, beta, begin, end, s );
   else
      fail( TNLMI_NOT_IMPLEMENTED, "windowed SpMV plan: values, vectors and compute type must all be FP64 or FP32" );
   TNLMI_HIP( hipGetLastError() );
}

}  // namespace tnlmi
