// spmv_window.hip -- the windowed SpMV plan of a CSR matrix: every x gather
// from LDS, every HBM access a stream.
//
// Random column indexes make every nonzero an 8-byte gather into x; through
// the L2 each one costs a 128-byte line and an L2 request.  The plan splits
// the columns into windows of `win` entries (8192 by default: 64 KiB of FP64
// x, one LDS allocation; 10240, 16384 and 20480 are also built, up to the
// 160 KiB of LDS a gfx950 workgroup can hold) and the nonzeros into row blocks of at most `cap` (FP64 10000,
// FP32 20000): runs of whole consecutive rows, or one piece of a row longer
// than cap (such a row is cut into pieces of cap nonzeros):
//
//   phase 1  one workgroup per window: x[window] -> LDS, then stream the
//            window's nonzeros (values + 13-bit local columns, in CSR order)
//            and write each product v * x[c] to prod[] (window-major)
//   phase 2  one workgroup per row block: copy the block's products from each
//            window's contiguous segment into LDS at their CSR position
//            (16-bit slot), then sum every row in CSR order (a piece: its
//            partial sum)
//   fixup    rows cut into pieces: the pieces' partial sums, added in order
//
// Per nonzero: 8 + 2 bytes read and 8 written in phase 1, 8 + 2 read in
// phase 2 -- about 1.9 GB for the 67 M-nonzero power-law matrix, all of it
// coalesced, and no gather leaves LDS.  Products are rounded to the compute
// type and each row is summed from zero left to right in column order, as
// TNL's host loop does (ReducingOperations_CSR.h:46-74 with
// SparseMatrixBase.hpp:331-389), so every row of at most kLongRow = 256
// nonzeros, summed by one thread, matches it bit for bit; longer rows are
// summed by a wave (pieces: by the workgroup) in a fixed, deterministic order.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <string>
#include <vector>

#include "segments.hpp"
#include "spmv.hpp"
#include "spmv_window.hpp"

namespace tnlmi {
namespace {

constexpr int kWinDefault = 8192;    // x entries per window
constexpr int kP1 = 1024;            // phase-1 threads
constexpr int kP2 = 512;             // phase-2 threads
// a row block has at most 2 kP2 rows (two per thread) and kP2 / 64 rows longer
// than kLongRow (one per wave)
constexpr int kLongRow = 256;        // longer rows are summed by a whole wave
constexpr int kPad = 32;             // padding entries of prod and dest (phase-2 reads past a segment)

// row-block capacity: the phase-2 LDS buffer of 80,000 bytes, so two phase-2
// workgroups share a CU's 160 KiB (measured against 8160 FP64 nonzeros in
// 64 KiB: 467-480 against 481-496 us per power-law SpMV, the chunks of a
// window's products in a block being 1.2 times longer)
template< typename C >
constexpr int cap_of()
{
   return sizeof( C ) == 8 ? 10000 : 20000;
}

// window of each nonzero (padding slots, column -1, go to an extra window K)
__global__ void w_key( int64_t nnz, const int32_t* __restrict__ cols, int32_t K, int32_t win,
                       int32_t* __restrict__ key, int32_t* __restrict__ ids )
{
   const int64_t e = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( e >= nnz )
      return;
   const int32_t c = cols[ e ];
   key[ e ] = c >= 0 ? c / win : K;
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

// the row block holding nonzero e: the last b < B with E[b] <= e
__device__ __forceinline__ int64_t block_of( int64_t B, const int64_t* __restrict__ E, int64_t e )
{
   int64_t lo = 0, hi = B;
   while( hi - lo > 1 ) {
      const int64_t mid = ( lo + hi ) / 2;
      if( E[ mid ] <= e )
         lo = mid;
      else
         hi = mid;
   }
   return lo;
}

// the sorted arrays: value, local column; and the slot within the row block
// of each nonzero of window k < K (chunk (b, k)), at i + dbase[b K + k]
template< typename V >
__global__ void w_fill( int64_t nnz, int32_t win, const int32_t* __restrict__ perm, const int32_t* __restrict__ cols,
                        const V* __restrict__ vals, int64_t B, const int64_t* __restrict__ E,
                        const int32_t* __restrict__ skey, int32_t K, const int32_t* __restrict__ dbase,
                        V* __restrict__ pv, uint16_t* __restrict__ pc, uint16_t* __restrict__ dest )
{
   const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( i >= nnz )
      return;
   const int32_t e = perm[ i ];
   const int32_t c = cols[ e ];
   pv[ i ] = vals[ e ];
   pc[ i ] = (uint16_t) ( c >= 0 ? c % win : 0 );
   const int32_t k = skey[ i ];
   if( k < K ) {  // window K: padding slots, in no chunk
      const int64_t b = block_of( B, E, e );
      dest[ i + dbase[ b * K + k ] ] = (uint16_t) ( e - E[ b ] );
   }
}

// seg[b * K + k] = first position of window k holding a nonzero e >= E[b]
// (a window's nonzeros are in CSR order: the sort is stable)
__global__ void w_seg( int64_t B, int32_t K, const int64_t* __restrict__ E, const int64_t* __restrict__ wstart,
                       const int32_t* __restrict__ perm, int32_t* __restrict__ seg )
{
   const int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( q >= ( B + 1 ) * K )
      return;
   const int64_t b = q / K;
   const int32_t k = (int32_t) ( q % K );
   int64_t lo = wstart[ k ], hi = wstart[ k + 1 ];
   const int64_t e = E[ b ];
   while( lo < hi ) {
      const int64_t mid = ( lo + hi ) / 2;
      if( perm[ mid ] < e )
         lo = mid + 1;
      else
         hi = mid;
   }
   seg[ q ] = (int32_t) lo;
}

// phase 1: prod[i] = v[i] * x[window base + pc[i]] for the window's nonzeros
// (U nonzeros per thread and batch; a software-pipelined form that requests
// the next batch before the current one's products are stored measured
// slower: 272 against 235 us per power-law SpMV)
template< typename V, typename C, int WIN, int U >
__global__ void __launch_bounds__( kP1 ) w_phase1( int64_t ncols, const int64_t* __restrict__ wstart,
                                                   const V* __restrict__ pv, const uint16_t* __restrict__ pc,
                                                   const C* __restrict__ x, C* __restrict__ prod )
{
   __shared__ C xw[ WIN ];
   const int64_t k = blockIdx.x;
   const int64_t base = k * WIN;
   for( int t = threadIdx.x; t < WIN; t += kP1 )
      xw[ t ] = base + t < ncols ? x[ base + t ] : (C) 0;
   __syncthreads();
   const int64_t a = wstart[ k ], b = wstart[ k + 1 ];
   // plain (cached) loads and stores: the nontemporal forms measured slower
   // here, the stored products in particular (phase 2 reads them next)
   for( int64_t i0 = a + threadIdx.x; i0 < b; i0 += U * kP1 ) {
      C p[ U ];
#pragma unroll
      for( int u = 0; u < U; u++ ) {
         const int64_t i = i0 + u * kP1;
         p[ u ] = i < b ? (C) pv[ i ] * xw[ pc[ i ] ] : (C) 0;
      }
#pragma unroll
      for( int u = 0; u < U; u++ ) {
         const int64_t i = i0 + u * kP1;
         if( i < b )
            prod[ i ] = p[ u ];
      }
   }
}

// phase 2: row block b -> LDS in CSR order -> row sums -> y.  One workgroup
// per row block.  The products of a block's segment (window k) are loaded by
// a group of GRP lanes together, PER consecutive terms each: one load
// instruction covers GRP * PER terms of a segment, so the lanes of a group
// touch consecutive cache lines (a lane per segment would touch a line of its
// own in every load, and the L1 tag lookups, one per line, would bound the
// kernel).  All rounds of a thread (kP2 / GRP groups, R = SEGS / (kP2 /
// GRP) segments each) are issued before the first is stored.  SEGS, the
// segments (windows) per sweep of a block, is 512 (default shape: 8-lane
// groups of 4 terms per lane), or 256 for plans of at most 256 windows
// (16-lane groups of 2).  (1024-thread workgroups at 64 VGPRs, two per CU,
// measured no faster than 512.)
constexpr int kGrp = 16;  // the default group width (GRP)

// per-block header: 8 int32 -- rows [r0, r1), nonzeros [e0, e1), the long
// row a piece belongs to (-1: a run of whole rows), long-row list [l0, l1)
struct BlockHdr
{
   int32_t r0, r1, e0, e1, piece, l0, l1, pad;
};

template< typename C >
__device__ __forceinline__ C sum_row( const C* buf, int32_t q0, int32_t q1 )
{
   // from zero, left to right; the LDS reads of kSumU terms issued before
   // their adds (a long serial add chain holds the whole block: unrolling
   // 4 -> 16 measured 487 -> 470 us per power-law SpMV; issuing the last,
   // shorter run's reads together as well measured slower)
   constexpr int kSumU = 16;
   C sum = 0;
   int32_t q = q0;
   for( ; q + kSumU <= q1; q += kSumU ) {
      C t[ kSumU ];
#pragma unroll
      for( int u = 0; u < kSumU; u++ )
         t[ u ] = buf[ q + u ];
#pragma unroll
      for( int u = 0; u < kSumU; u++ )
         sum += t[ u ];
   }
   for( ; q < q1; q++ )
      sum += buf[ q ];
   return sum;
}

// PER terms at i (products at i, their 16-bit slots at i + di), loaded
// unconditionally: prod and dest carry kPad entries of padding, so a read past
// a segment's end stays in bounds (and what it reads is not stored); di is a
// multiple of 4, so both reads have the same alignment
template< typename C, int PER >
__device__ __forceinline__ void get_pair( const C* prod, const uint16_t* dest, int32_t i, int32_t di,
                                          C ( &v )[ PER ], uint32_t ( &d )[ PER / 2 ] )
{
#pragma unroll
   for( int u = 0; u < PER; u++ )
      v[ u ] = prod[ i + u ];
   __builtin_memcpy( d, dest + ( i + di ), sizeof( d ) );
}

template< typename C, int PER >
__device__ __forceinline__ void put_pair( C* buf, const C ( &v )[ PER ], const uint32_t ( &d )[ PER / 2 ], int32_t i,
                                          int32_t z, int32_t a = INT32_MIN )
{
#pragma unroll
   for( int u = 0; u < PER; u++ )
      if( i + u < z && i + u >= a )
         buf[ ( d[ u / 2 ] >> ( 16 * ( u % 2 ) ) ) & 0xffff ] = v[ u ];
}

template< typename C, int CAP, bool BETA, int SEGS, int PER, bool ALIGN, int GRP = kGrp >
__global__ void __launch_bounds__( kP2 )
   w_phase2( int64_t B, int32_t K, const BlockHdr* __restrict__ hdr, const int32_t* __restrict__ seg,
             const C* __restrict__ prod, const uint16_t* __restrict__ dest, const int32_t* __restrict__ dbase,
             const int32_t* __restrict__ off,
             const int32_t* __restrict__ lrows, const uint16_t* __restrict__ rorder, C* __restrict__ part,
             C* __restrict__ y, C alpha, C beta,
             int64_t begin, int64_t end, int32_t across )
{
   __shared__ C buf[ CAP ];
   constexpr int P2 = kP2;
   __shared__ C red[ P2 / 64 ];
   // SEGS segments per sweep: P2 / GRP groups of R segments each
   constexpr int R = SEGS * GRP / P2;
   constexpr int kSpan = GRP * PER;  // terms of a segment per load round
   static_assert( R >= 1 && R <= GRP, "a group's lanes hold its segments' bounds" );
   // a read of PER terms starts at most at a segment's last term (or, for an
   // empty segment, its start, at most nnz): it ends inside the kPad padding
   static_assert( PER % 2 == 0 && PER <= kPad, "slot words hold two slots; reads stay in the padding" );
   // consecutive row blocks on one XCD: their chunks of a window's products
   // (about 120 bytes each, 30 of slot indexes) share cache lines in its L2
   const int64_t b = xcd_contiguous( blockIdx.x, gridDim.x );
   const BlockHdr H = hdr[ b ];  // uniform
   if( ( H.piece < 0 ? H.r1 : H.r0 + 1 ) <= begin || H.r0 >= end )
      return;  // no row of this block is asked for
   const int t = threadIdx.x;
   const int lane = t & 63, wid = wave_uniform( (int32_t) ( t >> 6 ) );
   const int gl = lane % GRP, gb = lane - gl;
   // segment sx (window sx) of the block: its bounds, held by lane gb + i
   // (i < R) of the group's lanes.  A wave takes G R consecutive segments;
   // across: in round i its G groups read segments G i .. G i + G - 1 of
   // them, whose slots lie side by side (the slot array is block-major);
   // else group g reads segment g R + i
   constexpr int G = 64 / GRP;
   const int32_t sx = wid * G * R
                      + ( across ? std::min( gl, R - 1 ) * G + lane / GRP : ( lane / GRP ) * R + std::min( gl, R - 1 ) );
   const int32_t kt = sx < K ? sx : K - 1;
   // both bounds loaded unconditionally (a load under the condition would
   // wait for the first before it is issued)
   const int32_t a = seg[ b * K + kt ];
   const int32_t zk = seg[ ( b + 1 ) * K + kt ];
   const int32_t z = sx < K ? zk : a;
   const int32_t da = dbase[ b * K + kt ];
   // the rows of this thread: ranks t and t + P2 of the block's rows ordered
   // by length, longest first (a block has at most 2 P2 rows), so the rows
   // of a wave, each summed serially by its lane, are about equally long;
   // and its wave's long row (at most one per wave).  Unconditional loads at
   // clamped indexes, requested after the products (no product request
   // waits on them; the offsets arrive while the products are scattered).
   constexpr int J = 2;
   const int32_t nrows = H.r1 - H.r0;
   int32_t rr[ J ], ob[ J ], oe[ J ];
   // round i: the group's segment i, kSpan terms.  ALIGN: a group's reads
   // start at the segment's first term rounded down to a multiple of PER, so
   // every lane's PER products (8 PER bytes) and slots (2 PER bytes) are
   // naturally aligned; a term before the segment is read but not stored.
   constexpr int32_t kAl = ALIGN ? ~( PER - 1 ) : ~0;
   C v[ R ][ PER ];
   uint32_t d[ R ][ PER / 2 ];
#pragma unroll
   for( int i = 0; i < R; i++ ) {
      // lanes past the segment's end read its last term again (no extra
      // cache line per request); what they read is not stored
      const int32_t ai = __shfl( a, gb + i, 64 ), zi = __shfl( z, gb + i, 64 ), di = __shfl( da, gb + i, 64 );
      get_pair< C, PER >( prod, dest, std::min( ( ai & kAl ) + PER * gl, std::max( zi - 1, ai ) & kAl ), di,
                          v[ i ], d[ i ] );
   }
#pragma unroll
   for( int j = 0; j < J; j++ )
      rr[ j ] = H.r0 + rorder[ H.r0 + std::min( t + j * P2, std::max( nrows - 1, 0 ) ) ];
#pragma unroll
   for( int j = 0; j < J; j++ ) {
      ob[ j ] = off[ rr[ j ] ];
      oe[ j ] = off[ rr[ j ] + 1 ];
   }
   const int32_t li = std::min( H.l0 + wid, std::max( H.l1 - 1, 0 ) );
   const int32_t lr = lrows[ 3 * li ], lq0 = lrows[ 3 * li + 1 ], lq1 = lrows[ 3 * li + 2 ];
   // clear (padding slots, column -1, have no product: they read as zero)
   const int32_t len = H.e1 - H.e0;
   for( int32_t q = t; q < len; q += P2 )
      buf[ q ] = (C) 0;
   __syncthreads();
#pragma unroll
   for( int i = 0; i < R; i++ ) {
      const int32_t ai = __shfl( a, gb + i, 64 ), zi = __shfl( z, gb + i, 64 );
      const int32_t a0 = ai & kAl;
      put_pair< C, PER >( buf, v[ i ], d[ i ], a0 + PER * gl, zi, ai );
      // on demand: the rest of a segment longer than kSpan
      const int32_t di = __shfl( da, gb + i, 64 );
      for( int32_t i0 = a0 + kSpan + PER * gl; i0 < zi; i0 += kSpan ) {
         C w[ PER ];
         uint32_t dw[ PER / 2 ];
         get_pair< C, PER >( prod, dest, i0, di, w, dw );
         put_pair< C, PER >( buf, w, dw, i0, zi );
      }
   }
   // on demand: windows past SEGS (segment k handled as segment k % SEGS)
   for( int32_t k0 = SEGS; k0 < K; k0 += SEGS ) {
      const int32_t k = std::min( k0 + sx, K - 1 );
      const int32_t ak = seg[ b * K + k ], zk = k0 + sx < K ? seg[ ( b + 1 ) * K + k ] : ak;
      const int32_t dk = dbase[ b * K + k ];
      for( int i = 0; i < R; i++ ) {
         const int32_t ai = __shfl( ak, gb + i, 64 ), zi = __shfl( zk, gb + i, 64 ), di = __shfl( dk, gb + i, 64 );
         for( int32_t i0 = ( ai & kAl ) + PER * gl; i0 < zi; i0 += kSpan ) {
            C w[ PER ];
            uint32_t dw[ PER / 2 ];
            get_pair< C, PER >( prod, dest, i0, di, w, dw );
            put_pair< C, PER >( buf, w, dw, i0, zi, ai );
         }
      }
   }
   __syncthreads();
   const int32_t e0 = H.e0;
   auto keep = [ & ]( int32_t r, C sum ) {
      if( r >= begin && r < end )
         y[ r ] = BETA ? alpha * sum + beta * y[ r ] : alpha * sum;
   };
   if( H.piece >= 0 ) {
      // a piece of a row longer than cap: its partial sum (strided parts,
      // then a fixed-order reduction), added to the other pieces' by w_fixup
      C p = 0;
      for( int32_t q = t; q < len; q += P2 )
         p += buf[ q ];
      const C sum = block_sum< P2 >( p, red );
      if( t == 0 )
         part[ b ] = sum;
      return;
   }
   // rows of at most kLongRow nonzeros: one thread each, summed from zero
   // left to right
#pragma unroll
   for( int j = 0; j < J; j++ )
      if( t + j * P2 < nrows && oe[ j ] - ob[ j ] <= kLongRow )
         keep( rr[ j ], sum_row( buf, ob[ j ] - e0, oe[ j ] - e0 ) );
   // the long row of this wave (if any): lane l sums terms l, l + 64, ... and
   // an xor butterfly adds the 64 parts
   if( H.l0 + wid < H.l1 ) {
      C p = 0;
      for( int32_t q = lq0 + lane; q < lq1; q += 64 )
         p += buf[ q ];
      const C sum = group_sum< 64 >( p );
      if( lane == 0 )
         keep( lr, sum );
   }
}

// rows cut into pieces: y[row] from the pieces' partial sums, in piece order
// (huge holds (row, first piece block, pieces) per row)
template< typename C >
__global__ void w_fixup( int64_t H, const int32_t* __restrict__ huge, const C* __restrict__ part, C* __restrict__ y,
                         C alpha, C beta, int64_t begin, int64_t end )
{
   const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( i >= H )
      return;
   const int32_t r = huge[ 3 * i ], first = huge[ 3 * i + 1 ], count = huge[ 3 * i + 2 ];
   if( r < begin || r >= end )
      return;
   C sum = 0;
   for( int32_t j = 0; j < count; j++ )
      sum += part[ first + j ];
   y[ r ] = beta == (C) 0 ? alpha * sum : alpha * sum + beta * y[ r ];
}

template< typename V >
void build( WindowPlan& W, const tnlmi_matrix& A, hipStream_t s )
{
   const tnlmi_segments& seg = *A.segments;
   const int64_t n = A.rows, nnz = seg.elements;
   const int32_t* off = seg.offsets.ptr;
   const int32_t* cols = A.column_indexes;
   const V* vals = (const V*) A.values;
   W.K = (int32_t) std::max< int64_t >( 1, ceil_div( A.columns, W.win ) );
   // row blocks: runs of whole consecutive rows with at most cap nonzeros
   // together, or pieces of cap nonzeros of a longer row
   std::vector< int32_t > hoff( n + 1 );
   TNLMI_HIP( hipMemcpyAsync( hoff.data(), off, sizeof( int32_t ) * ( n + 1 ), hipMemcpyDeviceToHost, s ) );
   TNLMI_HIP( hipStreamSynchronize( s ) );
   const int64_t cap = W.cap;
   std::vector< int32_t > rb, piece, huge;
   std::vector< int64_t > E;
   for( int64_t r = 0; r < n; ) {
      const int64_t start = hoff[ r ], len = hoff[ r + 1 ] - start;
      if( len > cap ) {
         const int64_t count = ceil_div( len, cap );
         const int32_t h = (int32_t) ( huge.size() / 3 );
         huge.insert( huge.end(), { (int32_t) r, (int32_t) rb.size(), (int32_t) count } );
         for( int64_t j = 0; j < count; j++ ) {
            rb.push_back( (int32_t) r );
            E.push_back( start + j * cap );
            piece.push_back( h );
         }
         r++;
         continue;
      }
      // at most cap nonzeros, 2 kP2 rows and kP2 / 64 rows longer than kLongRow
      int64_t e = r, longs = 0;
      while( e < n && hoff[ e + 1 ] - start <= cap && e - r < 2 * kP2 ) {
         const bool lg = hoff[ e + 1 ] - hoff[ e ] > kLongRow;
         if( lg && longs == kP2 / 64 )
            break;
         longs += lg;
         e++;
      }
      rb.push_back( (int32_t) r );
      E.push_back( start );
      piece.push_back( -1 );
      r = e;
   }
   rb.push_back( (int32_t) n );
   E.push_back( nnz );
   W.B = (int64_t) piece.size();
   W.H = (int64_t) huge.size() / 3;
   // block headers; rows longer than kLongRow inside runs of whole rows (summed
   // by a wave) as (row, first slot, end slot)
   std::vector< int32_t > hdr, lrows;
   for( int64_t i = 0; i < W.B; i++ ) {
      const int32_t l0 = (int32_t) ( lrows.size() / 3 );
      if( piece[ i ] < 0 )
         for( int32_t r = rb[ i ]; r < rb[ i + 1 ]; r++ )
            if( hoff[ r + 1 ] - hoff[ r ] > kLongRow )
               lrows.insert( lrows.end(), { r, (int32_t) ( hoff[ r ] - E[ i ] ), (int32_t) ( hoff[ r + 1 ] - E[ i ] ) } );
      hdr.insert( hdr.end(), { rb[ i ], rb[ i + 1 ], (int32_t) E[ i ], (int32_t) E[ i + 1 ], piece[ i ], l0,
                               (int32_t) ( lrows.size() / 3 ), 0 } );
   }
   auto upload = [ & ]( auto& dst, const auto& src ) {
      dst.allocate( std::max< int64_t >( 1, (int64_t) src.size() ) );
      if( !src.empty() )
         TNLMI_HIP( hipMemcpyAsync( dst.ptr, src.data(), sizeof( src[ 0 ] ) * src.size(), hipMemcpyHostToDevice,
                                    s ) );
   };
   DeviceArray< int64_t > dE;  // build only
   upload( dE, E );
   upload( W.hdr, hdr );
   upload( W.huge, huge );
   if( lrows.empty() )
      lrows.assign( 3, 0 );  // phase 2 reads one entry at a clamped index
   upload( W.lrows, lrows );
   // each run's rows by length, longest first (ties: row order); a piece
   // block's row keeps index 0
   {
      std::vector< uint16_t > order( std::max< int64_t >( 1, n ), 0 );
      std::vector< uint16_t > idx;
      for( int64_t i = 0; i < W.B; i++ ) {
         if( piece[ i ] >= 0 )
            continue;
         const int32_t r0 = rb[ i ], cnt = rb[ i + 1 ] - r0;
         idx.resize( cnt );
         for( int32_t q = 0; q < cnt; q++ )
            idx[ q ] = (uint16_t) q;
         std::stable_sort( idx.begin(), idx.end(), [ & ]( uint16_t a, uint16_t b ) {
            return hoff[ r0 + a + 1 ] - hoff[ r0 + a ] > hoff[ r0 + b + 1 ] - hoff[ r0 + b ];
         } );
         std::copy( idx.begin(), idx.end(), order.begin() + r0 );
      }
      upload( W.rorder, order );
   }
   W.part.allocate( std::max< int64_t >( 1, W.B ) * (int64_t) sizeof( V ) );
   // sort the nonzeros by window (stable: CSR order inside a window)
   DeviceArray< int32_t > key, skey, ids, perm;
   for( auto* a : { &key, &skey, &ids, &perm } )
      a->allocate( std::max< int64_t >( 1, nnz ) );
   const dim3 b256( 256 );
   auto g = [ & ]( int64_t c ) { return dim3( (unsigned) std::max< int64_t >( 1, ceil_div( c, 256 ) ) ); };
   hipLaunchKernelGGL( w_key, g( nnz ), b256, 0, s, nnz, cols, W.K, W.win, key.ptr, ids.ptr );
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
   W.seg.allocate( ( W.B + 1 ) * W.K );
   hipLaunchKernelGGL( w_seg, g( ( W.B + 1 ) * W.K ), b256, 0, s, W.B, W.K, dE.ptr, W.wstart.ptr, perm.ptr,
                       W.seg.ptr );
   // the slot array, block-major: chunk (b, k) (its window-k nonzeros, in
   // prod at [seg[b K + k], seg[(b + 1) K + k])) gets the 4-slot groups
   // [c, c + ceil((r + len) / 4)) after the block's previous chunk, r = its
   // first index mod 4, and its slots start at 4 c + r, so a phase-2 read of
   // a chunk's slots has the alignment of its read of the products, and a
   // row block's slots are one contiguous run.  TNLMI_WINDOW_SLOTS=window
   // keeps them in prod's order (dbase 0).
   std::vector< int32_t > dbase( std::max< int64_t >( 1, W.B * W.K ), 0 );
   int64_t slots = nnz;
   const char* lay = std::getenv( "TNLMI_WINDOW_SLOTS" );
   if( ! ( lay && std::string( lay ) == "window" ) && W.B > 0 ) {
      std::vector< int32_t > hseg( ( W.B + 1 ) * W.K );
      TNLMI_HIP( hipMemcpyAsync( hseg.data(), W.seg.ptr, sizeof( int32_t ) * hseg.size(), hipMemcpyDeviceToHost,
                                 s ) );
      TNLMI_HIP( hipStreamSynchronize( s ) );
      int64_t c = kPad / 4;  // reads may start up to 3 slots before a chunk
      for( int64_t b = 0; b < W.B; b++ )
         for( int64_t k = 0; k < W.K; k++ ) {
            const int64_t a = hseg[ b * W.K + k ], len = hseg[ ( b + 1 ) * W.K + k ] - a, r = a % 4;
            dbase[ b * W.K + k ] = (int32_t) ( 4 * c + r - a );
            c += ( r + len + 3 ) / 4;
         }
      slots = 4 * c;
      if( slots + kPad > INT32_MAX )
         fail( TNLMI_NOT_IMPLEMENTED, "windowed SpMV plan: more than 2^31 slot entries" );
   }
   upload( W.dbase, dbase );
   W.dest.allocate( slots + kPad );
   TNLMI_HIP( hipMemsetAsync( W.dest.ptr, 0, sizeof( uint16_t ) * ( slots + kPad ), s ) );
   W.pv.allocate( std::max< int64_t >( 1, nnz ) * (int64_t) sizeof( V ) );
   W.pc.allocate( std::max< int64_t >( 1, nnz ) );
   hipLaunchKernelGGL( ( w_fill< V > ), g( nnz ), b256, 0, s, nnz, W.win, perm.ptr, cols, vals, W.B, dE.ptr,
                       skey.ptr, W.K, W.dbase.ptr, (V*) W.pv.ptr, W.pc.ptr, W.dest.ptr );
   W.prod.allocate( ( nnz + kPad ) * (int64_t) sizeof( V ) );
   TNLMI_HIP( hipMemsetAsync( W.prod.ptr + nnz * (int64_t) sizeof( V ), 0, sizeof( V ) * kPad, s ) );
   TNLMI_HIP( hipGetLastError() );
   TNLMI_HIP( hipStreamSynchronize( s ) );
}

template< typename C, int WIN >
void phase1( const WindowPlan& W, const C* x, hipStream_t s )
{
   if( W.unroll == 8 )
      hipLaunchKernelGGL( ( w_phase1< C, C, WIN, 8 > ), dim3( (unsigned) W.K ), dim3( kP1 ), 0, s, W.columns,
                          W.wstart.ptr, (const C*) W.pv.ptr, W.pc.ptr, x, (C*) W.prod.ptr );
   else
      hipLaunchKernelGGL( ( w_phase1< C, C, WIN, 4 > ), dim3( (unsigned) W.K ), dim3( kP1 ), 0, s, W.columns,
                          W.wstart.ptr, (const C*) W.pv.ptr, W.pc.ptr, x, (C*) W.prod.ptr );
}

template< typename C, int SEGS, int PER, int GRP = kGrp >
void phase2( const WindowPlan& W, C* y, double alpha, double beta, int64_t begin, int64_t end, hipStream_t s )
{
   if( W.align )
      hipLaunchKernelGGL( ( beta == 0.0 ? w_phase2< C, cap_of< C >(), false, SEGS, PER, true, GRP >
                                        : w_phase2< C, cap_of< C >(), true, SEGS, PER, true, GRP > ),
                          dim3( (unsigned) W.B ), dim3( kP2 ), 0, s, W.B, W.K, (const BlockHdr*) W.hdr.ptr, W.seg.ptr,
                          (const C*) W.prod.ptr, W.dest.ptr, W.dbase.ptr, W.offs.ptr, W.lrows.ptr, W.rorder.ptr, (C*) W.part.ptr, y, (C) alpha,
                          (C) beta, begin, end, (int32_t) W.across );
   else
      hipLaunchKernelGGL( ( beta == 0.0 ? w_phase2< C, cap_of< C >(), false, SEGS, PER, false, GRP >
                                        : w_phase2< C, cap_of< C >(), true, SEGS, PER, false, GRP > ),
                          dim3( (unsigned) W.B ), dim3( kP2 ), 0, s, W.B, W.K, (const BlockHdr*) W.hdr.ptr, W.seg.ptr,
                          (const C*) W.prod.ptr, W.dest.ptr, W.dbase.ptr, W.offs.ptr, W.lrows.ptr, W.rorder.ptr, (C*) W.part.ptr, y, (C) alpha,
                          (C) beta, begin, end, (int32_t) W.across );
}

template< typename C >
void run( const WindowPlan& W, const C* x, C* y, double alpha, double beta, int64_t begin, int64_t end,
          hipStream_t s )
{
   if( W.nnz > 0 ) {
      switch( W.win ) {
         case 8192: phase1< C, 8192 >( W, x, s ); break;
         case 10240: phase1< C, 10240 >( W, x, s ); break;
         case 16384: phase1< C, 16384 >( W, x, s ); break;
         case 20480: phase1< C, 20480 >( W, x, s ); break;
         default: fail( TNLMI_RUNTIME_ERROR, "windowed SpMV plan: unsupported window width" );
      }
   }
   if( W.B > 0 ) {
      const int key = W.segs * 10 + W.per;
      switch( key ) {
         case 5122: phase2< C, 512, 2 >( W, y, alpha, beta, begin, end, s ); break;
         case 2562: phase2< C, 256, 2 >( W, y, alpha, beta, begin, end, s ); break;
         case 2564: phase2< C, 256, 4 >( W, y, alpha, beta, begin, end, s ); break;
         case 5124: phase2< C, 512, 4, 8 >( W, y, alpha, beta, begin, end, s ); break;  // 8-lane groups
         default: fail( TNLMI_RUNTIME_ERROR, "windowed SpMV plan: unsupported phase-2 shape" );
      }
   }
   if( W.H > 0 )
      hipLaunchKernelGGL( ( w_fixup< C > ), dim3( (unsigned) ceil_div( W.H, 256 ) ), dim3( 256 ), 0, s, W.H,
                          W.huge.ptr, (const C*) W.part.ptr, y, (C) alpha, (C) beta, begin, end );
}

}  // namespace

void window_build( WindowPlan& W, const tnlmi_matrix& A, hipStream_t s )
{
   W.rows = A.rows;
   W.columns = A.columns;
   W.nnz = A.segments->elements;
   W.value_type = A.value_type;
   // window width and kernel shapes: TNLMI_WINDOW_COLS (8192, 10240, 16384,
   // 20480), TNLMI_WINDOW_PER (phase 2: 2, 4) and TNLMI_WINDOW_UNROLL (phase 1:
   // 4, 8) in the environment override them
   W.win = kWinDefault;
   if( const char* e = std::getenv( "TNLMI_WINDOW_COLS" ) ) {
      const int v = std::atoi( e );
      if( v == 8192 || v == 10240 || v == 16384 || v == 20480 )
         W.win = v;
   }
   W.K = (int32_t) std::max< int64_t >( 1, ceil_div( A.columns, W.win ) );
   W.segs = W.K <= 256 ? 256 : 512;
   // phase 2 with 512 windows per sweep: 8-lane groups of 4 terms per lane
   // (449 against 455 us per power-law SpMV with 16-lane groups of 2 terms,
   // 460 with 4-lane groups of 8); with 256: 16-lane groups of 2 terms
   W.per = W.segs == 512 ? 4 : 2;
   if( const char* e = std::getenv( "TNLMI_WINDOW_PER" ) )
      if( std::atoi( e ) == 2 || std::atoi( e ) == 4 )
         W.per = std::atoi( e );
   W.unroll = 4;
   if( const char* e = std::getenv( "TNLMI_WINDOW_UNROLL" ) )
      if( std::atoi( e ) == 8 )
         W.unroll = 8;
   // TNLMI_WINDOW_ALIGN=0: phase-2 reads start at each segment's first term
   // instead of the aligned position before it
   W.align = true;
   if( const char* e = std::getenv( "TNLMI_WINDOW_ALIGN" ) )
      W.align = e[ 0 ] != '0';
   // TNLMI_WINDOW_ROUNDS=grouped: phase-2 groups read their own consecutive
   // segments round by round instead of a wave reading adjacent ones
   W.across = true;
   if( const char* e = std::getenv( "TNLMI_WINDOW_ROUNDS" ) )
      W.across = std::string( e ) != "grouped";
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
