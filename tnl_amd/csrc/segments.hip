// segments.hip -- format converters (setSegmentsSizes of every TNL segment
// format), matrix assembly into the format storage and the C ABI for both.
//
// GPU converters: CSR offsets (scan), Ellpack (max), SlicedEllpack (per-slice
// max + scan), BiEllpack (per-strip stable sort + group widths + scan),
// SortedSegments (stable radix sort).  ChunkedEllpack slices and AdaptiveCSR
// descriptors are sequential greedy passes in the reference and are built on
// the host here as there (ChunkedEllpack.hpp:159-161, AdaptiveCSR.hpp:198-231),
// then uploaded.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "segments.hpp"
#include "spmv.hpp"

namespace tnlmi {

// ------------------------------------------------------------------ scan
namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

__global__ void __launch_bounds__( kScanBlock ) scan_partials( const int32_t* in, int64_t n, int64_t* part )
{
   __shared__ int64_t sm[ kScanBlock / 64 ];
   const int64_t base = (int64_t) blockIdx.x * kScanTile;
   int64_t acc = 0;
   for( int i = 0; i < kScanItems; i++ ) {
      const int64_t idx = base + (int64_t) i * kScanBlock + threadIdx.x;
      if( idx < n )
         acc += in[ idx ];
   }
   acc = block_sum< kScanBlock >( acc, sm );
   if( threadIdx.x == 0 )
      part[ blockIdx.x ] = acc;
}

// single workgroup: exclusive scan of the partials (in place), total -> *total
__global__ void __launch_bounds__( 1024 ) scan_partials_scan( int64_t* part, int64_t nb, int64_t* total )
{
   __shared__ int64_t sm[ 1024 ];
   int64_t carry = 0;
   for( int64_t b0 = 0; b0 < nb; b0 += 1024 ) {
      const int64_t i = b0 + threadIdx.x;
      const int64_t v = i < nb ? part[ i ] : 0;
      sm[ threadIdx.x ] = v;
      __syncthreads();
      for( int o = 1; o < 1024; o *= 2 ) {
         const int64_t t = threadIdx.x >= o ? sm[ threadIdx.x - o ] : 0;
         __syncthreads();
         sm[ threadIdx.x ] += t;
         __syncthreads();
      }
      if( i < nb )
         part[ i ] = carry + sm[ threadIdx.x ] - v;
      const int64_t blocksum = sm[ 1023 ];
      __syncthreads();
      carry += blocksum;
   }
   if( threadIdx.x == 0 )
      *total = carry;
}

__global__ void __launch_bounds__( kScanBlock ) scan_apply( const int32_t* in, int32_t* out, int64_t n,
                                                            const int64_t* part, const int64_t* total )
{
   __shared__ int64_t sm[ kScanBlock ];
   const int64_t base = (int64_t) blockIdx.x * kScanTile + (int64_t) threadIdx.x * kScanItems;
   int32_t v[ kScanItems ];
   int64_t acc = 0;
#pragma unroll
   for( int i = 0; i < kScanItems; i++ ) {
      v[ i ] = base + i < n ? in[ base + i ] : 0;
      acc += v[ i ];
   }
   sm[ threadIdx.x ] = acc;
   __syncthreads();
   for( int o = 1; o < kScanBlock; o *= 2 ) {
      const int64_t t = threadIdx.x >= o ? sm[ threadIdx.x - o ] : 0;
      __syncthreads();
      sm[ threadIdx.x ] += t;
      __syncthreads();
   }
   int64_t run = part[ blockIdx.x ] + sm[ threadIdx.x ] - acc;
#pragma unroll
   for( int i = 0; i < kScanItems; i++ ) {
      if( base + i < n )
         out[ base + i ] = (int32_t) run;
      run += v[ i ];
   }
   if( blockIdx.x == 0 && threadIdx.x == 0 )
      out[ n ] = (int32_t) *total;
}

}  // namespace

int64_t exclusive_scan_i32( const int32_t* d_in, int32_t* d_out, int64_t n, hipStream_t s )
{
   const int64_t nb = std::max< int64_t >( 1, ceil_div( n, kScanTile ) );
   DeviceArray< int64_t > part;
   part.allocate( nb + 1 );
   int64_t* d_total = part.ptr + nb;
   hipLaunchKernelGGL( scan_partials, dim3( nb ), dim3( kScanBlock ), 0, s, d_in, n, part.ptr );
   hipLaunchKernelGGL( scan_partials_scan, dim3( 1 ), dim3( 1024 ), 0, s, part.ptr, nb, d_total );
   hipLaunchKernelGGL( scan_apply, dim3( nb ), dim3( kScanBlock ), 0, s, d_in, d_out, n, part.ptr, d_total );
   TNLMI_HIP( hipGetLastError() );
   int64_t total = 0;
   TNLMI_HIP( hipMemcpyAsync( &total, d_total, sizeof( int64_t ), hipMemcpyDeviceToHost, s ) );
   TNLMI_HIP( hipStreamSynchronize( s ) );
   return total;
}

SegView make_view( const tnlmi_segments& s )
{
   SegView v{};
   v.format = s.format;
   v.col_major = s.org == TNLMI_COLUMN_MAJOR;
   v.n = s.n;
   v.sizes = s.sizes.ptr;
   v.offsets = s.offsets.ptr;
   v.ell_width = s.ell_width;
   v.ell_aligned = s.ell_aligned;
   v.slice_size = s.param;
   v.slice_offsets = s.slice_offsets.ptr;
   v.slice_seg_sizes = s.slice_seg_sizes.ptr;
   v.perm = s.perm.ptr;
   v.group_ptrs = s.group_ptrs.ptr;
   v.inv_perm = s.inv_perm.ptr;
   v.warp = s.format == TNLMI_BIELLPACK ? s.param : 32;
   v.logw = s.logw;
   v.ch_slices = s.ch_slices.ptr;
   v.seg_to_chunk = s.seg_to_chunk.ptr;
   v.seg_to_slice = s.seg_to_slice.ptr;
   v.chunks_in_slice = s.chunks_in_slice;
   v.sorted_offsets = s.sorted_offsets.ptr;
   return v;
}

// ------------------------------------------------------------ converters
namespace {

__global__ void max_i32( const int32_t* in, int64_t n, int32_t* out )
{
   int32_t m = 0;
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x )
      m = max( m, in[ i ] );
   for( int o = 32; o > 0; o /= 2 )
      m = max( m, __shfl_xor( m, o ) );
   if( ( threadIdx.x & 63 ) == 0 )
      atomicMax( out, m );
}

// CSR tile plan: tile_rows[t] = first row whose offset >= t*T.
__global__ void csr_tile_rows( const int32_t* offsets, int64_t n, int64_t tiles, int32_t T, int32_t* tile_rows )
{
   const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( t > tiles )
      return;
   if( t == tiles ) {
      tile_rows[ t ] = (int32_t) n;
      return;
   }
   const int64_t key = t * (int64_t) T;
   int64_t lo = 0, hi = n;  // lower_bound over offsets[0..n]
   while( lo < hi ) {
      const int64_t mid = ( lo + hi ) / 2;
      if( offsets[ mid ] < key )
         lo = mid + 1;
      else
         hi = mid;
   }
   tile_rows[ t ] = (int32_t) lo;
}

// SlicedEllpack.hpp:122-145: per-slice maximum (virtual rows count as 0).
__global__ void sliced_max( const int32_t* sizes, int64_t n, int32_t S, int64_t slices, int32_t* seg_sizes,
                            int32_t* scaled )
{
   const int64_t s = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( s >= slices )
      return;
   int32_t m = 0;
   const int64_t e = min( n, ( s + 1 ) * (int64_t) S );
   for( int64_t r = s * S; r < e; r++ )
      m = max( m, sizes[ r ] );
   seg_sizes[ s ] = m;
   scaled[ s ] = m * S;
}

// BiEllpack.hpp:245-340: per strip, a stable descending sort of the sizes
// (what the reference's strict-'<' cocktail sort produces), the permutation
// perm[row] = absolute sorted position, and the group sizes width*height.
template< int W >
__global__ void biellpack_strips( const int32_t* sizes, int64_t n, int64_t strips, int logw, int32_t* perm,
                                  int32_t* groups )
{
   const int64_t s = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( s >= strips )
      return;
   const int64_t b = s * W;
   const int cnt = (int) min( (int64_t) W, n - b );
   int32_t sz[ W ];
   int32_t ord[ W ];
   for( int i = 0; i < cnt; i++ ) {
      sz[ i ] = sizes[ b + i ];
      int32_t j = i - 1;
      while( j >= 0 && sz[ ord[ j ] ] < sz[ i ] ) {
         ord[ j + 1 ] = ord[ j ];
         j--;
      }
      ord[ j + 1 ] = i;
   }
   for( int p = 0; p < cnt; p++ )
      perm[ b + ord[ p ] ] = (int32_t) ( b + p );
   int empty = 0;
   if( s == strips - 1 )
      while( cnt <= ( 1 << ( logw - 1 - empty ) ) - 1 )
         empty++;
   int32_t* g = groups + s * ( logw + 1 );
   for( int i = 0; i <= logw; i++ )
      g[ i ] = 0;
   int32_t allocated = 0;
   for( int gi = empty; gi < logw; gi++ ) {
      const int pos = ( 1 << ( logw - 1 - gi ) ) - 1;
      const int32_t len = sz[ ord[ pos ] ];
      g[ gi ] = ( len - allocated ) * ( 1 << ( logw - gi ) );
      allocated = len;
   }
}

__global__ void iota_i32( int32_t* a, int64_t n )
{
   const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( i < n )
      a[ i ] = (int32_t) i;
}

__global__ void invert_perm( const int32_t* inv, int64_t n, int32_t* perm )
{
   const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( i < n )
      perm[ inv[ i ] ] = (int32_t) i;
}

__global__ void negate_i32( const int32_t* a, int64_t n, int32_t* out )
{
   const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( i < n )
      out[ i ] = -a[ i ];
}

template< typename T >
void upload( DeviceArray< T >& d, const std::vector< T >& h, hipStream_t s )
{
   d.allocate( (int64_t) h.size() );
   if( ! h.empty() )
      TNLMI_HIP( hipMemcpyAsync( d.ptr, h.data(), sizeof( T ) * h.size(), hipMemcpyHostToDevice, s ) );
}

// ChunkedEllpack.hpp:156-216 / :281-380 on the host.
void chunked_host( tnlmi_segments& seg, const std::vector< int32_t >& sz, hipStream_t s )
{
   const int64_t n = (int64_t) sz.size();
   const int32_t C = seg.chunks_in_slice;
   const int32_t desired = C * seg.desired_chunk_size;
   std::vector< int32_t > sl, s2c( n ), s2s( n ), sp( n + 1, 0 );
   // slice boundaries: add rows until the slice holds >= desired elements or C rows
   {
      int64_t r = 0;
      while( r < n ) {
         int32_t cnt = 0, alloc = 0;
         do {
            alloc += sz[ r ];
            cnt++;
            r++;
         } while( alloc < desired && r < n && cnt < C );
         sl.insert( sl.end(), { cnt, 0, (int32_t) ( r - cnt ), alloc } );
      }
   }
   const int64_t ns = (int64_t) sl.size() / 4;
   int32_t elements = 0;
   for( int64_t si = 0; si < ns; si++ ) {
      int32_t* S = &sl[ 4 * si ];
      const int32_t cnt = S[ 0 ], b = S[ 2 ], alloc = S[ 3 ], e = b + cnt;
      int32_t free_chunks = C - cnt;
      int32_t added_total = 0, maxlen = sz[ b ];
      for( int32_t i = b; i < e; i++ ) {
         const double ratio = alloc != 0 ? (double) sz[ i ] / (double) alloc : 0.0;
         const int32_t add = (int32_t) ( free_chunks * ratio );  // truncation as in the reference
         s2c[ i ] = 1 + add;
         added_total += add;
         maxlen = std::max( maxlen, sz[ i ] );
      }
      free_chunks -= added_total;
      while( free_chunks )  // leftovers round-robin to the longest rows
         for( int32_t i = b; i < e && free_chunks; i++ )
            if( sz[ i ] == maxlen ) {
               s2c[ i ]++;
               free_chunks--;
            }
      int32_t chunk = 0;
      for( int32_t i = b; i < e; i++ )
         chunk = std::max( chunk, ( sz[ i ] + s2c[ i ] - 1 ) / s2c[ i ] );
      S[ 1 ] = chunk;
      S[ 3 ] = elements;
      elements += C * chunk;
      for( int32_t i = b; i < e; i++ ) {
         s2s[ i ] = (int32_t) si;
         sp[ i + 1 ] = chunk * s2c[ i ];
      }
      for( int32_t i = b + 1; i < e; i++ )
         s2c[ i ] += s2c[ i - 1 ];
   }
   for( int64_t i = 1; i <= n; i++ )
      sp[ i ] += sp[ i - 1 ];
   std::vector< int32_t > c2s( ns * C, 0 );
   int64_t c = 0;
   for( int64_t r = 0; r < n; r++ ) {
      const int32_t first = r != sl[ 4 * s2s[ r ] + 2 ] ? s2c[ r - 1 ] : 0;
      for( int32_t i = first; i < s2c[ r ]; i++ )
         c2s[ c++ ] = (int32_t) r;
   }
   seg.ch_nslices = ns;
   seg.storage = elements;
   upload( seg.ch_slices, sl, s );
   upload( seg.seg_to_chunk, s2c, s );
   upload( seg.seg_to_slice, s2s, s );
   upload( seg.seg_ptrs, sp, s );
   upload( seg.chunks_to_seg, c2s, s );
}

// AdaptiveCSR.hpp:165-231 on the host, with the wave64 device semantics:
// rows per STREAM block bounded by 24576/valueSize/(256/32) elements (the
// compile-time warp size of the HIP host pass is 32), LONG rows padded to the
// end of a 256-thread block of 64-wide waves (4 descriptors per block).
std::vector< int32_t > adaptive_host( const std::vector< int32_t >& off, int value_size )
{
   std::vector< int32_t > out;
   const int64_t n = (int64_t) off.size() - 1;
   if( n <= 0 || off[ n ] == 0 )
      return out;
   const int64_t limit = 24576 / value_size / ( 256 / 32 );
   const int64_t wpb = 256 / kWave;
   auto push = [ & ]( int32_t t, int64_t f, int64_t bs, int64_t ns, int64_t wi, int64_t wc ) {
      out.insert( out.end(), { t, (int32_t) f, (int32_t) bs, (int32_t) ns, (int32_t) wi, (int32_t) wc } );
   };
   int64_t start = 0, next = 0;
   while( next != n ) {
      int type = 1;
      int64_t sum = 0, cur = start;
      next = n;
      for( ; cur < n; cur++ ) {
         sum += off[ cur + 1 ] - off[ cur ];
         if( sum > limit ) {
            if( cur > start )
               next = cur;
            else {
               type = sum <= 2 * 512 ? 2 : 0;
               next = cur + 1;
            }
            break;
         }
      }
      if( type == 0 ) {
         const int64_t cnt = (int64_t) out.size() / 6;
         int64_t left = ceil_div( cnt, wpb ) * wpb - cnt;
         if( left == 0 )
            left = wpb;
         for( int64_t i = 0; i < left; i++ )
            push( 0, start, 0, 0, i, left );
      }
      else
         push( type, start, off[ next ] - off[ start ], next - start, 0, 0 );
      start = next;
   }
   push( 2, next, 0, 0, 0, 0 );
   return out;
}

template< typename T >
std::vector< T > download( const T* d, int64_t n, hipStream_t s )
{
   std::vector< T > h( n );
   if( n )
      TNLMI_HIP( hipMemcpyAsync( h.data(), d, sizeof( T ) * n, hipMemcpyDeviceToHost, s ) );
   TNLMI_HIP( hipStreamSynchronize( s ) );
   return h;
}

void build( tnlmi_segments& seg, const int32_t* d_sizes, hipStream_t s )
{
   const int64_t n = seg.n;
   seg.sizes.allocate( n );
   if( n )
      TNLMI_HIP( hipMemcpyAsync( seg.sizes.ptr, d_sizes, sizeof( int32_t ) * n, hipMemcpyDeviceToDevice, s ) );
   seg.offsets.allocate( n + 1 );
   const int64_t total = exclusive_scan_i32( seg.sizes.ptr, seg.offsets.ptr, n, s );
   if( total > INT32_MAX )
      fail( TNLMI_OVERFLOW_ERROR, "segments: the number of elements exceeds the int32 index range" );
   seg.elements = total;
   seg.storage = total;
   {
      DeviceArray< int32_t > m;
      m.allocate( 1 );
      TNLMI_HIP( hipMemsetAsync( m.ptr, 0, 4, s ) );
      if( n )
         hipLaunchKernelGGL( max_i32, dim3( std::min< int64_t >( 1024, ceil_div( n, 256 ) ) ), dim3( 256 ), 0, s,
                             seg.sizes.ptr, n, m.ptr );
      seg.max_size = download( m.ptr, 1, s )[ 0 ];
   }
   const int blocks256 = (int) std::max< int64_t >( 1, ceil_div( n, 256 ) );
   switch( seg.format ) {
      case TNLMI_CSR:
      case TNLMI_ADAPTIVE_CSR: {
         seg.tile_nnz = csr_tile_size( seg.max_size, total );
         seg.tiles = std::max< int64_t >( 1, ceil_div( total, seg.tile_nnz ) );
         seg.tile_rows.allocate( seg.tiles + 1 );
         hipLaunchKernelGGL( csr_tile_rows, dim3( ceil_div( seg.tiles + 1, 256 ) ), dim3( 256 ), 0, s,
                             seg.offsets.ptr, n, seg.tiles, seg.tile_nnz, seg.tile_rows.ptr );
         seg.carry_row.allocate( seg.tiles );
         seg.carry_val.allocate( seg.tiles );
         if( seg.format == TNLMI_ADAPTIVE_CSR ) {
            const auto off = download( seg.offsets.ptr, n + 1, s );
            const int vs[ 6 ] = { 1, 2, 4, 8, 16, 32 };
            for( int i = 0; i < 6; i++ ) {
               const auto b = adaptive_host( off, vs[ i ] );
               upload( seg.blocks[ i ], b, s );
               seg.nblocks[ i ] = (int64_t) b.size() / 6;
            }
         }
         break;
      }
      case TNLMI_ELLPACK: {
         // Ellpack.hpp:81-97
         seg.ell_width = seg.max_size;
         const int64_t aligned = seg.org == TNLMI_ROW_MAJOR ? n : ceil_div( n, 32 ) * 32;
         if( aligned * (int64_t) seg.ell_width > INT32_MAX )
            fail( TNLMI_OVERFLOW_ERROR,
                  "Ellpack: multiplication overflow - the storage size required for the segments is larger than "
                  "the maximal value of used index type." );
         seg.ell_aligned = (int32_t) aligned;
         seg.storage = aligned * seg.ell_width;
         break;
      }
      case TNLMI_SLICED_ELLPACK: {
         const int32_t S = seg.param;
         seg.slices = ceil_div( n, S );
         seg.slice_seg_sizes.allocate( seg.slices );
         seg.slice_offsets.allocate( seg.slices + 1 );
         DeviceArray< int32_t > scaled;
         scaled.allocate( std::max< int64_t >( 1, seg.slices ) );
         if( seg.slices )
            hipLaunchKernelGGL( sliced_max, dim3( ceil_div( seg.slices, 256 ) ), dim3( 256 ), 0, s, seg.sizes.ptr, n,
                                S, seg.slices, seg.slice_seg_sizes.ptr, scaled.ptr );
         const int64_t st = exclusive_scan_i32( scaled.ptr, seg.slice_offsets.ptr, seg.slices, s );
         if( st > INT32_MAX )
            fail( TNLMI_OVERFLOW_ERROR, "SlicedEllpack: storage size exceeds the int32 index range" );
         seg.storage = st;
         break;
      }
      case TNLMI_BIELLPACK: {
         const int32_t W = seg.param;
         seg.logw = W == 32 ? 5 : 6;
         seg.strips = ceil_div( n, W );
         seg.perm.allocate( n );
         const int64_t ng = seg.strips * ( seg.logw + 1 );
         seg.group_ptrs.allocate( ng + 1 );
         DeviceArray< int32_t > groups;
         groups.allocate( std::max< int64_t >( 1, ng ) );
         if( seg.strips ) {
            if( W == 32 )
               hipLaunchKernelGGL( biellpack_strips< 32 >, dim3( ceil_div( seg.strips, 64 ) ), dim3( 64 ), 0, s,
                                   seg.sizes.ptr, n, seg.strips, seg.logw, seg.perm.ptr, groups.ptr );
            else
               hipLaunchKernelGGL( biellpack_strips< 64 >, dim3( ceil_div( seg.strips, 64 ) ), dim3( 64 ), 0, s,
                                   seg.sizes.ptr, n, seg.strips, seg.logw, seg.perm.ptr, groups.ptr );
         }
         const int64_t st = exclusive_scan_i32( groups.ptr, seg.group_ptrs.ptr, ng, s );
         if( st > INT32_MAX )
            fail( TNLMI_OVERFLOW_ERROR, "BiEllpack: storage size exceeds the int32 index range" );
         seg.storage = st;
         break;
      }
      case TNLMI_CHUNKED_ELLPACK: {
         chunked_host( seg, download( seg.sizes.ptr, n, s ), s );
         break;
      }
      case TNLMI_SORTED_CSR: {
         // SortedSegments.hpp:129-200: stable sort by size descending (ties keep
         // index order), inverse permutation, embedded CSR of the sorted sizes.
         seg.perm.allocate( n );
         seg.inv_perm.allocate( n );
         seg.sorted_offsets.allocate( n + 1 );
         DeviceArray< int32_t > keys_out, vals_in;
         keys_out.allocate( std::max< int64_t >( 1, n ) );
         vals_in.allocate( std::max< int64_t >( 1, n ) );
         if( n ) {
            hipLaunchKernelGGL( iota_i32, dim3( blocks256 ), dim3( 256 ), 0, s, vals_in.ptr, n );
            size_t tmp_bytes = 0;
            TNLMI_HIP( hipcub::DeviceRadixSort::SortPairsDescending( nullptr, tmp_bytes, seg.sizes.ptr, keys_out.ptr,
                                                                    vals_in.ptr, seg.inv_perm.ptr, (int) n, 0, 32,
                                                                    s ) );
            DeviceArray< char > tmp;
            tmp.allocate( (int64_t) tmp_bytes + 1 );
            TNLMI_HIP( hipcub::DeviceRadixSort::SortPairsDescending( tmp.ptr, tmp_bytes, seg.sizes.ptr, keys_out.ptr,
                                                                    vals_in.ptr, seg.inv_perm.ptr, (int) n, 0, 32,
                                                                    s ) );
            hipLaunchKernelGGL( invert_perm, dim3( blocks256 ), dim3( 256 ), 0, s, seg.inv_perm.ptr, n, seg.perm.ptr );
         }
         exclusive_scan_i32( keys_out.ptr, seg.sorted_offsets.ptr, n, s );
         // the tiled CSR plan over the embedded (sorted) CSR
         seg.tile_nnz = csr_tile_size( seg.max_size, total );
         seg.tiles = std::max< int64_t >( 1, ceil_div( total, seg.tile_nnz ) );
         seg.tile_rows.allocate( seg.tiles + 1 );
         hipLaunchKernelGGL( csr_tile_rows, dim3( ceil_div( seg.tiles + 1, 256 ) ), dim3( 256 ), 0, s,
                             seg.sorted_offsets.ptr, n, seg.tiles, seg.tile_nnz, seg.tile_rows.ptr );
         seg.carry_row.allocate( seg.tiles );
         seg.carry_val.allocate( seg.tiles );
         break;
      }
      default:
         fail( TNLMI_INVALID_ARGUMENT, "segments: unknown format" );
   }
   TNLMI_HIP( hipGetLastError() );
   TNLMI_HIP( hipStreamSynchronize( s ) );
}

// ---------------------------------------------------------------- fill
template< typename S, typename D >
__device__ __forceinline__ D convert_value( S v );
template<>
__device__ __forceinline__ double convert_value< double, double >( double v )
{
   return v;
}
template<>
__device__ __forceinline__ float convert_value< double, float >( double v )
{
   return (float) v;
}
template<>
__device__ __forceinline__ bf16 convert_value< double, bf16 >( double v )
{
   return bf16{ float_to_bf16( (float) v ) };
}
template<>
__device__ __forceinline__ double convert_value< float, double >( float v )
{
   return v;
}
template<>
__device__ __forceinline__ float convert_value< float, float >( float v )
{
   return v;
}
template<>
__device__ __forceinline__ bf16 convert_value< float, bf16 >( float v )
{
   return bf16{ float_to_bf16( v ) };
}

template< typename D >
__global__ void clear_storage( D* vals, int32_t* cols, int64_t storage )
{
   for( int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < storage; i += (int64_t) gridDim.x * blockDim.x ) {
      vals[ i ] = D{};
      cols[ i ] = -1;
   }
}

template< typename S, typename D >
__global__ void fill_rows( SegView v, int64_t storage, const int32_t* csr_off, const int32_t* csr_cols,
                           const S* csr_vals, D* vals, int32_t* cols )
{
   const int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( r >= v.n )
      return;
   const int32_t b = csr_off[ r ];
   const int32_t e = min( csr_off[ r + 1 ], b + v.sizes[ r ] );  // never beyond the row capacity
   for( int32_t k = 0; k < e - b; k++ ) {
      const int64_t g = global_index( v, r, k );
      if( g < 0 || g >= storage )
         continue;
      cols[ g ] = csr_cols[ b + k ];
      vals[ g ] = convert_value< S, D >( csr_vals[ b + k ] );
   }
}

__global__ void gidx_rows( SegView v, int32_t* out )
{
   const int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( r >= v.n )
      return;
   const int32_t b = v.offsets[ r ], e = v.offsets[ r + 1 ];
   for( int32_t k = 0; k < e - b; k++ )
      out[ b + k ] = (int32_t) global_index( v, r, k );
}

template< typename S, typename D >
void fill_typed( const tnlmi_segments& seg, const int32_t* off, const int32_t* c, const void* v, void* dv, int32_t* dc,
                 hipStream_t s )
{
   if( seg.storage )
      hipLaunchKernelGGL( clear_storage< D >, dim3( std::min< int64_t >( 4096, ceil_div( seg.storage, 256 ) ) ),
                          dim3( 256 ), 0, s, (D*) dv, dc, seg.storage );
   if( seg.n )
      hipLaunchKernelGGL( ( fill_rows< S, D > ), dim3( ceil_div( seg.n, 256 ) ), dim3( 256 ), 0, s, make_view( seg ),
                          seg.storage, off, c, (const S*) v, (D*) dv, dc );
   TNLMI_HIP( hipGetLastError() );
}

}  // namespace

// ------------------------------------------------------------ file arrays
namespace {

[[noreturn]] void bad_file( const std::string& what )
{
   fail( TNLMI_RUNTIME_ERROR, "matrix file: " + what );
}

bool from_zero_nondecreasing( const std::vector< int32_t >& a )
{
   if( a.empty() || a[ 0 ] != 0 )
      return false;
   for( size_t i = 1; i < a.size(); i++ )
      if( a[ i ] < a[ i - 1 ] )
         return false;
   return true;
}

}  // namespace

void segments_export( const tnlmi_segments& seg, SegmentsFileData& f, hipStream_t s )
{
   const int64_t n = seg.n;
   f = SegmentsFileData{};
   f.elements = seg.file_size >= 0 ? seg.file_size : seg.elements;
   f.storage = seg.storage;
   switch( seg.format ) {
      case TNLMI_CSR:
      case TNLMI_ADAPTIVE_CSR:
         f.offsets = download( seg.offsets.ptr, n + 1, s );
         if( seg.format == TNLMI_ADAPTIVE_CSR )
            for( int i = 0; i < 6; i++ )
               f.blocks[ i ] = download( seg.blocks[ i ].ptr, seg.nblocks[ i ] * 6, s );
         break;
      case TNLMI_ELLPACK:
         f.ell_width = seg.ell_width;
         f.ell_aligned = seg.ell_aligned;
         f.ell_count = (int32_t) n;
         break;
      case TNLMI_SLICED_ELLPACK:
         f.slice_offsets = download( seg.slice_offsets.ptr, seg.slices + 1, s );
         f.slice_seg_sizes = download( seg.slice_seg_sizes.ptr, seg.slices, s );
         break;
      case TNLMI_BIELLPACK:
         f.perm = download( seg.perm.ptr, n, s );
         f.group_ptrs = download( seg.group_ptrs.ptr, seg.group_ptrs.size, s );
         break;
      case TNLMI_CHUNKED_ELLPACK:
         f.ch_nslices = (int32_t) seg.ch_nslices;
         f.chunks_in_slice = seg.chunks_in_slice;
         f.desired_chunk_size = seg.desired_chunk_size;
         f.seg_to_chunk = download( seg.seg_to_chunk.ptr, n, s );
         f.seg_to_slice = download( seg.seg_to_slice.ptr, n, s );
         f.chunks_to_seg = download( seg.chunks_to_seg.ptr, seg.chunks_to_seg.size, s );
         f.seg_ptrs = download( seg.seg_ptrs.ptr, n + 1, s );
         f.ch_slices = download( seg.ch_slices.ptr, 4 * seg.ch_nslices, s );
         // TNL sizes its slice array with one entry per segment
         // (ChunkedEllpack.hpp:163); the entries past the last slice are unused
         f.ch_slices.resize( 4 * std::max< int64_t >( n, seg.ch_nslices ), 0 );
         break;
      case TNLMI_SORTED_CSR:
         f.offsets = download( seg.sorted_offsets.ptr, n + 1, s );
         f.perm = download( seg.perm.ptr, n, s );
         f.inv_perm = download( seg.inv_perm.ptr, n, s );
         break;
   }
}

void segments_import( tnlmi_segments& seg, const SegmentsFileData& f, hipStream_t s )
{
   const int64_t n = seg.n;
   auto check = []( bool ok, const char* what ) {
      if( ! ok )
         bad_file( what );
   };
   check( f.storage >= 0 && f.storage <= INT32_MAX && f.elements >= 0, "storage size out of the int32 index range" );
   std::vector< int32_t > sizes( n );
   SegView v{};  // host view of the file's arrays, for the index check below
   v.format = seg.format;
   v.col_major = seg.org == TNLMI_COLUMN_MAJOR;
   v.n = n;
   switch( seg.format ) {
      case TNLMI_CSR:
      case TNLMI_ADAPTIVE_CSR:
         check( (int64_t) f.offsets.size() == n + 1 && from_zero_nondecreasing( f.offsets ),
                "CSR offsets are not rows + 1 nondecreasing values from 0" );
         for( int64_t r = 0; r < n; r++ )
            sizes[ r ] = f.offsets[ r + 1 ] - f.offsets[ r ];
         break;
      case TNLMI_ELLPACK: {
         const int64_t aligned = seg.org == TNLMI_ROW_MAJOR ? n : ceil_div( n, 32 ) * 32;
         check( f.ell_count == n, "Ellpack segment count differs from the matrix rows" );
         check( f.ell_width >= 0 && f.ell_aligned == aligned, "Ellpack segment size or aligned size out of range" );
         check( aligned * f.ell_width <= INT32_MAX, "Ellpack storage out of the int32 index range" );
         std::fill( sizes.begin(), sizes.end(), f.ell_width );
         v.ell_width = f.ell_width;
         v.ell_aligned = f.ell_aligned;
         break;
      }
      case TNLMI_SLICED_ELLPACK: {
         const int64_t S = seg.param, ns = ceil_div( n, S );
         check( (int64_t) f.slice_seg_sizes.size() == ns && (int64_t) f.slice_offsets.size() == ns + 1
                   && from_zero_nondecreasing( f.slice_offsets ) && f.slice_offsets.back() == f.storage,
                "SlicedEllpack slice arrays do not match the rows and the storage size" );
         for( int64_t q = 0; q < ns; q++ )  // SlicedEllpack.hpp:137-141: slice q holds S rows of its width
            check( f.slice_seg_sizes[ q ] >= 0
                      && (int64_t) f.slice_offsets[ q + 1 ] - f.slice_offsets[ q ] == S * f.slice_seg_sizes[ q ],
                   "SlicedEllpack slice offsets do not match the slice widths" );
         for( int64_t r = 0; r < n; r++ )
            sizes[ r ] = f.slice_seg_sizes[ r / S ];
         v.slice_size = (int32_t) S;
         v.slice_offsets = f.slice_offsets.data();
         v.slice_seg_sizes = f.slice_seg_sizes.data();
         break;
      }
      case TNLMI_BIELLPACK: {
         // BiEllpack.h getSegmentSizeDirect: the widths of the row's active groups
         const int32_t W = seg.param, logw = W == 32 ? 5 : 6;
         const int64_t strips = ceil_div( n, W );
         check( (int64_t) f.perm.size() == n && (int64_t) f.group_ptrs.size() == strips * ( logw + 1 ) + 1
                   && from_zero_nondecreasing( f.group_ptrs ) && f.group_ptrs.back() == f.storage,
                "BiEllpack arrays do not match the rows and the storage size" );
         std::vector< char > seen( W );
         for( int64_t st = 0; st < strips; st++ ) {
            std::fill( seen.begin(), seen.end(), 0 );
            for( int64_t r = st * W; r < std::min< int64_t >( n, ( st + 1 ) * W ); r++ ) {
               const int64_t p = (int64_t) f.perm[ r ] - st * W;
               check( p >= 0 && p < W && ! seen[ p ], "BiEllpack permutation is not a permutation of its strip" );
               seen[ p ] = 1;
            }
         }
         for( int64_t r = 0; r < n; r++ ) {
            const int64_t st = r / W, g0 = st * ( logw + 1 );
            const int32_t rsp = f.perm[ r ] - (int32_t) ( st * W );
            int groups = 1;
            for( int i = 0, bis = 1; i < logw + 1; i++, bis *= 2 )
               if( rsp < bis ) {
                  groups = logw + 1 - i;
                  break;
               }
            int64_t size = 0;
            for( int g = 0, height = W; g < groups; g++, height /= 2 )
               size += ( f.group_ptrs[ g0 + g + 1 ] - f.group_ptrs[ g0 + g ] ) / height;
            check( size <= INT32_MAX, "BiEllpack segment size out of range" );
            sizes[ r ] = (int32_t) size;
         }
         v.perm = f.perm.data();
         v.group_ptrs = f.group_ptrs.data();
         v.warp = W;
         v.logw = logw;
         break;
      }
      case TNLMI_CHUNKED_ELLPACK: {
         // ChunkedEllpack.h getSegmentSizeDirect: chunks of the row x the slice's chunk size
         const int64_t ns = f.ch_nslices, C = f.chunks_in_slice;
         check( ns >= 0 && C > 0 && f.desired_chunk_size > 0 && (int64_t) f.seg_to_chunk.size() == n
                   && (int64_t) f.seg_to_slice.size() == n && (int64_t) f.seg_ptrs.size() == n + 1
                   && (int64_t) f.chunks_to_seg.size() == ns * C && (int64_t) f.ch_slices.size() >= 4 * ns,
                "ChunkedEllpack arrays do not match the rows and the slice count" );
         // the SpMV kernel sums every chunk of a slice (one thread per chunk,
         // one per row): TNL's 256 chunks per slice, slices covering the rows
         // in order, each slice's chunk region inside the storage
         check( C == 256, "ChunkedEllpack with other than 256 chunks per slice" );
         int64_t next = 0;
         for( int64_t si = 0; si < ns; si++ ) {
            const int32_t* S = &f.ch_slices[ 4 * si ];
            check( S[ 0 ] >= 0 && S[ 0 ] <= C && S[ 2 ] == next && S[ 1 ] >= 0 && S[ 3 ] >= 0
                      && (int64_t) S[ 3 ] + C * (int64_t) S[ 1 ] <= f.storage,
                   "ChunkedEllpack slice descriptor out of range" );
            next += S[ 0 ];
         }
         check( next == n, "ChunkedEllpack slices do not cover the rows" );
         for( int64_t r = 0; r < n; r++ ) {
            const int64_t si = f.seg_to_slice[ r ];
            check( si >= 0 && si < ns, "ChunkedEllpack segment-to-slice mapping out of range" );
            const int32_t* S = &f.ch_slices[ 4 * si ];
            check( S[ 2 ] <= r && r < S[ 2 ] + S[ 0 ], "ChunkedEllpack segment-to-slice mapping out of range" );
            const int64_t first = r != S[ 2 ] ? f.seg_to_chunk[ r - 1 ] : 0, last = f.seg_to_chunk[ r ];
            check( first >= 0 && first <= last && last <= C, "ChunkedEllpack segment-to-chunk mapping out of range" );
            check( ( last - first ) * S[ 1 ] <= INT32_MAX, "ChunkedEllpack segment size out of range" );
            sizes[ r ] = (int32_t) ( ( last - first ) * S[ 1 ] );
         }
         v.ch_slices = f.ch_slices.data();
         v.seg_to_chunk = f.seg_to_chunk.data();
         v.seg_to_slice = f.seg_to_slice.data();
         v.chunks_in_slice = f.chunks_in_slice;
         break;
      }
      case TNLMI_SORTED_CSR: {
         check( (int64_t) f.offsets.size() == n + 1 && from_zero_nondecreasing( f.offsets )
                   && (int64_t) f.perm.size() == n && (int64_t) f.inv_perm.size() == n,
                "SortedSegments arrays do not match the rows" );
         for( int64_t r = 0; r < n; r++ )
            check( f.perm[ r ] >= 0 && f.perm[ r ] < n && f.inv_perm[ f.perm[ r ] ] == r,
                   "SortedSegments permutations are not inverse permutations" );
         for( int64_t r = 0; r < n; r++ )
            sizes[ r ] = f.offsets[ f.perm[ r ] + 1 ] - f.offsets[ f.perm[ r ] ];
         break;
      }
      default:
         fail( TNLMI_INVALID_ARGUMENT, "segments: unknown format" );
   }
   // every element of every segment inside the storage (the kernels index
   // with these arrays unchecked)
   if( seg.format == TNLMI_CSR || seg.format == TNLMI_ADAPTIVE_CSR || seg.format == TNLMI_SORTED_CSR )
      check( f.offsets.back() == f.storage, "CSR offsets do not end at the storage size" );
   else
      for( int64_t r = 0; r < n; r++ )
         for( int32_t k = 0; k < sizes[ r ]; k++ ) {
            const int64_t g = global_index( v, r, k );
            check( g >= 0 && g < f.storage, "a segment element lies outside the storage" );
         }
   // the format built from these sizes, then the file's own arrays
   DeviceArray< int32_t > d;
   upload( d, sizes, s );
   build( seg, d.ptr, s );
   switch( seg.format ) {
      case TNLMI_ELLPACK:
         seg.ell_width = f.ell_width;
         seg.ell_aligned = f.ell_aligned;
         seg.storage = (int64_t) f.ell_aligned * f.ell_width;
         break;
      case TNLMI_SLICED_ELLPACK:
         upload( seg.slice_offsets, f.slice_offsets, s );
         upload( seg.slice_seg_sizes, f.slice_seg_sizes, s );
         seg.storage = f.storage;
         seg.file_size = f.elements;
         break;
      case TNLMI_BIELLPACK:
         upload( seg.perm, f.perm, s );
         upload( seg.group_ptrs, f.group_ptrs, s );
         seg.storage = f.storage;
         seg.file_size = f.elements;
         break;
      case TNLMI_CHUNKED_ELLPACK: {
         seg.ch_nslices = f.ch_nslices;
         seg.chunks_in_slice = f.chunks_in_slice;
         seg.desired_chunk_size = f.desired_chunk_size;
         upload( seg.seg_to_chunk, f.seg_to_chunk, s );
         upload( seg.seg_to_slice, f.seg_to_slice, s );
         upload( seg.chunks_to_seg, f.chunks_to_seg, s );
         upload( seg.seg_ptrs, f.seg_ptrs, s );
         upload( seg.ch_slices, std::vector< int32_t >( f.ch_slices.begin(), f.ch_slices.begin() + 4 * f.ch_nslices ), s );
         seg.storage = f.storage;
         seg.file_size = f.elements;
         break;
      }
      case TNLMI_SORTED_CSR:
         // the file's order (a TNL object sorted in pieces keeps its own)
         upload( seg.perm, f.perm, s );
         upload( seg.inv_perm, f.inv_perm, s );
         upload( seg.sorted_offsets, f.offsets, s );
         if( seg.tiles > 0 )
            hipLaunchKernelGGL( csr_tile_rows, dim3( ceil_div( seg.tiles + 1, 256 ) ), dim3( 256 ), 0, s,
                                seg.sorted_offsets.ptr, n, seg.tiles, seg.tile_nnz, seg.tile_rows.ptr );
         break;
      default:
         break;
   }
   TNLMI_HIP( hipGetLastError() );
   TNLMI_HIP( hipStreamSynchronize( s ) );
}

}  // namespace tnlmi

using namespace tnlmi;

extern "C" {

tnlmi_status tnlmi_segments_create( tnlmi_format format, tnlmi_organization organization, int32_t param,
                                    const int32_t* d_sizes, int64_t n, tnlmi_stream stream, tnlmi_segments** out )
{
   return guarded( [ & ] {
      if( ! out )
         fail( TNLMI_INVALID_ARGUMENT, "segments_create: out is NULL" );
      *out = nullptr;
      if( n < 0 || n > INT32_MAX )
         fail( TNLMI_INVALID_ARGUMENT, "segments_create: segment count out of range" );
      if( n > 0 && ! d_sizes )
         fail( TNLMI_INVALID_ARGUMENT, "segments_create: sizes is NULL" );
      if( organization != TNLMI_COLUMN_MAJOR && organization != TNLMI_ROW_MAJOR )
         fail( TNLMI_INVALID_ARGUMENT, "segments_create: unknown organization" );
      auto* seg = new tnlmi_segments;
      seg->format = format;
      seg->org = organization;
      seg->n = n;
      if( format == TNLMI_SLICED_ELLPACK ) {
         seg->param = param ? param : 32;
         if( seg->param <= 0 || ( seg->param & ( seg->param - 1 ) ) ) {
            delete seg;
            fail( TNLMI_INVALID_ARGUMENT, "SlicedEllpack: slice size must be a positive power of two" );
         }
      }
      else if( format == TNLMI_BIELLPACK ) {
         seg->param = param ? param : 32;
         if( seg->param != 32 && seg->param != 64 ) {
            delete seg;
            fail( TNLMI_NOT_IMPLEMENTED, "BiEllpack: WarpSize must be 32 or 64" );
         }
      }
      try {
         build( *seg, d_sizes, as_stream( stream ) );
      }
      catch( ... ) {
         delete seg;
         throw;
      }
      *out = seg;
   } );
}

void tnlmi_segments_destroy( tnlmi_segments* segments )
{
   delete segments;
}

int64_t tnlmi_segments_count( const tnlmi_segments* s )
{
   return s ? s->n : 0;
}
int64_t tnlmi_segments_storage_size( const tnlmi_segments* s )
{
   return s ? s->storage : 0;
}
int64_t tnlmi_segments_element_count( const tnlmi_segments* s )
{
   return s ? s->elements : 0;
}
tnlmi_format tnlmi_segments_format( const tnlmi_segments* s )
{
   return s ? s->format : TNLMI_CSR;
}

tnlmi_status tnlmi_segments_get_view( const tnlmi_segments* s, tnlmi_segments_view* out )
{
   return guarded( [ & ] {
      if( ! s || ! out )
         fail( TNLMI_INVALID_ARGUMENT, "segments_get_view: NULL argument" );
      *out = tnlmi::make_view( *s );
   } );
}

tnlmi_status tnlmi_segments_get_array( const tnlmi_segments* s, tnlmi_segments_array which, const int32_t** d_array,
                                       int64_t* length )
{
   return guarded( [ & ] {
      if( ! s || ! d_array || ! length )
         fail( TNLMI_INVALID_ARGUMENT, "segments_get_array: NULL argument" );
      const DeviceArray< int32_t >* a = nullptr;
      const tnlmi_format f = s->format;
      const bool csr_like = f == TNLMI_CSR || f == TNLMI_ADAPTIVE_CSR;
      switch( which ) {
         case TNLMI_ARR_OFFSETS:
            a = f == TNLMI_SORTED_CSR ? &s->sorted_offsets : csr_like ? &s->offsets : nullptr;
            break;
         case TNLMI_ARR_SLICE_OFFSETS:
            a = f == TNLMI_SLICED_ELLPACK ? &s->slice_offsets : nullptr;
            break;
         case TNLMI_ARR_SLICE_SEGMENT_SIZES:
            a = f == TNLMI_SLICED_ELLPACK ? &s->slice_seg_sizes : nullptr;
            break;
         case TNLMI_ARR_PERMUTATION:
            a = f == TNLMI_BIELLPACK || f == TNLMI_SORTED_CSR ? &s->perm : nullptr;
            break;
         case TNLMI_ARR_GROUP_POINTERS:
            a = f == TNLMI_BIELLPACK ? &s->group_ptrs : nullptr;
            break;
         case TNLMI_ARR_INVERSE_PERMUTATION:
            a = f == TNLMI_SORTED_CSR ? &s->inv_perm : nullptr;
            break;
         case TNLMI_ARR_CHUNK_SLICES:
            a = f == TNLMI_CHUNKED_ELLPACK ? &s->ch_slices : nullptr;
            break;
         case TNLMI_ARR_SEGMENT_TO_CHUNK:
            a = f == TNLMI_CHUNKED_ELLPACK ? &s->seg_to_chunk : nullptr;
            break;
         case TNLMI_ARR_SEGMENT_TO_SLICE:
            a = f == TNLMI_CHUNKED_ELLPACK ? &s->seg_to_slice : nullptr;
            break;
         case TNLMI_ARR_SEGMENT_POINTERS:
            a = f == TNLMI_CHUNKED_ELLPACK ? &s->seg_ptrs : nullptr;
            break;
         case TNLMI_ARR_CHUNKS_TO_SEGMENTS:
            a = f == TNLMI_CHUNKED_ELLPACK ? &s->chunks_to_seg : nullptr;
            break;
         case TNLMI_ARR_ADAPTIVE_BLOCKS:
            a = f == TNLMI_ADAPTIVE_CSR ? &s->blocks[ 3 ] : nullptr;
            break;
         case TNLMI_ARR_SIZES:
            a = &s->sizes;
            break;
         case TNLMI_ARR_TILE_ROWS:
            a = csr_like || f == TNLMI_SORTED_CSR ? &s->tile_rows : nullptr;
            break;
      }
      if( ! a )
         fail( TNLMI_INVALID_ARGUMENT, "segments_get_array: the format has no such array" );
      *d_array = a->ptr;
      *length = a->size;
   } );
}

tnlmi_status tnlmi_segments_ellpack_dims( const tnlmi_segments* s, int32_t* segment_size, int32_t* aligned_size )
{
   return guarded( [ & ] {
      if( ! s || s->format != TNLMI_ELLPACK )
         fail( TNLMI_INVALID_ARGUMENT, "segments_ellpack_dims: not an Ellpack segments object" );
      *segment_size = s->ell_width;
      *aligned_size = s->ell_aligned;
   } );
}

tnlmi_status tnlmi_segments_adaptive_blocks( const tnlmi_segments* s, int32_t value_bytes, const int32_t** d_blocks,
                                             int64_t* count )
{
   return guarded( [ & ] {
      if( ! s || s->format != TNLMI_ADAPTIVE_CSR )
         fail( TNLMI_INVALID_ARGUMENT, "segments_adaptive_blocks: not an AdaptiveCSR segments object" );
      const int l = value_bytes <= 1 ? 0 : value_bytes <= 2 ? 1 : value_bytes <= 4 ? 2 : value_bytes <= 8 ? 3
                                                             : value_bytes <= 16 ? 4 : 5;
      *d_blocks = s->blocks[ l ].ptr;
      *count = s->nblocks[ l ];
   } );
}

tnlmi_status tnlmi_segments_global_indexes( const tnlmi_segments* s, int32_t* d_gidx, tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( ! s || ( s->elements && ! d_gidx ) )
         fail( TNLMI_INVALID_ARGUMENT, "segments_global_indexes: NULL argument" );
      if( s->n )
         hipLaunchKernelGGL( gidx_rows, dim3( ceil_div( s->n, 256 ) ), dim3( 256 ), 0, as_stream( stream ),
                             make_view( *s ), d_gidx );
      TNLMI_HIP( hipGetLastError() );
   } );
}

tnlmi_status tnlmi_matrix_fill( const tnlmi_segments* seg, const int32_t* d_csr_offsets, const int32_t* d_csr_columns,
                                const void* d_csr_values, tnlmi_dtype csr_type, void* d_values, tnlmi_dtype value_type,
                                int32_t* d_column_indexes, tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( ! seg )
         fail( TNLMI_INVALID_ARGUMENT, "matrix_fill: segments is NULL" );
      if( seg->storage && ( ! d_values || ! d_column_indexes ) )
         fail( TNLMI_INVALID_ARGUMENT, "matrix_fill: storage arrays are NULL" );
      if( seg->n && ! d_csr_offsets )
         fail( TNLMI_INVALID_ARGUMENT, "matrix_fill: CSR offsets are NULL" );
      hipStream_t s = as_stream( stream );
      drop_auto_plans( *seg, d_values, d_column_indexes );
      if( csr_type == TNLMI_F64 && value_type == TNLMI_F64 )
         fill_typed< double, double >( *seg, d_csr_offsets, d_csr_columns, d_csr_values, d_values, d_column_indexes, s );
      else if( csr_type == TNLMI_F64 && value_type == TNLMI_F32 )
         fill_typed< double, float >( *seg, d_csr_offsets, d_csr_columns, d_csr_values, d_values, d_column_indexes, s );
      else if( csr_type == TNLMI_F64 && value_type == TNLMI_BF16 )
         fill_typed< double, bf16 >( *seg, d_csr_offsets, d_csr_columns, d_csr_values, d_values, d_column_indexes, s );
      else if( csr_type == TNLMI_F32 && value_type == TNLMI_F32 )
         fill_typed< float, float >( *seg, d_csr_offsets, d_csr_columns, d_csr_values, d_values, d_column_indexes, s );
      else if( csr_type == TNLMI_F32 && value_type == TNLMI_F64 )
         fill_typed< float, double >( *seg, d_csr_offsets, d_csr_columns, d_csr_values, d_values, d_column_indexes, s );
      else if( csr_type == TNLMI_F32 && value_type == TNLMI_BF16 )
         fill_typed< float, bf16 >( *seg, d_csr_offsets, d_csr_columns, d_csr_values, d_values, d_column_indexes, s );
      else
         fail( TNLMI_NOT_IMPLEMENTED, "matrix_fill: unsupported value type combination" );
   } );
}

}  // extern "C"
