/*
 * oracle.cpp -- CPU restatement of TNL's SpMV / segments / Krylov hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Built with -ffp-contract=off so
 * that, like TNL's default x86-64 build (no -march=native), no multiply-add is
 * fused.  Reference paths are relative to src/TNL/ of the TNL tree.
 */
#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>
#include <random>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using I = int32_t;

int ilog2( int w )
{
   int l = 0;
   while( ( 1 << l ) < w )
      l++;
   return l;
}

// Algorithms/Segments/BiEllpackBase.hpp:194-200
int64_t biell_virtual( int64_t n, int warp )
{
   return n % warp ? n + warp - n % warp : n;
}

// Algorithms/Segments/detail/BiEllpack.h getActiveGroupsCount
int biell_active_groups( I rowStripPerm, int logw )
{
   I bisection = 1;
   for( int i = 0; i < logw + 1; i++ ) {
      if( rowStripPerm < bisection )
         return logw + 1 - i;
      bisection *= 2;
   }
   return -1;
}

}  // namespace

extern "C" {

/* ------------------------------------------------------------------ CSR */
void orc_csr_offsets( const int32_t* sizes, int64_t n, int32_t* offsets )
{
   // CSR.hpp:86-102: offsets = [sizes..., 0], inplaceExclusiveScan
   int32_t acc = 0;
   for( int64_t i = 0; i < n; i++ ) {
      offsets[ i ] = acc;
      acc += sizes[ i ];
   }
   offsets[ n ] = acc;
}

/* -------------------------------------------------------------- Ellpack */
int orc_ellpack_dims( const int32_t* sizes, int64_t n, int col_major, int alignment, int32_t* segment_size,
                      int32_t* aligned_size )
{
   // Ellpack.hpp:72-97: segmentSize = max(sizes); row-major alignedSize = n,
   // column-major alignedSize = roundUpDivision(n, Alignment) * Alignment.
   // max() of an empty vector is the identity of the Max reduction, the
   // lowest int32 value; the storage size is 0 either way.
   int32_t w = n ? 0 : INT32_MIN;
   for( int64_t i = 0; i < n; i++ )
      w = std::max( w, sizes[ i ] );
   int64_t aligned = col_major ? ( n + alignment - 1 ) / alignment * alignment : n;
   *segment_size = w;
   *aligned_size = (int32_t) aligned;
   // integerMultiplyOverflow(alignedSize, segmentSize) -> std::overflow_error
   if( (int64_t) aligned * (int64_t) w > (int64_t) INT32_MAX || (int64_t) aligned * (int64_t) w < (int64_t) INT32_MIN )
      return 3;
   return 0;
}

void orc_gidx_ellpack( const int32_t* sizes, int64_t n, int32_t width, int32_t aligned, int col_major, int32_t* gidx )
{
   // EllpackBase.hpp:108-117
   int64_t p = 0;
   for( int64_t r = 0; r < n; r++ )
      for( int32_t k = 0; k < sizes[ r ]; k++ )
         gidx[ p++ ] = col_major ? (int32_t) ( r + (int64_t) aligned * k ) : (int32_t) ( r * width + k );
}

/* -------------------------------------------------------- SlicedEllpack */
void orc_sliced_ellpack( const int32_t* sizes, int64_t n, int slice_size, int32_t* slice_offsets,
                         int32_t* slice_seg_sizes )
{
   // SlicedEllpack.hpp:116-152: per-slice max of sizes (0 for virtual rows),
   // sliceOffsets[i] = max*SliceSize, then exclusive scan.
   const int64_t slices = ( n + slice_size - 1 ) / slice_size;
   for( int64_t s = 0; s < slices; s++ ) {
      int32_t m = 0;
      for( int64_t r = s * slice_size; r < ( s + 1 ) * slice_size; r++ )
         m = std::max( m, r < n ? sizes[ r ] : 0 );
      slice_seg_sizes[ s ] = m;
      slice_offsets[ s ] = m * slice_size;
   }
   slice_offsets[ slices ] = 0;
   int32_t acc = 0;
   for( int64_t s = 0; s <= slices; s++ ) {
      const int32_t v = slice_offsets[ s ];
      slice_offsets[ s ] = acc;
      acc += v;
   }
}

void orc_gidx_sliced( const int32_t* sizes, int64_t n, int slice_size, const int32_t* slice_offsets,
                      const int32_t* slice_seg_sizes, int col_major, int32_t* gidx )
{
   // SlicedEllpackBase.hpp:119-146
   int64_t p = 0;
   for( int64_t r = 0; r < n; r++ ) {
      const int64_t s = r / slice_size, in = r % slice_size;
      for( int32_t k = 0; k < sizes[ r ]; k++ )
         gidx[ p++ ] = col_major ? (int32_t) ( slice_offsets[ s ] + in + (int64_t) slice_size * k )
                                 : (int32_t) ( slice_offsets[ s ] + in * slice_seg_sizes[ s ] + k );
   }
}

/* ------------------------------------------------------------- BiEllpack */
static void biell_groups( const int32_t* sizes, int64_t n, int warp, const int32_t* perm, int32_t* group_ptrs )
{
   // BiEllpack.hpp:304-340 (initGroupPointers) followed by the exclusive scan
   // of :186.  Each strip owns logw+1 slots; slot logw is always 0.
   const int logw = ilog2( warp );
   const int64_t strips = biell_virtual( n, warp ) / warp;
   const int64_t total = strips * ( logw + 1 ) + 1;
   std::fill( group_ptrs, group_ptrs + total, 0 );
   for( int64_t strip = 0; strip < strips; strip++ ) {
      const int64_t first = strip * warp;
      const int64_t gbegin = strip * ( logw + 1 );
      int empty = 0;
      if( strip == strips - 1 ) {
         const int64_t cnt = n - first;
         while( cnt <= ( int64_t( 1 ) << ( logw - 1 - empty ) ) - 1 )
            empty++;
      }
      int32_t allocated = 0;
      for( int g = empty; g < logw; g++ ) {
         const int64_t pos = ( int64_t( 1 ) << ( logw - 1 - g ) ) - 1;
         int64_t ps = 0;
         while( perm[ ps + first ] != pos + first )
            ps++;
         const int32_t width = sizes[ ps + first ] - allocated;
         const int32_t height = 1 << ( logw - g );
         allocated = sizes[ ps + first ];
         group_ptrs[ g + gbegin ] = width * height;
      }
   }
   int32_t acc = 0;
   for( int64_t i = 0; i < total; i++ ) {
      const int32_t v = group_ptrs[ i ];
      group_ptrs[ i ] = acc;
      acc += v;
   }
}

void orc_biellpack( const int32_t* sizes, int64_t n, int warp, int32_t* perm, int32_t* group_ptrs )
{
   // BiEllpack.hpp:245-298: the strict-'<' cocktail sort of each strip is a
   // stable sort by size, descending; perm[row] = absolute sorted position.
   const int64_t strips = biell_virtual( n, warp ) / warp;
   std::vector< int64_t > idx;
   for( int64_t s = 0; s < strips; s++ ) {
      const int64_t b = s * warp, e = std::min< int64_t >( n, b + warp );
      idx.resize( e - b );
      std::iota( idx.begin(), idx.end(), b );
      std::stable_sort( idx.begin(), idx.end(), [ & ]( int64_t a, int64_t c ) { return sizes[ a ] > sizes[ c ]; } );
      for( int64_t p = 0; p < e - b; p++ )
         perm[ idx[ p ] ] = (int32_t) ( b + p );
   }
   biell_groups( sizes, n, warp, perm, group_ptrs );
}

void orc_biellpack_bruteforce( const int32_t* sizes, int64_t n, int warp, int32_t* perm, int32_t* group_ptrs )
{
   // Loop-for-loop restatement of BiEllpack.hpp:258-297 (O(warp^3) per strip);
   // exists only to pin orc_biellpack on small inputs.
   if( n == 0 ) {
      group_ptrs[ 0 ] = 0;
      return;
   }
   for( int64_t i = 0; i < n; i++ )
      perm[ i ] = (int32_t) i;
   const int64_t strips = biell_virtual( n, warp ) / warp;
   for( int64_t i = 0; i < strips; i++ ) {
      int64_t begin = i * warp;
      int64_t end = ( i + 1 ) * warp - 1;
      if( n - 1 < end )
         end = n - 1;
      bool sorted = false;
      int64_t p1 = 0, p2 = 0, offset = 0;
      while( ! sorted ) {
         sorted = true;
         for( int64_t j = begin + offset; j < end - offset; j++ ) {
            for( int64_t k = begin; k < end + 1; k++ ) {
               if( perm[ k ] == j )
                  p1 = k;
               if( perm[ k ] == j + 1 )
                  p2 = k;
            }
            if( sizes[ p1 ] < sizes[ p2 ] ) {
               std::swap( perm[ p1 ], perm[ p2 ] );
               sorted = false;
            }
         }
         for( int64_t j = end - 1 - offset; j > begin + offset; j-- ) {
            for( int64_t k = begin; k < end + 1; k++ ) {
               if( perm[ k ] == j )
                  p1 = k;
               if( perm[ k ] == j - 1 )
                  p2 = k;
            }
            if( sizes[ p2 ] < sizes[ p1 ] ) {
               std::swap( perm[ p1 ], perm[ p2 ] );
               sorted = false;
            }
         }
         offset++;
      }
   }
   biell_groups( sizes, n, warp, perm, group_ptrs );
}

void orc_gidx_biellpack( const int32_t* sizes, int64_t n, int warp, const int32_t* perm, const int32_t* group_ptrs,
                         int col_major, int32_t* gidx )
{
   // detail/BiEllpack.h getGlobalIndex
   const int logw = ilog2( warp );
   int64_t p = 0;
   for( int64_t r = 0; r < n; r++ ) {
      const int64_t strip = r / warp;
      const int64_t g0 = strip * ( logw + 1 );
      const I rsp = perm[ r ] - (I) ( strip * warp );
      const int groups = biell_active_groups( rsp, logw );
      for( int32_t k0 = 0; k0 < sizes[ r ]; k0++ ) {
         I local = k0;
         I global = group_ptrs[ g0 ];
         I height = warp;
         I out = -1;
         for( int g = 0; g < groups; g++ ) {
            const I gsize = group_ptrs[ g0 + g + 1 ] - group_ptrs[ g0 + g ];
            if( gsize ) {
               const I gwidth = gsize / height;
               if( local >= gwidth ) {
                  local -= gwidth;
                  global += gsize;
               }
               else {
                  out = col_major ? global + rsp + local * height : global + rsp * gwidth + local;
                  break;
               }
            }
            height /= 2;
         }
         gidx[ p++ ] = out;
      }
   }
}

/* -------------------------------------------------------- ChunkedEllpack */
int64_t orc_chunked_ellpack( const int32_t* sizes, int64_t n, int chunks_in_slice, int desired_chunk_size,
                             int32_t* slices, int32_t* seg_to_chunk, int32_t* seg_to_slice, int32_t* seg_ptrs,
                             int32_t* chunks_to_seg, int64_t chunks_cap, int32_t* storage )
{
   auto S = [ & ]( int64_t s, int f ) -> int32_t& { return slices[ 4 * s + f ]; };  // size, chunkSize, first, pointer
   // resolveSliceSizes, ChunkedEllpack.hpp:281-310
   const int32_t desired = chunks_in_slice * desired_chunk_size;
   int64_t nslices = 0;
   {
      int64_t seg = 0;
      int32_t slice_size = 0, allocated = 0;
      while( seg < n ) {
         allocated += sizes[ seg ];
         slice_size++;
         seg++;
         if( allocated < desired )
            if( seg < n && slice_size < chunks_in_slice )
               continue;
         S( nslices, 0 ) = slice_size;
         S( nslices, 2 ) = (int32_t) ( seg - slice_size );
         S( nslices, 3 ) = allocated;
         slice_size = 0;
         nslices++;
         allocated = 0;
      }
   }
   // setSlice, ChunkedEllpack.hpp:313-380
   seg_ptrs[ 0 ] = 0;
   int32_t elements = 0;
   for( int64_t si = 0; si < nslices; si++ ) {
      const int32_t slice_size = S( si, 0 ), begin = S( si, 2 ), alloc = S( si, 3 );
      const int32_t end = begin + slice_size;
      int32_t free_chunks = chunks_in_slice - slice_size;
      for( int32_t i = begin; i < end; i++ )
         seg_to_chunk[ i ] = 1;
      int total_added = 0;
      int max_len = sizes[ begin ];
      for( int32_t i = begin; i < end; i++ ) {
         double ratio = 0.0;
         if( alloc != 0 )
            ratio = (double) sizes[ i ] / (double) alloc;
         const int32_t added = (int32_t) ( free_chunks * ratio );  // Index addedChunks = freeChunks * segmentRatio
         total_added += added;
         seg_to_chunk[ i ] += added;
         if( max_len < sizes[ i ] )
            max_len = sizes[ i ];
      }
      free_chunks -= total_added;
      while( free_chunks )
         for( int32_t i = begin; i < end && free_chunks; i++ )
            if( sizes[ i ] == max_len ) {
               seg_to_chunk[ i ]++;
               free_chunks--;
            }
      int32_t max_chunk = 0;
      for( int32_t i = begin; i < end; i++ )
         max_chunk = std::max( max_chunk, ( sizes[ i ] + seg_to_chunk[ i ] - 1 ) / seg_to_chunk[ i ] );
      S( si, 1 ) = max_chunk;
      S( si, 3 ) = elements;
      elements += chunks_in_slice * max_chunk;
      for( int32_t i = begin; i < end; i++ )
         seg_to_slice[ i ] = (int32_t) si;
      for( int32_t i = begin; i < end; i++ )
         seg_ptrs[ i + 1 ] = max_chunk * seg_to_chunk[ i ];
      for( int32_t j = begin + 1; j < end; j++ )
         seg_to_chunk[ j ] += seg_to_chunk[ j - 1 ];
   }
   // inplaceInclusiveScan(segmentPointers), :170
   for( int64_t i = 1; i <= n; i++ )
      seg_ptrs[ i ] += seg_ptrs[ i - 1 ];
   // chunksToSegmentsMapping, :172-184
   const int64_t nchunks = nslices * chunks_in_slice;
   if( nchunks > chunks_cap )
      return -1;
   int64_t c = 0;
   for( int64_t seg = 0; seg < n; seg++ ) {
      const int32_t si = seg_to_slice[ seg ];
      int32_t first = 0;
      if( seg != S( si, 2 ) )
         first = seg_to_chunk[ seg - 1 ];
      const int32_t cnt = seg_to_chunk[ seg ] - first;
      for( int32_t i = 0; i < cnt; i++ )
         chunks_to_seg[ c++ ] = (int32_t) seg;
   }
   // entries past c keep whatever the caller put there (the reference leaves
   // them uninitialised)
   *storage = elements;
   return nslices;
}

void orc_gidx_chunked( const int32_t* sizes, int64_t n, int chunks_in_slice, const int32_t* slices,
                       const int32_t* seg_to_chunk, const int32_t* seg_to_slice, int col_major, int32_t* gidx )
{
   // detail/ChunkedEllpack.h getGlobalIndex
   int64_t p = 0;
   for( int64_t r = 0; r < n; r++ ) {
      const int32_t si = seg_to_slice[ r ];
      int32_t first = 0;
      if( r != slices[ 4 * si + 2 ] )
         first = seg_to_chunk[ r - 1 ];
      const int32_t off = slices[ 4 * si + 3 ], cs = slices[ 4 * si + 1 ];
      for( int32_t k = 0; k < sizes[ r ]; k++ )
         gidx[ p++ ] = col_major ? off + ( k % cs ) * chunks_in_slice + first + k / cs : off + first * cs + k;
   }
}

/* ---------------------------------------------------------- AdaptiveCSR */
int64_t orc_adaptive_csr_blocks( const int32_t* offsets, int64_t n, int value_size, int host_warp,
                                 int device_warp, int32_t* blocks, int64_t cap )
{
   // AdaptiveCSR.hpp:102-126: no blocks when max(offsets) == 0
   if( n == 0 || offsets[ n ] == 0 )
      return 0;
   // CSRAdaptiveKernelParameters.h: 24576 B of shared memory, block 256,
   // warps = 256 / getWarpSize() (host pass), MaxAdaptiveElementsPerWarp 512.
   const size_t per_warp = 24576 / (size_t) value_size / (size_t) ( 256 / host_warp );
   const int64_t rows = n + 1;  // offsets.getSize()
   int64_t count = 0;
   auto emit = [ & ]( int32_t type, int32_t first, int32_t bsize, int32_t nseg, int32_t widx, int32_t wcnt ) {
      if( count < cap ) {
         int32_t* d = blocks + 6 * count;
         d[ 0 ] = type;
         d[ 1 ] = first;
         d[ 2 ] = bsize;
         d[ 3 ] = nseg;
         d[ 4 ] = widx;
         d[ 5 ] = wcnt;
      }
      count++;
   };
   int64_t start = 0, next = 0;
   while( next != rows - 1 ) {
      // findLimit, AdaptiveCSR.hpp:165-192
      int type = 1;
      size_t sum = 0;
      int64_t cur = start;
      bool found = false;
      for( ; cur < rows - 1; cur++ ) {
         sum += offsets[ cur + 1 ] - offsets[ cur ];
         if( sum > per_warp ) {
            if( cur - start > 0 ) {
               type = 1;  // STREAM
               next = cur;
            }
            else {
               type = sum <= 2 * 512 ? 2 : 0;  // VECTOR : LONG
               next = cur + 1;
            }
            found = true;
            break;
         }
      }
      if( ! found ) {
         type = 1;
         next = rows - 1;
      }
      if( type == 0 ) {
         const int64_t wpb = 256 / device_warp;
         int64_t left = ( count + wpb - 1 ) / wpb * wpb - count;
         if( left == 0 )
            left = wpb;
         for( int64_t i = 0; i < left; i++ )
            emit( 0, (int32_t) start, 0, 0, (int32_t) i, (int32_t) left );
      }
      else
         emit( type, (int32_t) start, offsets[ next ] - offsets[ start ], (int32_t) ( next - start ), 0, 0 );
      start = next;
   }
   emit( 2, (int32_t) next, 0, 0, 0, 0 );  // terminal descriptor (VECTOR by default argument)
   return count <= cap ? count : -1;
}

/* -------------------------------------------------------- SortedSegments */
void orc_sorted_segments( const int32_t* sizes, int64_t n, int64_t sigma, int32_t* sorted_sizes,
                          int32_t* inverse_perm, int32_t* perm )
{
   // SortedSegments.hpp:129-200
   std::vector< std::pair< int32_t, int32_t > > aux( n );
   for( int64_t i = 0; i < n; i++ )
      aux[ i ] = { sizes[ i ], (int32_t) i };
   auto cmp = []( const std::pair< int32_t, int32_t >& a, const std::pair< int32_t, int32_t >& b ) {
      return a.first > b.first || ( a.first == b.first && a.second < b.second );
   };
   if( sigma == -1 )
      std::sort( aux.begin(), aux.end(), cmp );
   if( sigma > 0 )
      for( int64_t i = 0; i < n; i += sigma )
         std::sort( aux.begin() + i, aux.begin() + std::min( i + sigma, n ), cmp );
   for( int64_t i = 0; i < n; i++ ) {
      sorted_sizes[ i ] = aux[ i ].first;
      inverse_perm[ i ] = aux[ i ].second;
   }
   for( int64_t i = 0; i < n; i++ )
      perm[ inverse_perm[ i ] ] = (int32_t) i;
}

}  // extern "C"

/* ----------------------------------------------------------------- SpMV */
namespace {

template< typename V, typename X, typename C >
C row_sum( const int32_t* row_ptr, const int32_t* cols, const V* vals, const X* x, int64_t r )
{
   // ReducingOperations_CSR.h:46-60 with the fetch of SparseMatrixBase.hpp:331-345
   C aux = 0;
   for( int32_t g = row_ptr[ r ]; g < row_ptr[ r + 1 ]; g++ ) {
      const int32_t c = cols[ g ];
      const C v = c == -1 ? C( 0 ) : C( vals[ g ] * x[ c ] );
      aux = aux + v;
   }
   return aux;
}

template< typename V, typename X, typename Y, typename C >
void spmv( int64_t rows, const int32_t* row_ptr, const int32_t* cols, const V* vals, const X* x, Y* y, C alpha, C beta,
           int64_t begin, int64_t end )
{
   if( end == 0 )
      end = rows;
   for( int64_t r = begin; r < end; r++ ) {
      const C v = row_sum< V, X, C >( row_ptr, cols, vals, x, r );
      // SparseMatrixBase.hpp:346-389
      if( beta == C( 0 ) )
         y[ r ] = alpha == C( 1 ) ? Y( v ) : Y( alpha * v );
      else
         y[ r ] = alpha == C( 1 ) ? Y( beta * y[ r ] + v ) : Y( beta * y[ r ] + alpha * v );
   }
}

}  // namespace

extern "C" {

void orc_spmv_f64( int64_t rows, const int32_t* row_ptr, const int32_t* cols, const double* vals, const double* x,
                   double* y, double alpha, double beta, int64_t begin, int64_t end )
{
   spmv< double, double, double, double >( rows, row_ptr, cols, vals, x, y, alpha, beta, begin, end );
}

void orc_spmv_f32( int64_t rows, const int32_t* row_ptr, const int32_t* cols, const float* vals, const float* x,
                   float* y, float alpha, float beta, int64_t begin, int64_t end )
{
   spmv< float, float, float, float >( rows, row_ptr, cols, vals, x, y, alpha, beta, begin, end );
}

void orc_spmv_f64_omp( int64_t rows, const int32_t* row_ptr, const int32_t* cols, const double* vals, const double* x,
                       double* y, int threads )
{
   // ReducingOperations_CSR.h:69-73: omp parallel for schedule(dynamic, 100)
#pragma omp parallel for schedule( dynamic, 100 ) num_threads( threads )
   for( int64_t r = 0; r < rows; r++ )
      y[ r ] = row_sum< double, double, double >( row_ptr, cols, vals, x, r );
}

// SparseMatrixBase.hpp:298-329, a SymmetricMatrix whose CSR arrays hold the
// lower triangle and the diagonal: y *= beta (whole vector, beta != 1), then
// per row r of [begin, end), in element order, the mirrored term
// alpha * v * x[r] into y[c] for c < r and v * x[c] into the row's sum, which
// adds alpha * sum to y[r] (the host order of the reference with one thread).
void orc_spmv_symmetric( int64_t rows, const int32_t* row_ptr, const int32_t* cols, const double* vals,
                         const double* x, double* y, double alpha, double beta, int64_t begin, int64_t end )
{
   if( end == 0 )
      end = rows;
   if( beta != 1.0 )
      for( int64_t i = 0; i < rows; i++ )
         y[ i ] *= beta;
   for( int64_t r = begin; r < end; r++ ) {
      double sum = 0;
      for( int64_t g = row_ptr[ r ]; g < row_ptr[ r + 1 ]; g++ ) {
         const int32_t c = cols[ g ];
         if( c == -1 ) {
            sum += 0.0;  // the padding fetch returns 0
            continue;
         }
         if( c < r )
            y[ c ] += alpha * vals[ g ] * x[ r ];
         sum += vals[ g ] * x[ c ];
      }
      y[ r ] += alpha * sum;
   }
}

// SparseMatrixBase.hpp:541-588 (general matrix): y *= beta (whole vector,
// beta != 1), then y[c] += alpha * x[r] * v for every element with c in
// [begin, end), rows in order.
// SparseMatrixBase.hpp:166-231: the row's slots in order; a stored column is
// updated (multiplicator * old + value, two roundings), else the first padding
// slot or the first larger column takes the element -- for the latter the
// slots from it to the row end move one place right first, the last slot's
// content falling off (the reference does not check for a full row there)
int64_t orc_add_elements( int64_t n, const int32_t* off, int32_t* cols, double* vals, int symmetric,
                          const int64_t* op_pos, const double* op_val, int64_t nops )
{
   (void) n;
   for( int64_t i = 0; i < nops; i++ ) {
      int64_t r = op_pos[ 2 * i ], c = op_pos[ 2 * i + 1 ];
      const double value = op_val[ 2 * i ], mult = op_val[ 2 * i + 1 ];
      if( symmetric && r < c )
         std::swap( r, c );
      const int64_t b = off[ r ], size = off[ r + 1 ] - b;
      int64_t k = 0;
      bool done = false;
      for( ; k < size; k++ ) {
         const int32_t col = cols[ b + k ];
         if( col == c ) {
            const double scaled = mult * vals[ b + k ];
            vals[ b + k ] = scaled + value;
            done = true;
            break;
         }
         if( col == -1 || col > c )
            break;
      }
      if( done )
         continue;
      if( k == size )
         return i;
      if( cols[ b + k ] != -1 )
         for( int64_t j = size - 1; j > k; j-- ) {
            cols[ b + j ] = cols[ b + j - 1 ];
            vals[ b + j ] = vals[ b + j - 1 ];
         }
      cols[ b + k ] = (int32_t) c;
      vals[ b + k ] = value;
   }
   return -1;
}

void orc_spmv_transposed( int64_t rows, int64_t columns, const int32_t* row_ptr, const int32_t* cols,
                          const double* vals, const double* x, double* y, double alpha, double beta, int64_t begin,
                          int64_t end )
{
   if( end == 0 )
      end = columns;
   if( beta != 1.0 )
      for( int64_t i = 0; i < columns; i++ )
         y[ i ] *= beta;
   for( int64_t r = 0; r < rows; r++ )
      for( int64_t g = row_ptr[ r ]; g < row_ptr[ r + 1 ]; g++ ) {
         const int32_t c = cols[ g ];
         if( c >= begin && c < end && c != -1 )
            y[ c ] += alpha * x[ r ] * vals[ g ];
      }
}

}  // extern "C"

/* -------------------------------------------------------------- solvers */
namespace {

struct Csr
{
   int64_t n;
   const int32_t* ptr;
   const int32_t* col;
   const double* val;
   void mv( const double* x, double* y ) const
   {
      spmv< double, double, double, double >( n, ptr, col, val, x, y, 1.0, 0.0, 0, 0 );
   }
   double get( int64_t r, int64_t c ) const
   {
      for( int32_t g = ptr[ r ]; g < ptr[ r + 1 ]; g++ )
         if( col[ g ] == c )
            return val[ g ];
      return 0.0;
   }
};

double dot( const double* a, const double* b, int64_t n )
{
   double s = 0;
   for( int64_t i = 0; i < n; i++ )
      s += a[ i ] * b[ i ];
   return s;
}
double nrm2( const double* a, int64_t n )
{
   return std::sqrt( dot( a, a, n ) );
}

struct Ilu0
{
   std::vector< int32_t > lp, lc, up, uc;
   std::vector< double > lv, uv;
};

void ilu0_factor( const Csr& A, Ilu0& f )
{
   // Preconditioners/ILU0.hpp:12-96 (IKJ variant, minColumn = 0).  U rows are
   // stored reversed: U row N-1-i holds row i, diagonal entry first.
   const int64_t N = A.n;
   std::vector< int32_t > Lr( N ), Ur( N );
   for( int64_t i = 0; i < N; i++ ) {
      int32_t le = 0, ue = 0;
      for( int32_t g = A.ptr[ i ]; g < A.ptr[ i + 1 ]; g++ ) {
         const int32_t c = A.col[ g ];
         if( c < 0 )
            continue;  // padding never occurs in a compressed row
         if( c < i )
            le++;
         else if( c < N )
            ue++;
         else
            break;
      }
      Lr[ i ] = le;
      Ur[ N - 1 - i ] = ue;
   }
   f.lp.assign( N + 1, 0 );
   f.up.assign( N + 1, 0 );
   for( int64_t i = 0; i < N; i++ ) {
      f.lp[ i + 1 ] = f.lp[ i ] + Lr[ i ];
      f.up[ i + 1 ] = f.up[ i ] + Ur[ i ];
   }
   f.lc.assign( f.lp[ N ], -1 );
   f.lv.assign( f.lp[ N ], 0.0 );
   f.uc.assign( f.up[ N ], -1 );
   f.uv.assign( f.up[ N ], 0.0 );
   auto Uget = [ & ]( int64_t urow, int64_t c ) -> double {
      for( int32_t g = f.up[ urow ]; g < f.up[ urow + 1 ]; g++ )
         if( f.uc[ g ] == c )
            return f.uv[ g ];
      return 0.0;
   };
   for( int64_t i = 0; i < N; i++ ) {
      const int32_t* columns = A.col + A.ptr[ i ];
      const double* values = A.val + A.ptr[ i ];
      const int32_t le = Lr[ i ], ue = Ur[ N - 1 - i ];
      const int64_t ub = f.up[ N - 1 - i ];
      for( int32_t j = 0; j < ue; j++ ) {
         f.uc[ ub + j ] = columns[ le + j ];
         f.uv[ ub + j ] = values[ le + j ];
      }
      if( le > 0 ) {
         const int64_t lb = f.lp[ i ];
         for( int32_t j = 0; j < le; j++ ) {
            f.lc[ lb + j ] = columns[ j ];
            f.lv[ lb + j ] = values[ j ];
         }
         for( int32_t ck = 0; ck < le; ck++ ) {
            const int32_t k = f.lc[ lb + ck ];
            const double Lik = f.lv[ lb + ck ] / Uget( N - 1 - k, k );
            f.lv[ lb + ck ] = Lik;
            for( int32_t cj = ck + 1; cj < le; cj++ )
               f.lv[ lb + cj ] = f.lv[ lb + cj ] - Lik * Uget( N - 1 - k, f.lc[ lb + cj ] );
            for( int32_t cj = 0; cj < ue; cj++ )
               f.uv[ ub + cj ] = f.uv[ ub + cj ] - Lik * Uget( N - 1 - k, f.uc[ ub + cj ] );
         }
      }
   }
}

void ilu0_apply( int64_t N, const int32_t* lp, const int32_t* lc, const double* lv, const int32_t* up,
                 const int32_t* uc, const double* uv, const double* b, double* x )
{
   // Utils/TriangularSolve.h: triangularSolveLower<true>, then
   // triangularSolveUpper<reversedRows=true, fullStorage=true> in place.
   for( int64_t i = 0; i < N; i++ ) {
      double xi = b[ i ];
      for( int32_t g = lp[ i ]; g < lp[ i + 1 ]; g++ )
         xi -= lv[ g ] * x[ lc[ g ] ];
      x[ i ] = xi;
   }
   for( int64_t k = 0; k < N; k++ ) {
      const int64_t i = N - 1 - k;
      double xi = x[ i ];
      const double uii = uv[ up[ k ] ];
      for( int32_t g = up[ k ] + 1; g < up[ k + 1 ]; g++ )
         xi -= uv[ g ] * x[ uc[ g ] ];
      x[ i ] = xi / uii;
   }
}

struct Precond
{
   int kind = 0;
   std::vector< double > diag;
   Ilu0 ilu;
   void update( const Csr& A )
   {
      if( kind == 1 ) {
         diag.resize( A.n );
         for( int64_t i = 0; i < A.n; i++ )
            diag[ i ] = A.get( i, i );  // Diagonal.hpp:27-30
      }
      else if( kind == 2 )
         ilu0_factor( A, ilu );
   }
   void solve( const double* b, double* x, int64_t n ) const
   {
      if( kind == 1 )
         for( int64_t i = 0; i < n; i++ )
            x[ i ] = b[ i ] / diag[ i ];  // Diagonal.hpp:38-43
      else if( kind == 2 )
         ilu0_apply( n, ilu.lp.data(), ilu.lc.data(), ilu.lv.data(), ilu.up.data(), ilu.uc.data(), ilu.uv.data(),
                     b, x );
      else
         std::memcpy( x, b, sizeof( double ) * n );
   }
};

// IterativeSolver.hpp:98-143
struct Iter
{
   orc_solver_state* st;
   int64_t it = 0;
   double res = 0;
   bool check() const
   {
      return ! ( std::isnan( res ) || it > st->max_iterations
                 || ( res > st->divergence_residue && it >= st->min_iterations )
                 || ( res < st->convergence_residue && it >= st->min_iterations ) );
   }
   bool next()
   {
      bool r = check();
      it++;
      return r;
   }
   bool converged() const
   {
      if( std::isnan( res ) )
         return false;
      if( res > st->divergence_residue && it > st->min_iterations )
         return false;
      if( it >= st->max_iterations )
         return false;
      if( res > st->convergence_residue )
         return false;
      return true;
   }
   void finish( bool conv )
   {
      st->iterations = it;
      st->residue = res;
      st->converged = conv;
   }
};

}  // namespace

extern "C" {

void orc_cg( int64_t n, const int32_t* row_ptr, const int32_t* cols, const double* vals, const double* b, double* x,
             int precond, orc_solver_state* st )
{
   // CG.hpp:10-108
   Csr A{ n, row_ptr, cols, vals };
   Precond M;
   M.kind = precond;
   M.update( A );
   std::vector< double > r( n ), p( n ), Ap( n ), z( n );
   Iter S{ st };
   double normb;
   if( precond ) {
      M.solve( b, r.data(), n );
      normb = nrm2( r.data(), n );
   }
   else
      normb = nrm2( b, n );
   if( normb == 0 ) {
      std::fill( x, x + n, 0.0 );
      S.finish( true );
      return;
   }
   A.mv( x, r.data() );
   for( int64_t i = 0; i < n; i++ )
      r[ i ] = b[ i ] - r[ i ];
   double s1, s2;
   if( precond ) {
      M.solve( r.data(), z.data(), n );
      p = z;
      s1 = dot( r.data(), z.data(), n );
   }
   else {
      p = r;
      s1 = dot( r.data(), r.data(), n );
   }
   S.res = std::sqrt( s1 ) / normb;
   while( S.next() ) {
      A.mv( p.data(), Ap.data() );
      s2 = dot( Ap.data(), p.data(), n );
      if( s2 == 0.0 ) {
         S.res = 0.0;
         break;
      }
      const double alpha = s1 / s2;
      for( int64_t i = 0; i < n; i++ )
         x[ i ] += alpha * p[ i ];
      for( int64_t i = 0; i < n; i++ )
         r[ i ] -= alpha * Ap[ i ];
      if( precond ) {
         M.solve( r.data(), z.data(), n );
         s2 = s1;
         s1 = dot( r.data(), z.data(), n );
      }
      else {
         s2 = s1;
         s1 = dot( r.data(), r.data(), n );
      }
      const double beta = s2 == 0.0 ? 0.0 : s1 / s2;
      const double* src = precond ? z.data() : r.data();
      for( int64_t i = 0; i < n; i++ )
         p[ i ] = src[ i ] + beta * p[ i ];
      S.res = std::sqrt( s1 ) / normb;
   }
   S.finish( S.converged() );
}

void orc_jacobi( int64_t n, const int32_t* row_ptr, const int32_t* cols, const double* vals, const double* b,
                 double* x, double omega, int64_t residue_period, orc_solver_state* st )
{
   // Jacobi.hpp:65-125
   Csr A{ n, row_ptr, cols, vals };
   std::vector< double > d( n, 0.0 ), aux( n ), tmp( n );
   for( int64_t i = 0; i < n; i++ )
      for( int32_t g = row_ptr[ i ]; g < row_ptr[ i + 1 ]; g++ )
         if( cols[ g ] == i )
            d[ i ] = vals[ g ];
   Iter S{ st };
   S.res = st->convergence_residue + 1.0;
   const double bnorm = nrm2( b, n );
   auto residue = [ & ]() {
      A.mv( x, tmp.data() );
      double s = 0;
      for( int64_t i = 0; i < n; i++ )
         s += ( tmp[ i ] - b[ i ] ) * ( tmp[ i ] - b[ i ] );
      return std::sqrt( s ) / bnorm;
   };
   auto step = [ & ]( const double* in, double* out ) {
      for( int64_t r = 0; r < n; r++ ) {
         double v = 0;
         for( int32_t g = row_ptr[ r ]; g < row_ptr[ r + 1 ]; g++ )
            if( cols[ g ] >= 0 )
               v = v + vals[ g ] * in[ cols[ g ] ];
         out[ r ] = in[ r ] + omega / d[ r ] * ( b[ r ] - v );
      }
   };
   std::copy( x, x + n, aux.begin() );
   while( S.next() ) {
      step( x, aux.data() );
      if( S.it % residue_period == 0 )
         S.res = residue();
      S.it++;
      step( aux.data(), x );
      if( S.it % residue_period == 0 )
         S.res = residue();
   }
   S.res = residue();
   S.finish( S.converged() );
}

void orc_ilu0_factor( int64_t n, const int32_t* row_ptr, const int32_t* cols, const double* vals, int32_t* l_ptr,
                      int32_t* l_cols, double* l_vals, int32_t* u_ptr, int32_t* u_cols, double* u_vals )
{
   Csr A{ n, row_ptr, cols, vals };
   Ilu0 f;
   ilu0_factor( A, f );
   std::copy( f.lp.begin(), f.lp.end(), l_ptr );
   std::copy( f.up.begin(), f.up.end(), u_ptr );
   std::copy( f.lc.begin(), f.lc.end(), l_cols );
   std::copy( f.lv.begin(), f.lv.end(), l_vals );
   std::copy( f.uc.begin(), f.uc.end(), u_cols );
   std::copy( f.uv.begin(), f.uv.end(), u_vals );
}

void orc_ilu0_solve( int64_t n, const int32_t* l_ptr, const int32_t* l_cols, const double* l_vals,
                     const int32_t* u_ptr, const int32_t* u_cols, const double* u_vals, const double* b, double* x )
{
   ilu0_apply( n, l_ptr, l_cols, l_vals, u_ptr, u_cols, u_vals, b, x );
}

double orc_true_residual( int64_t n, const int32_t* row_ptr, const int32_t* cols, const double* vals,
                          const double* x, const double* b )
{
   Csr A{ n, row_ptr, cols, vals };
   std::vector< double > v( n );
   A.mv( x, v.data() );
   double s = 0;
   for( int64_t i = 0; i < n; i++ )
      s += ( v[ i ] - b[ i ] ) * ( v[ i ] - b[ i ] );
   return std::sqrt( s ) / nrm2( b, n );
}

}  // extern "C"

/* ---------------------------------------------------------------- GMRES */
namespace {

struct Gmres
{
   // Solvers/Linear/GMRES.hpp and GMRES.h:122-232
   Csr A;
   Precond M;
   int variant, rmin, rmax, dmin, dmax;
   int64_t n, ld;
   Iter S;
   std::vector< double > r, w, tmp, z, V, Y, H, s, cs, sn, T, YL;

   void matvec_prec( double* out, const double* v )
   {
      // preconditioned_matvec, :451-465
      if( M.kind ) {
         A.mv( v, tmp.data() );
         M.solve( tmp.data(), out, n );
      }
      else
         A.mv( v, out );
   }
   void residue_vec( const double* x, const double* b )
   {
      // compute_residue, :433-449
      if( M.kind ) {
         A.mv( x, tmp.data() );
         for( int64_t i = 0; i < n; i++ )
            tmp[ i ] = b[ i ] - tmp[ i ];
         M.solve( tmp.data(), r.data(), n );
      }
      else {
         A.mv( x, r.data() );
         for( int64_t i = 0; i < n; i++ )
            r[ i ] = b[ i ] - r[ i ];
      }
   }
   static void gen_rot( double dx, double dy, double& c, double& sn_ )
   {
      // generatePlaneRotation, :729-746
      if( dy == 0.0 ) {
         c = 1.0;
         sn_ = 0.0;
      }
      else if( std::fabs( dy ) > std::fabs( dx ) ) {
         const double t = dx / dy;
         sn_ = 1.0 / std::sqrt( 1.0 + t * t );
         c = t * sn_;
      }
      else {
         const double t = dy / dx;
         c = 1.0 / std::sqrt( 1.0 + t * t );
         sn_ = t * c;
      }
   }
   static void app_rot( double& dx, double& dy, double c, double sn_ )
   {
      // applyPlaneRotation, :748-756
      const double t = c * dx + sn_ * dy;
      dy = c * dy - sn_ * dx;
      dx = t;
   }
   void givens( int i, int m )
   {
      // apply_givens_rotations, :758-770
      for( int k = 0; k < i; k++ )
         app_rot( H[ k + i * ( m + 1 ) ], H[ k + 1 + i * ( m + 1 ) ], cs[ k ], sn[ k ] );
      if( H[ i + 1 + i * ( m + 1 ) ] != 0.0 ) {
         gen_rot( H[ i + i * ( m + 1 ) ], H[ i + 1 + i * ( m + 1 ) ], cs[ i ], sn[ i ] );
         app_rot( H[ i + i * ( m + 1 ) ], H[ i + 1 + i * ( m + 1 ) ], cs[ i ], sn[ i ] );
         app_rot( s[ i ], s[ i + 1 ], cs[ i ], sn[ i ] );
      }
   }
   // gemv of Matrices/MatrixOperations.h: y = alpha*A(:,0..k-1)*x + beta*y
   void gemv( int k, double alpha, const double* Am, const double* xv, double beta, double* yv )
   {
      for( int64_t i = 0; i < n; i++ ) {
         double acc = 0;
         for( int j = 0; j < k; j++ )
            acc += Am[ i + j * ld ] * xv[ j ];
         yv[ i ] = beta == 0.0 ? alpha * acc : alpha * acc + beta * yv[ i ];
      }
   }
   int ortho_gs( int m, double normb, double beta, bool classical )
   {
      // orthogonalize_CGS :182-254, orthogonalize_MGS :256-343
      for( int64_t i = 0; i < n; i++ )
         V[ i ] = ( 1.0 / beta ) * r[ i ];
      std::fill( H.begin(), H.end(), 0.0 );
      std::fill( s.begin(), s.end(), 0.0 );
      s[ 0 ] = beta;
      const bool reorth = variant == 1 || variant == 3;
      for( int i = 0; i < m && S.next(); i++ ) {
         matvec_prec( w.data(), &V[ i * ld ] );
         for( int k = 0; k <= i; k++ )
            H[ k + i * ( m + 1 ) ] = 0.0;
         for( int l = 0; l < ( reorth ? 2 : 1 ); l++ ) {
            if( classical ) {
               std::vector< double > Hl( i + 1 );
               for( int k = 0; k <= i; k++ )
                  Hl[ k ] = dot( &V[ k * ld ], w.data(), n );
               for( int k = 0; k <= i; k++ )
                  H[ k + i * ( m + 1 ) ] += Hl[ k ];
               gemv( i + 1, -1.0, V.data(), Hl.data(), 1.0, w.data() );
            }
            else
               for( int k = 0; k <= i; k++ ) {
                  const double hk = dot( w.data(), &V[ k * ld ], n );
                  H[ k + i * ( m + 1 ) ] += hk;
                  for( int64_t q = 0; q < n; q++ )
                     w[ q ] = w[ q ] - hk * V[ k * ld + q ];
               }
         }
         const double normw = nrm2( w.data(), n );
         H[ i + 1 + i * ( m + 1 ) ] = normw;
         for( int64_t q = 0; q < n; q++ )
            V[ ( i + 1 ) * ld + q ] = ( 1.0 / normw ) * w[ q ];
         givens( i, m );
         S.res = std::fabs( s[ i + 1 ] ) / normb;
         if( ! S.check() )
            return i;
      }
      return m;
   }
   // Householder helpers for CWY, :467-643
   void hh_generate( int i, double* yi )
   {
      for( int64_t j = 0; j < n; j++ )
         yi[ j ] = j < i ? 0.0 : z[ j ];
      const double normz = nrm2( yi, n );
      const double yii = yi[ i ];
      yi[ i ] = yii > 0.0 ? yii + normz : yii - normz;
      const double nys = 2 * ( normz * normz + std::fabs( yii ) * normz );
      const double ti = 2.0 / nys;
      const int R = rmax + 1;
      T[ i + i * R ] = ti;
      if( i > 0 ) {
         std::vector< double > aux( i );
         for( int k = 0; k < i; k++ )
            aux[ k ] = dot( &Y[ k * ld ], yi, n );
         for( int k = 0; k < i; k++ ) {
            T[ k + i * R ] = 0.0;
            for( int j = k; j < i; j++ )
               T[ k + i * R ] -= T[ k + j * R ] * ( ti * aux[ j ] );
         }
      }
   }
   void hh_apply_trunc( double* out, int i, const double* yi )
   {
      const int R = rmax + 1;
      for( int k = 0; k < R && k < n; k++ )
         YL[ i * R + k ] = yi[ k ];
      for( int k = 0; k <= i; k++ )
         out[ k ] = z[ k ] - YL[ i * R + k ] * 1.0;
   }
   void hh_cwy( double* v, int i )
   {
      const int R = rmax + 1;
      std::vector< double > aux( i + 1 );
      for( int k = 0; k <= i; k++ )
         aux[ k ] = YL[ i + k * R ];
      for( int k = 0; k <= i; k++ ) {
         double a2 = 0.0;
         for( int j = k; j <= i; j++ )
            a2 += T[ k + j * R ] * aux[ j ];
         aux[ k ] = a2;
      }
      gemv( i + 1, -1.0, Y.data(), aux.data(), 0.0, v );
      v[ i ] = 1.0 + v[ i ];
   }
   void hh_cwy_t( int i )
   {
      const int R = rmax + 1;
      std::vector< double > aux( i + 1 );
      for( int k = 0; k <= i; k++ )
         aux[ k ] = dot( &Y[ k * ld ], w.data(), n );
      for( int k = i; k >= 0; k-- ) {
         double a2 = 0.0;
         for( int j = 0; j <= k; j++ )
            a2 += T[ j + k * R ] * aux[ j ];
         aux[ k ] = a2;
      }
      z = w;
      gemv( i + 1, -1.0, Y.data(), aux.data(), 1.0, z.data() );
   }
   int ortho_cwy( int m, double normb )
   {
      // orthogonalize_CWY, :345-431
      z = r;
      std::fill( H.begin(), H.end(), 0.0 );
      std::fill( s.begin(), s.end(), 0.0 );
      std::fill( T.begin(), T.end(), 0.0 );
      for( int i = 0; i <= m && S.next(); i++ ) {
         double* yi = &Y[ i * ld ];
         hh_generate( i, yi );
         if( i == 0 )
            hh_apply_trunc( s.data(), i, yi );
         else
            hh_apply_trunc( &H[ ( i - 1 ) * ( m + 1 ) ], i, yi );
         hh_cwy( V.data(), i );
         if( i < m ) {
            matvec_prec( w.data(), V.data() );
            hh_cwy_t( i );
         }
         if( i > 0 )
            givens( i - 1, m );
         S.res = std::fabs( s[ i ] ) / normb;
         if( i > 0 && ! S.check() )
            return i - 1;
      }
      return m;
   }
   void update( int k, int m, double* x )
   {
      // update, :645-727
      std::vector< double > y( s.begin(), s.begin() + m + 1 );
      for( int i = k; i >= 0; i-- ) {
         y[ i ] /= H[ i + i * ( m + 1 ) ];
         for( int j = i - 1; j >= 0; j-- )
            y[ j ] -= H[ j + i * ( m + 1 ) ] * y[ i ];
      }
      if( variant != 4 )
         gemv( k + 1, 1.0, V.data(), y.data(), 1.0, x );
      else {
         const int R = rmax + 1;
         std::vector< double > aux( k + 1, 0.0 );
         for( int j = 0; j <= k; j++ )
            for( int i = 0; i <= j; i++ )
               aux[ i ] += y[ j ] * YL[ j + i * R ];
         for( int i = 0; i <= k; i++ ) {
            double a2 = 0.0;
            for( int j = i; j <= k; j++ )
               a2 += T[ i + j * R ] * aux[ j ];
            aux[ i ] = a2;
         }
         gemv( k + 1, -1.0, Y.data(), aux.data(), 1.0, x );
         for( int j = 0; j <= k; j++ )
            x[ j ] = x[ j ] + y[ j ];
      }
   }
   bool solve( const double* b, double* x )
   {
      // solve, :71-180
      if( rmin <= 0 || rmax <= 0 || rmin > rmax || dmin < 0 || dmax < 0 || dmin > dmax )
         return false;
      ld = ( n + 31 ) / 32 * 32 + 1;  // setSize, :772-806 (host layout)
      const int m0 = rmax;
      r.assign( n, 0 );
      w.assign( n, 0 );
      tmp.assign( n, 0 );
      cs.assign( m0 + 1, 0 );
      sn.assign( m0 + 1, 0 );
      H.assign( ( m0 + 1 ) * m0, 0 );
      s.assign( m0 + 1, 0 );
      if( variant == 4 ) {
         z.assign( n, 0 );
         Y.assign( ld * ( m0 + 1 ), 0 );
         T.assign( ( m0 + 1 ) * ( m0 + 1 ), 0 );
         YL.assign( ( m0 + 1 ) * ( m0 + 1 ), 0 );
         V.assign( n, 0 );
      }
      else
         V.assign( ld * ( m0 + 1 ), 0 );
      double normb;
      if( M.kind ) {
         M.solve( b, tmp.data(), n );
         normb = nrm2( tmp.data(), n );
      }
      else
         normb = nrm2( b, n );
      if( normb == 0 ) {
         std::fill( x, x + n, 0.0 );
         return true;
      }
      residue_vec( x, b );
      double beta = nrm2( r.data(), n );
      S.it = 0;
      S.res = beta / normb;
      double beta_ratio = 1;
      int cycles = 0;
      int m = rmax;
      while( S.check() ) {
         if( rmax > rmin && cycles > 0 ) {
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
         int o = variant == 4 ? ortho_cwy( m, normb ) : ortho_gs( m, normb, beta, variant <= 1 );
         if( o < m ) {
            update( o, m, x );
            return S.converged();
         }
         update( m - 1, m, x );
         residue_vec( x, b );
         const double old = beta;
         beta = nrm2( r.data(), n );
         S.res = beta / normb;
         ++cycles;
         beta_ratio = beta / old;
      }
      return S.converged();
   }
};

}  // namespace

extern "C" void orc_gmres( int64_t n, const int32_t* row_ptr, const int32_t* cols, const double* vals,
                           const double* b, double* x, int precond, int variant, int restart_min, int restart_max,
                           int step_min, int step_max, orc_solver_state* st )
{
   Gmres g;
   g.A = Csr{ n, row_ptr, cols, vals };
   g.M.kind = precond;
   g.M.update( g.A );
   g.variant = variant;
   g.rmin = restart_min;
   g.rmax = restart_max;
   g.dmin = step_min;
   g.dmax = step_max;
   g.n = n;
   g.S.st = st;
   const bool conv = g.solve( b, x );
   g.S.finish( conv );
}

/* ------------------------------------------------------ BiCGStab / TFQMR */
namespace {

// M^{-1} A src (PreconditionedMatvec of BICGStab.hpp:147-157 / TFQMR.hpp:46-53)
void pmv( const Csr& A, const Precond& M, int precond, const double* src, double* dst, std::vector< double >& tmp )
{
   if( precond ) {
      A.mv( src, tmp.data() );
      M.solve( tmp.data(), dst, A.n );
   }
   else
      A.mv( src, dst );
}

// r = M^{-1} (b - A x) (BICGStab.hpp:131-145)
void pres( const Csr& A, const Precond& M, int precond, const double* x, const double* b, double* r,
           std::vector< double >& tmp )
{
   const int64_t n = A.n;
   A.mv( x, tmp.data() );
   for( int64_t i = 0; i < n; i++ )
      tmp[ i ] = b[ i ] - tmp[ i ];
   if( precond )
      M.solve( tmp.data(), r, n );
   else
      std::memcpy( r, tmp.data(), sizeof( double ) * n );
}

}  // namespace

extern "C" {

void orc_bicgstab( int64_t n, const int32_t* row_ptr, const int32_t* cols, const double* vals, const double* b,
                   double* x, int precond, int exact_residue, orc_solver_state* st )
{
   // BICGStab.hpp:33-127
   Csr A{ n, row_ptr, cols, vals };
   Precond M;
   M.kind = precond;
   M.update( A );
   std::vector< double > r( n ), r_ast( n ), p( n ), s( n ), Ap( n ), As( n ), tmp( n );
   Iter S{ st };
   double b_norm;
   if( precond ) {
      M.solve( b, tmp.data(), n );
      b_norm = nrm2( tmp.data(), n );
   }
   else
      b_norm = nrm2( b, n );
   if( b_norm == 0 ) {
      std::fill( x, x + n, 0.0 );
      S.finish( true );
      return;
   }
   pres( A, M, precond, x, b, r.data(), tmp );
   p = r_ast = r;
   std::fill( s.begin(), s.end(), 0.0 );
   double rho = dot( r.data(), r_ast.data(), n ), r_ast_sqnorm = rho;
   const double eps2 = std::numeric_limits< double >::epsilon() * std::numeric_limits< double >::epsilon();
   S.res = std::sqrt( rho ) / b_norm;
   while( S.next() ) {
      pmv( A, M, precond, p.data(), Ap.data(), tmp );
      const double alpha = rho / dot( Ap.data(), r_ast.data(), n );
      for( int64_t i = 0; i < n; i++ )
         s[ i ] = r[ i ] - alpha * Ap[ i ];
      pmv( A, M, precond, s.data(), As.data(), tmp );
      const double omega = dot( As.data(), s.data(), n ) / dot( As.data(), As.data(), n );
      for( int64_t i = 0; i < n; i++ )
         x[ i ] += alpha * p[ i ] + omega * s[ i ];
      for( int64_t i = 0; i < n; i++ )
         r[ i ] = s[ i ] - omega * As[ i ];
      const double rho_old = rho;
      rho = dot( r.data(), r_ast.data(), n );
      if( std::abs( rho ) < eps2 * r_ast_sqnorm ) {
         pres( A, M, precond, x, b, r.data(), tmp );
         r_ast = r;
         r_ast_sqnorm = rho = dot( r.data(), r_ast.data(), n );
      }
      const double beta = ( rho / rho_old ) * ( alpha / omega );
      for( int64_t i = 0; i < n; i++ )
         p[ i ] = r[ i ] + beta * p[ i ] - ( beta * omega ) * Ap[ i ];
      if( exact_residue ) {
         pres( A, M, precond, x, b, s.data(), tmp );
         S.res = nrm2( s.data(), n ) / b_norm;
      }
      else
         S.res = nrm2( r.data(), n ) / b_norm;
   }
   S.finish( S.converged() );
}

void orc_tfqmr( int64_t n, const int32_t* row_ptr, const int32_t* cols, const double* vals, const double* b,
                double* x, int precond, orc_solver_state* st )
{
   // TFQMR.hpp:12-121
   Csr A{ n, row_ptr, cols, vals };
   Precond M;
   M.kind = precond;
   M.update( A );
   std::vector< double > r( n ), r_ast( n ), w( n ), u( n ), v( n ), d( n ), Au( n ), tmp( n );
   Iter S{ st };
   double b_norm;
   if( precond ) {
      M.solve( b, tmp.data(), n );
      b_norm = nrm2( tmp.data(), n );
   }
   else
      b_norm = nrm2( b, n );
   pres( A, M, precond, x, b, r.data(), tmp );
   if( b_norm == 0 ) {
      std::fill( x, x + n, 0.0 );
      S.finish( true );
      return;
   }
   w = u = r;
   pmv( A, M, precond, u.data(), Au.data(), tmp );
   v = Au;
   std::fill( d.begin(), d.end(), 0.0 );
   double tau = nrm2( r.data(), n ), theta = 0.0, eta = 0.0, alpha = 0.0;
   r_ast = r;
   double rho = dot( r_ast.data(), r.data(), n );
   S.res = tau / b_norm;
   while( S.next() ) {
      const int64_t iter = S.it;
      if( iter % 2 == 1 )
         alpha = rho / dot( v.data(), r_ast.data(), n );
      else
         pmv( A, M, precond, u.data(), Au.data(), tmp );
      for( int64_t i = 0; i < n; i++ )
         w[ i ] -= alpha * Au[ i ];
      const double dc = theta * theta * eta / alpha;
      for( int64_t i = 0; i < n; i++ )
         d[ i ] = u[ i ] + dc * d[ i ];
      const double w_norm = nrm2( w.data(), n );
      theta = w_norm / tau;
      const double c = 1.0 / std::sqrt( 1.0 + theta * theta );
      tau = tau * theta * c;
      eta = c * c * alpha;
      for( int64_t i = 0; i < n; i++ )
         x[ i ] += eta * d[ i ];
      S.res = tau * std::sqrt( (double) ( iter + 1 ) ) / b_norm;
      if( iter > st->min_iterations && S.res < st->convergence_residue )
         break;
      if( iter % 2 == 0 ) {
         const double rho_new = dot( w.data(), r_ast.data(), n );
         const double beta = rho_new / rho;
         rho = rho_new;
         for( int64_t i = 0; i < n; i++ ) {
            u[ i ] = w[ i ] + beta * u[ i ];
            v[ i ] = beta * Au[ i ] + ( beta * beta ) * v[ i ];
         }
         pmv( A, M, precond, u.data(), Au.data(), tmp );
         for( int64_t i = 0; i < n; i++ )
            v[ i ] += Au[ i ];
      }
      else
         for( int64_t i = 0; i < n; i++ )
            u[ i ] -= alpha * v[ i ];
   }
   S.finish( S.converged() );
}


// BICGStabL.hpp:37-257: BiCGStab(ell) -- ell Bi-CG steps on the columns of R
// and U (column j is R_j, U_j; n entries each), then a minimal-residual step
// over r_1..r_ell orthogonalised by modified Gram-Schmidt.  One iteration is
// counted per Bi-CG step.
void orc_bicgstabl( int64_t n, const int32_t* row_ptr, const int32_t* cols, const double* vals, const double* b,
                    double* x, int precond, int ell, int exact_residue, orc_solver_state* st )
{
   Csr A{ n, row_ptr, cols, vals };
   Precond M;
   M.kind = precond;
   M.update( A );
   Iter S{ st };
   std::vector< double > R( ( ell + 1 ) * n ), U( ( ell + 1 ) * n ), r_ast( n ), tmp( n ), res( n );
   std::vector< double > T( ell * ell ), sigma( ell + 1 ), g0( ell + 1 ), g1( ell + 1 ), g2( ell + 1 );
   auto col = [ & ]( std::vector< double >& V, int j ) { return V.data() + (int64_t) j * n; };
   double b_norm;
   if( precond ) {
      M.solve( b, tmp.data(), n );
      b_norm = nrm2( tmp.data(), n );
   }
   else
      b_norm = nrm2( b, n );
   if( b_norm == 0 ) {
      std::fill( x, x + n, 0.0 );
      S.finish( true );
      return;
   }
   pres( A, M, precond, x, b, col( R, 0 ), tmp );
   sigma[ 0 ] = nrm2( col( R, 0 ), n );
   if( std::isnan( sigma[ 0 ] ) ) {
      S.res = sigma[ 0 ];
      S.finish( false );  // the reference throws "initial residue is NAN"
      return;
   }
   for( int64_t i = 0; i < n; i++ )
      r_ast[ i ] = col( R, 0 )[ i ] / sigma[ 0 ];
   double rho0 = 1.0, alpha = 0.0, omega = 1.0;
   std::fill( U.begin(), U.begin() + n, 0.0 );
   S.res = sigma[ 0 ] / b_norm;
   while( S.check() ) {
      rho0 = -omega * rho0;
      for( int j = 0; j < ell; j++ ) {
         S.next();
         const double rho1 = dot( r_ast.data(), col( R, j ), n );
         const double beta = alpha * rho1 / rho0;
         rho0 = rho1;
         for( int c = 0; c <= j; c++ )  // U_[0:j] = R_[0:j] - beta U_[0:j]
            for( int64_t i = 0; i < n; i++ )
               col( U, c )[ i ] = 1.0 * col( R, c )[ i ] + ( -beta ) * col( U, c )[ i ];
         pmv( A, M, precond, col( U, j ), col( U, j + 1 ), tmp );
         const double gamma = dot( r_ast.data(), col( U, j + 1 ), n );
         alpha = rho0 / gamma;
         for( int c = 0; c <= j; c++ )  // R_[0:j] = R_[0:j] - alpha U_[1:j+1]
            for( int64_t i = 0; i < n; i++ )
               col( R, c )[ i ] = 1.0 * col( R, c )[ i ] + ( -alpha ) * col( U, c + 1 )[ i ];
         pmv( A, M, precond, col( R, j ), col( R, j + 1 ), tmp );
         for( int64_t i = 0; i < n; i++ )
            x[ i ] = x[ i ] + alpha * col( U, 0 )[ i ];
      }
      for( int j = 1; j <= ell; j++ ) {
         for( int i = 1; i < j; i++ ) {
            const int ij = ( i - 1 ) + ( j - 1 ) * ell;
            T[ ij ] = dot( col( R, i ), col( R, j ), n ) / sigma[ i ];
            for( int64_t q = 0; q < n; q++ )
               col( R, j )[ q ] = col( R, j )[ q ] - T[ ij ] * col( R, i )[ q ];
         }
         sigma[ j ] = dot( col( R, j ), col( R, j ), n );
         g1[ j ] = dot( col( R, 0 ), col( R, j ), n ) / sigma[ j ];
      }
      omega = g1[ ell ];
      for( int j = ell; j >= 1; j-- ) {  // g0 = T^-1 g1
         g0[ j ] = g1[ j ];
         for( int i = j + 1; i <= ell; i++ )
            g0[ j ] -= T[ ( j - 1 ) + ( i - 1 ) * ell ] * g0[ i ];
      }
      for( int j = 1; j < ell; j++ ) {  // g2 = T S g0
         g2[ j ] = g0[ j + 1 ];
         for( int i = j + 1; i < ell; i++ )
            g2[ j ] += T[ ( j - 1 ) + ( i - 1 ) * ell ] * g0[ i + 1 ];
      }
      g2[ 0 ] = g0[ 1 ];
      // x += R_[0:ell-1] g2; r_0 -= R_[1:ell] g1[1:ell]; u_0 -= U_[1:ell] g0[1:ell]
      // (MatrixOperations::gemv: sum over the columns in order, then + y)
      for( int64_t i = 0; i < n; i++ ) {
         double ax = 0, ar = 0, au = 0;
         for( int c = 0; c < ell; c++ ) {
            ax += col( R, c )[ i ] * ( 1.0 * g2[ c ] );
            ar += col( R, c + 1 )[ i ] * ( -1.0 * g1[ c + 1 ] );
            au += col( U, c + 1 )[ i ] * ( -1.0 * g0[ c + 1 ] );
         }
         x[ i ] = ax + 1.0 * x[ i ];
         col( R, 0 )[ i ] = ar + 1.0 * col( R, 0 )[ i ];
         col( U, 0 )[ i ] = au + 1.0 * col( U, 0 )[ i ];
      }
      if( exact_residue ) {
         pres( A, M, precond, x, b, res.data(), tmp );
         sigma[ 0 ] = nrm2( res.data(), n );
      }
      else
         sigma[ 0 ] = nrm2( col( R, 0 ), n );
      S.res = sigma[ 0 ] / b_norm;
   }
   S.finish( S.converged() );
}

// IDRs.hpp:47-316 with setSize :334-386: IDR(s) with the shadow space P of s
// columns drawn from std::mt19937 seeded 0 through
// std::normal_distribution<double>(0, 1), each column having the scalar
// (P_i, P_j) subtracted from every entry for j < i and then scaled to unit
// norm, as the reference does; optional minimal-residual smoothing.  One
// iteration per G-space step (s per cycle) and one for the omega step.
void orc_idrs( int64_t n, const int32_t* row_ptr, const int32_t* cols, const double* vals, const double* b, double* x,
               int precond, int s, int smoothing, orc_solver_state* st )
{
   Csr A{ n, row_ptr, cols, vals };
   Precond M;
   M.kind = precond;
   M.update( A );
   Iter S{ st };
   std::vector< double > P( (int64_t) s * n ), G( (int64_t) s * n, 0.0 ), Uv( (int64_t) s * n, 0.0 );
   std::vector< double > r( n ), v( n ), t( n ), x_s, r_s, Mm( s * s, 0.0 ), f( s );
   auto colp = [ & ]( std::vector< double >& V, int j ) { return V.data() + (int64_t) j * n; };
   {
      std::mt19937 rng;
      rng.seed( 0 );
      std::normal_distribution< double > dist( 0, 1 );
      for( auto& p : P )
         p = dist( rng );
      for( int i = 0; i < s; i++ ) {
         for( int j = 0; j < i; j++ ) {
            const double d = dot( colp( P, i ), colp( P, j ), n );
            for( int64_t q = 0; q < n; q++ )
               colp( P, i )[ q ] = colp( P, i )[ q ] - d;
         }
         const double nr = nrm2( colp( P, i ), n );
         for( int64_t q = 0; q < n; q++ )
            colp( P, i )[ q ] /= nr;
      }
   }
   auto Mat = [ & ]( int i, int j ) -> double& { return Mm[ i * s + j ]; };
   auto psolve = [ & ]( const double* src, double* dst ) {
      if( precond )
         M.solve( src, dst, n );
      else if( src != dst )
         std::memcpy( dst, src, sizeof( double ) * n );
   };
   const double b_norm = nrm2( b, n );
   if( b_norm == 0 ) {
      std::fill( x, x + n, 0.0 );
      S.finish( true );
      return;
   }
   A.mv( x, r.data() );
   for( int64_t i = 0; i < n; i++ )
      r[ i ] = b[ i ] - r[ i ];
   double r_norm = nrm2( r.data(), n );
   if( std::isnan( r_norm ) ) {
      S.res = r_norm;
      S.finish( false );  // the reference throws "initial residue is NAN"
      return;
   }
   const double angle = 0.7;
   double omega = 1.0;
   for( int i = 0; i < s; i++ )
      Mat( i, i ) = 1;
   if( smoothing ) {
      x_s.assign( x, x + n );
      r_s = r;
   }
   auto smooth = [ & ]() {
      for( int64_t i = 0; i < n; i++ )
         t[ i ] = r_s[ i ] - r[ i ];
      const double gamma = dot( t.data(), r_s.data(), n ) / dot( t.data(), t.data(), n );
      for( int64_t i = 0; i < n; i++ ) {
         r_s[ i ] = r_s[ i ] - gamma * t[ i ];
         x_s[ i ] = x_s[ i ] - gamma * ( x_s[ i ] - x[ i ] );
      }
      r_norm = nrm2( r_s.data(), n );
   };
   S.res = r_norm / b_norm;
   while( S.check() ) {
      for( int i = 0; i < s; i++ )
         f[ i ] = dot( colp( P, i ), r.data(), n );
      for( int k = 0; k < s; k++ ) {
         // c = M[k:s, k:s]^-1 f[k:s] (LUsequential.h: no pivoting)
         const int m = s - k;
         std::vector< double > L( m * m ), c( m );
         for( int i = 0; i < m; i++ )
            for( int j = 0; j < m; j++ )
               L[ i * m + j ] = Mat( i + k, j + k );
         for( int q = 0; q < m; q++ )
            for( int j = q + 1; j < m; j++ ) {
               const double factor = L[ j * m + q ] / L[ q * m + q ];
               L[ j * m + q ] = factor;
               for( int i = q + 1; i < m; i++ )
                  L[ j * m + i ] = L[ j * m + i ] - factor * L[ q * m + i ];
            }
         for( int i = 0; i < m; i++ )
            c[ i ] = f[ k + i ];
         for( int q = 1; q < m; q++ )
            for( int j = 0; j < q; j++ )
               c[ q ] -= L[ q * m + j ] * c[ j ];
         for( int q = m - 1; q >= 0; q-- ) {
            for( int j = m - 1; j > q; j-- )
               c[ q ] -= L[ q * m + j ] * c[ j ];
            c[ q ] /= L[ q * m + q ];
         }
         // v = r - G[:, k:s] c (gemv: column sum, then + y)
         for( int64_t i = 0; i < n; i++ ) {
            double a = 0;
            for( int j = 0; j < m; j++ )
               a += colp( G, k + j )[ i ] * ( -1.0 * c[ j ] );
            v[ i ] = a + 1.0 * r[ i ];
         }
         psolve( v.data(), v.data() );
         // U_k = U[:, k:s] c + omega v (gemv with beta 0: U_k is read before it is written, entry by entry)
         double* Uk = colp( Uv, k );
         for( int64_t i = 0; i < n; i++ ) {
            double a = 0;
            for( int j = 0; j < m; j++ )
               a += colp( Uv, k + j )[ i ] * ( 1.0 * c[ j ] );
            Uk[ i ] = a;
         }
         for( int64_t i = 0; i < n; i++ )
            Uk[ i ] += omega * v[ i ];
         double* Gk = colp( G, k );
         A.mv( Uk, Gk );
         for( int i = 0; i < k; i++ ) {
            const double al = dot( colp( P, i ), Gk, n ) / Mat( i, i );
            for( int64_t q = 0; q < n; q++ ) {
               Uk[ q ] = Uk[ q ] - al * colp( Uv, i )[ q ];
               Gk[ q ] = Gk[ q ] - al * colp( G, i )[ q ];
            }
         }
         for( int i = k; i < s; i++ )
            Mat( i, k ) = dot( colp( P, i ), Gk, n );
         if( Mat( k, k ) == 0 ) {
            S.finish( false );  // breakdown
            return;
         }
         const double beta = f[ k ] / Mat( k, k );
         for( int64_t i = 0; i < n; i++ ) {
            x[ i ] = x[ i ] + beta * Uk[ i ];
            r[ i ] = r[ i ] - beta * Gk[ i ];
         }
         r_norm = nrm2( r.data(), n );
         if( smoothing )
            smooth();
         S.res = r_norm / b_norm;
         if( ! S.next() ) {
            if( smoothing )
               std::memcpy( x, x_s.data(), sizeof( double ) * n );
            S.finish( S.converged() );
            return;
         }
         for( int i = k + 1; i < s; i++ )
            f[ i ] = f[ i ] - beta * Mat( i, k );
      }
      psolve( r.data(), v.data() );
      A.mv( v.data(), t.data() );
      const double t_norm = nrm2( t.data(), n );
      const double tr = dot( t.data(), r.data(), n );
      const double rho = std::abs( tr / ( t_norm * r_norm ) );
      omega = tr / ( t_norm * t_norm );
      if( rho < angle )
         omega = omega * angle / rho;
      if( omega == 0 ) {
         S.finish( false );  // breakdown
         return;
      }
      for( int64_t i = 0; i < n; i++ ) {
         x[ i ] = x[ i ] + omega * v[ i ];
         r[ i ] = r[ i ] - omega * t[ i ];
      }
      r_norm = nrm2( r.data(), n );
      if( smoothing )
         smooth();
      S.res = r_norm / b_norm;
      S.next();
   }
   if( smoothing )
      std::memcpy( x, x_s.data(), sizeof( double ) * n );
   S.finish( S.converged() );
}

}  // extern "C"
