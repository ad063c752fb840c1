// spmv_plan.hip -- XCD column-split execution plan of a CSR matrix.
//
// On the power-law matrix the CSR product is bound by its x gathers: uniformly
// random columns over a 33.5 MB x miss the 4 MB L2 of every XCD almost
// always.  The plan cuts the columns into P ranges (P = 8: 4 MiB of FP64 x per
// range at 4M columns, one L2) and keeps, per range, a CSR matrix over all
// rows with the elements whose column lies in that range (a reordered copy of
// the CSR arrays, built once).  The dispatcher places workgroup b on XCD
// b mod 8, so workgroup b works on range b mod 8 (and b mod 8 + 8 for P = 16):
// each XCD gathers from one L2-resident slice of x.  Each range writes its own
// partial row sums; a final pass adds the P partials of every row in range
// order (deterministic) and applies the alpha/beta store of
// SparseMatrixBase.hpp:346-389.
//
// No reference counterpart: TNL's vectorProduct has no analysis phase.  The
// plain CSR path (spmv.hip) reads the CSR arrays themselves.
#include <algorithm>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <vector>

#include "segments.hpp"
#include "spmv.hpp"
#include "spmv_window.hpp"
#include "spmv_tile.hpp"

constexpr int kMaxParts = 16;

struct tnlmi_spmv_plan
{
   int64_t rows = 0, columns = 0, nnz = 0;
   int32_t parts = 8;
   int64_t part_width = 0;
   tnlmi_dtype value_type = TNLMI_F64;
   tnlmi::DeviceArray< int32_t > rowptr;  // parts x (rows+1), each relative to its range start
   tnlmi::DeviceArray< int32_t > cols;    // nnz, ranges concatenated
   tnlmi::DeviceArray< char > vals;       // nnz values of value_type
   tnlmi::DeviceArray< int32_t > tile_rows;
   tnlmi::DeviceArray< int32_t > carry_row;
   tnlmi::DeviceArray< double > carry_val;
   tnlmi::DeviceArray< double > ypart;  // parts x rows partial sums (compute type)
   int64_t part_base[ kMaxParts + 1 ] = {};
   int64_t tiles[ kMaxParts ] = {};
   int64_t tile_base[ kMaxParts + 1 ] = {};     // first global tile of each range
   int64_t tilerow_base[ kMaxParts + 1 ] = {};  // first tile_rows entry of each range
   std::unique_ptr< tnlmi::WindowPlan > window;  // parts == 0: the windowed plan
};

namespace tnlmi {
namespace {

struct PlanArgs
{
   int32_t parts;
   int64_t rows, nnz_part[ kMaxParts ], part_base[ kMaxParts ], tiles[ kMaxParts ], tile_base[ kMaxParts ],
      tilerow_base[ kMaxParts ];
   int64_t blocks_xcd0[ kXcds ];  // workgroups of the first range of each XCD
};

__global__ void count_parts( const int32_t* __restrict__ off, const int32_t* __restrict__ cols, int64_t n,
                             int64_t width, int parts, int32_t* __restrict__ cnt )
{
   const int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( r >= n )
      return;
   int32_t c[ kMaxParts ] = {};
   for( int64_t g = off[ r ]; g < off[ r + 1 ]; g++ ) {
      const int32_t col = cols[ g ];
      if( col >= 0 )
         c[ min( (int64_t) parts - 1, col / width ) ]++;
   }
   for( int p = 0; p < parts; p++ )
      cnt[ p * n + r ] = c[ p ];
}

template< typename V >
__global__ void scatter_parts( const int32_t* __restrict__ off, const int32_t* __restrict__ cols,
                               const V* __restrict__ vals, int64_t n, int64_t width, PlanArgs pa,
                               const int32_t* __restrict__ rowptr, int32_t* __restrict__ ocols, V* __restrict__ ovals )
{
   const int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( r >= n )
      return;
   int64_t pos[ kMaxParts ];
   for( int p = 0; p < pa.parts; p++ )
      pos[ p ] = pa.part_base[ p ] + rowptr[ p * ( n + 1 ) + r ];
   for( int64_t g = off[ r ]; g < off[ r + 1 ]; g++ ) {
      const int32_t col = cols[ g ];
      if( col < 0 )
         continue;
      const int p = (int) min( (int64_t) pa.parts - 1, col / width );
      ocols[ pos[ p ] ] = col;
      ovals[ pos[ p ] ] = vals[ g ];
      pos[ p ]++;
   }
}

__global__ void tile_rows_kernel( const int32_t* __restrict__ offsets, int64_t n, int64_t tiles,
                                  int32_t* __restrict__ tile_rows )
{
   const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( t > tiles )
      return;
   if( t == tiles ) {
      tile_rows[ t ] = (int32_t) n;
      return;
   }
   const int64_t key = t * (int64_t) kTileNnz;
   int64_t lo = 0, hi = n;
   while( lo < hi ) {
      const int64_t mid = ( lo + hi ) / 2;
      if( offsets[ mid ] < key )
         lo = mid + 1;
      else
         hi = mid;
   }
   tile_rows[ t ] = (int32_t) lo;
}

// Workgroup b: XCD x = b mod 8, j = b / 8.  XCD x owns range x, then range
// x + 8 (P = 16); j indexes its workgroups over the ranges in that order.
template< typename V, typename X, typename C >
__global__ void __launch_bounds__( 256 ) split_main( PlanArgs pa, const int32_t* __restrict__ rowptr,
                                                     const int32_t* __restrict__ cols, const V* __restrict__ vals,
                                                     const X* __restrict__ x, const int32_t* __restrict__ tile_rows,
                                                     int32_t* __restrict__ carry_row, C* __restrict__ carry_val,
                                                     C* __restrict__ ypart )
{
   __shared__ C lds[ 4 ][ kTileNnz ];
   const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
   const int xcd = blockIdx.x % kXcds;
   int64_t j = blockIdx.x / kXcds;
   int p = xcd;
   if( j >= pa.blocks_xcd0[ xcd ] ) {
      j -= pa.blocks_xcd0[ xcd ];
      p = xcd + kXcds;
      if( p >= pa.parts )
         return;
   }
   const int64_t t = j * 4 + wave;
   if( t >= pa.tiles[ p ] )
      return;
   const int32_t* off = rowptr + p * ( pa.rows + 1 );
   const int32_t* tr = tile_rows + pa.tilerow_base[ p ];
   C* yp = ypart + p * pa.rows;
   const int64_t gt = pa.tile_base[ p ] + t;
   csr_tile< V, X, C >(
      off, pa.rows, cols + pa.part_base[ p ], vals + pa.part_base[ p ], x, t, tr[ t ], tr[ t + 1 ], pa.nnz_part[ p ],
      lds[ wave ], lane, [ & ]( int64_t r, C v ) { yp[ r ] = v; },
      [ & ]( int64_t rc, C v ) {
         carry_row[ gt ] = (int32_t) rc;
         carry_val[ gt ] = v;
      } );
}

template< typename C >
__global__ void __launch_bounds__( 256 ) split_fixup( PlanArgs pa, int64_t total_tiles,
                                                      const int32_t* __restrict__ carry_row,
                                                      const C* __restrict__ carry_val, C* __restrict__ ypart )
{
   const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( t >= total_tiles )
      return;
   const int32_t r = carry_row[ t ];
   // the first tile of every range never carries, so runs stay inside a range
   if( r < 0 || ( t > 0 && carry_row[ t - 1 ] == r ) )
      return;
   int p = 0;
   while( p + 1 < pa.parts && pa.tile_base[ p + 1 ] <= t )
      p++;
   const int64_t end = p + 1 < pa.parts ? pa.tile_base[ p + 1 ] : total_tiles;
   C sum = 0;
   for( int64_t u = t; u < end && carry_row[ u ] == r; u++ )
      sum = sum + carry_val[ u ];
   ypart[ p * pa.rows + r ] += sum;
}

template< typename X, typename C >
__global__ void __launch_bounds__( 256 ) split_reduce( int parts, int64_t rows, const C* __restrict__ ypart,
                                                       X* __restrict__ y, C alpha, C beta, int64_t begin, int64_t end )
{
   const int64_t r = begin + (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( r >= end )
      return;
   C s = 0;
   for( int p = 0; p < parts; p++ )
      s = s + ypart[ p * rows + r ];
   keep( y, r, s, alpha, beta );
}

PlanArgs plan_args( const tnlmi_spmv_plan& P )
{
   PlanArgs pa{};
   pa.parts = P.parts;
   pa.rows = P.rows;
   for( int p = 0; p < P.parts; p++ ) {
      pa.nnz_part[ p ] = P.part_base[ p + 1 ] - P.part_base[ p ];
      pa.part_base[ p ] = P.part_base[ p ];
      pa.tiles[ p ] = P.tiles[ p ];
      pa.tile_base[ p ] = P.tile_base[ p ];
      pa.tilerow_base[ p ] = P.tilerow_base[ p ];
   }
   for( int x = 0; x < kXcds; x++ )
      pa.blocks_xcd0[ x ] = x < P.parts ? ceil_div( P.tiles[ x ], 4 ) : 0;
   return pa;
}

template< typename V, typename X, typename C >
void execute_typed( const tnlmi_spmv_plan& P, const void* x, void* y, double alpha, double beta, int64_t begin,
                    int64_t end, hipStream_t s )
{
   const PlanArgs pa = plan_args( P );
   int64_t per_xcd = 0;
   for( int xx = 0; xx < kXcds; xx++ ) {
      int64_t b = xx < P.parts ? ceil_div( P.tiles[ xx ], 4 ) : 0;
      if( xx + kXcds < P.parts )
         b += ceil_div( P.tiles[ xx + kXcds ], 4 );
      per_xcd = std::max( per_xcd, b );
   }
   C* cv = reinterpret_cast< C* >( P.carry_val.ptr );
   C* yp = reinterpret_cast< C* >( P.ypart.ptr );
   const int64_t total_tiles = P.tile_base[ P.parts ];
   hipLaunchKernelGGL( ( split_main< V, X, C > ), dim3( (unsigned) std::max< int64_t >( 1, per_xcd * kXcds ) ),
                       dim3( 256 ), 0, s, pa, P.rowptr.ptr, P.cols.ptr, (const V*) P.vals.ptr, (const X*) x,
                       P.tile_rows.ptr, P.carry_row.ptr, cv, yp );
   hipLaunchKernelGGL( ( split_fixup< C > ), dim3( (unsigned) std::max< int64_t >( 1, ceil_div( total_tiles, 256 ) ) ),
                       dim3( 256 ), 0, s, pa, total_tiles, P.carry_row.ptr, cv, yp );
   if( end > begin )
      hipLaunchKernelGGL( ( split_reduce< X, C > ), dim3( (unsigned) ceil_div( end - begin, 256 ) ), dim3( 256 ), 0, s,
                          P.parts, P.rows, yp, (X*) y, (C) alpha, (C) beta, begin, end );
   TNLMI_HIP( hipGetLastError() );
}

template< typename V >
void build_typed( tnlmi_spmv_plan& P, const tnlmi_matrix& A, hipStream_t s )
{
   const tnlmi_segments& seg = *A.segments;
   const int64_t n = A.rows;
   DeviceArray< int32_t > cnt;
   cnt.allocate( std::max< int64_t >( 1, P.parts * n ) );
   if( n )
      hipLaunchKernelGGL( count_parts, dim3( (unsigned) ceil_div( n, 256 ) ), dim3( 256 ), 0, s, seg.offsets.ptr,
                          A.column_indexes, n, P.part_width, P.parts, cnt.ptr );
   P.rowptr.allocate( P.parts * ( n + 1 ) );
   P.part_base[ 0 ] = 0;
   for( int p = 0; p < P.parts; p++ ) {
      const int64_t tot = exclusive_scan_i32( cnt.ptr + p * n, P.rowptr.ptr + p * ( n + 1 ), n, s );
      P.part_base[ p + 1 ] = P.part_base[ p ] + tot;
   }
   P.nnz = P.part_base[ P.parts ];
   P.cols.allocate( std::max< int64_t >( 1, P.nnz ) );
   P.vals.allocate( std::max< int64_t >( 1, P.nnz ) * (int64_t) sizeof( V ) );
   PlanArgs pa = plan_args( P );
   if( n )
      hipLaunchKernelGGL( ( scatter_parts< V > ), dim3( (unsigned) ceil_div( n, 256 ) ), dim3( 256 ), 0, s,
                          seg.offsets.ptr, A.column_indexes, (const V*) A.values, n, P.part_width, pa, P.rowptr.ptr,
                          P.cols.ptr, (V*) P.vals.ptr );
   P.tile_base[ 0 ] = 0;
   P.tilerow_base[ 0 ] = 0;
   for( int p = 0; p < P.parts; p++ ) {
      P.tiles[ p ] = std::max< int64_t >( 1, ceil_div( P.part_base[ p + 1 ] - P.part_base[ p ], kTileNnz ) );
      P.tile_base[ p + 1 ] = P.tile_base[ p ] + P.tiles[ p ];
      P.tilerow_base[ p + 1 ] = P.tilerow_base[ p ] + P.tiles[ p ] + 1;
   }
   P.tile_rows.allocate( P.tilerow_base[ P.parts ] );
   for( int p = 0; p < P.parts; p++ )
      hipLaunchKernelGGL( tile_rows_kernel, dim3( (unsigned) ceil_div( P.tiles[ p ] + 1, 256 ) ), dim3( 256 ), 0, s,
                          P.rowptr.ptr + p * ( n + 1 ), n, P.tiles[ p ], P.tile_rows.ptr + P.tilerow_base[ p ] );
   P.carry_row.allocate( P.tile_base[ P.parts ] );
   P.carry_val.allocate( P.tile_base[ P.parts ] );
   P.ypart.allocate( std::max< int64_t >( 1, P.parts * n ) );
   TNLMI_HIP( hipGetLastError() );
   TNLMI_HIP( hipStreamSynchronize( s ) );
}

}  // namespace
}  // namespace tnlmi

using namespace tnlmi;

extern "C" {

tnlmi_status tnlmi_spmv_plan_create( const tnlmi_matrix* A, int32_t parts, tnlmi_stream stream,
                                     tnlmi_spmv_plan** out )
{
   return guarded( [ & ] {
      if( ! out || ! A || ! A->segments )
         fail( TNLMI_INVALID_ARGUMENT, "spmv_plan_create: NULL argument" );
      *out = nullptr;
      if( A->segments->format != TNLMI_CSR )
         fail( TNLMI_INVALID_ARGUMENT, "spmv_plan_create: the column-split plan is built from a CSR matrix" );
      if( A->matrix_type == TNLMI_SYMMETRIC )
         fail( TNLMI_NOT_IMPLEMENTED, "spmv_plan_create: the plans serve general matrices" );
      if( parts != 0 && parts != 8 && parts != 16 )
         fail( TNLMI_INVALID_ARGUMENT, "spmv_plan_create: parts must be 0 (windowed plan), 8 or 16" );
      if( A->rows != A->segments->n )
         fail( TNLMI_INVALID_ARGUMENT, "spmv_plan_create: matrix rows do not match the segments count" );
      auto* P = new tnlmi_spmv_plan;
      P->rows = A->rows;
      P->columns = A->columns;
      P->parts = parts;
      P->part_width = parts > 0 ? std::max< int64_t >( 1, ceil_div( A->columns, parts ) ) : 0;
      P->value_type = A->value_type;
      try {
         hipStream_t s = as_stream( stream );
         if( parts == 0 ) {
            P->window = std::make_unique< tnlmi::WindowPlan >();
            tnlmi::window_build( *P->window, *A, s );
         }
         else if( A->value_type == TNLMI_F64 )
            build_typed< double >( *P, *A, s );
         else if( A->value_type == TNLMI_F32 )
            build_typed< float >( *P, *A, s );
         else
            build_typed< bf16 >( *P, *A, s );
      }
      catch( ... ) {
         delete P;
         throw;
      }
      *out = P;
   } );
}

void tnlmi_spmv_plan_destroy( tnlmi_spmv_plan* plan )
{
   delete plan;
}

}  // extern "C"

namespace tnlmi {

void plan_execute( const tnlmi_spmv_plan& P, tnlmi_dtype xt, const void* d_x, void* d_y, tnlmi_dtype ct,
                   double alpha, double beta, int64_t begin, int64_t end, hipStream_t s )
{
   if( P.window ) {
      window_execute( *P.window, xt, d_x, d_y, ct, alpha, beta, begin, end, s );
      return;
   }
   const tnlmi_dtype vt = P.value_type;
   if( vt == TNLMI_F64 && xt == TNLMI_F64 && ct == TNLMI_F64 )
      execute_typed< double, double, double >( P, d_x, d_y, alpha, beta, begin, end, s );
   else if( vt == TNLMI_F32 && xt == TNLMI_F32 && ct == TNLMI_F32 )
      execute_typed< float, float, float >( P, d_x, d_y, alpha, beta, begin, end, s );
   else if( vt == TNLMI_F32 && xt == TNLMI_F64 && ct == TNLMI_F64 )
      execute_typed< float, double, double >( P, d_x, d_y, alpha, beta, begin, end, s );
   else if( vt == TNLMI_BF16 && xt == TNLMI_F32 && ct == TNLMI_F32 )
      execute_typed< bf16, float, float >( P, d_x, d_y, alpha, beta, begin, end, s );
   else if( vt == TNLMI_BF16 && xt == TNLMI_F64 && ct == TNLMI_F64 )
      execute_typed< bf16, double, double >( P, d_x, d_y, alpha, beta, begin, end, s );
   else
      fail( TNLMI_NOT_IMPLEMENTED, "vectorProduct: unsupported (value, vector, compute) type combination" );
}

}  // namespace tnlmi

extern "C" tnlmi_status tnlmi_spmv_plan_execute( const tnlmi_spmv_plan* P, tnlmi_dtype x_type, const void* d_x,
                                                 void* d_y, tnlmi_dtype compute_type, double alpha, double beta,
                                                 int64_t begin, int64_t end, tnlmi_stream stream )
{
   return guarded( [ & ] {
      if( ! P )
         fail( TNLMI_INVALID_ARGUMENT, "spmv_plan_execute: plan is NULL" );
      if( end == 0 )
         end = P->rows;
      if( begin < 0 || end > P->rows || begin > end )
         fail( TNLMI_INVALID_ARGUMENT, "vectorProduct: row range [begin, end) is out of bounds" );
      if( ( P->columns && ! d_x ) || ( P->rows && ! d_y ) )
         fail( TNLMI_INVALID_ARGUMENT, "vectorProduct: vector pointer is NULL" );
      plan_execute( *P, x_type, d_x, d_y, compute_type, alpha, beta, begin, end, as_stream( stream ) );
   } );
}

namespace tnlmi {

// The automatic plans of one segments object.
struct AutoPlans
{
   struct Entry
   {
      const void* values;
      const int32_t* cols;
      int32_t value_type;
      int64_t rows, columns;
      tnlmi_spmv_plan* plan;  // NULL: A's own kernels were chosen
      uint64_t last_use;
   };
   std::mutex m;
   std::vector< Entry > entries;
   uint64_t clock = 0;
   ~AutoPlans()
   {
      for( auto& e : entries )
         tnlmi_spmv_plan_destroy( e.plan );
   }
};

namespace {

constexpr int64_t kAutoMinColumns = int64_t( 1 ) << 20;   // x (FP32) of 4 MiB: wider than an XCD's L2
constexpr int64_t kAutoMinElements = int64_t( 1 ) << 22;
constexpr int64_t kAutoSampleRows = 4096;
constexpr int64_t kAutoFar = 65536;                        // columns from the row's diagonal position
constexpr std::size_t kAutoEntries = 4;

// stored entries of evenly spaced sample rows, and those whose column lies
// more than kAutoFar from the row's position scaled to the columns
__global__ void column_spread( const int32_t* __restrict__ off, const int32_t* __restrict__ cols, int64_t rows,
                               int64_t columns, int64_t samples, unsigned long long* counts )
{
   const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if( i >= samples )
      return;
   const int64_t r = i * rows / samples;
   const double diag = (double) r * (double) columns / (double) rows;
   unsigned long long total = 0, far = 0;
   for( int32_t g = off[ r ]; g < off[ r + 1 ]; g++ ) {
      const int32_t c = cols[ g ];
      if( c < 0 )
         continue;
      total++;
      far += fabs( (double) c - diag ) > (double) kAutoFar;
   }
   atomicAdd( counts, total );
   atomicAdd( counts + 1, far );
}

bool columns_spread( const tnlmi_matrix& A, hipStream_t s )
{
   const int64_t samples = std::min< int64_t >( A.rows, kAutoSampleRows );
   DeviceArray< unsigned long long > counts;
   counts.allocate( 2 );
   TNLMI_HIP( hipMemsetAsync( counts.ptr, 0, 2 * sizeof( unsigned long long ), s ) );
   hipLaunchKernelGGL( column_spread, dim3( (unsigned) ceil_div( samples, 256 ) ), dim3( 256 ), 0, s,
                       A.segments->offsets.ptr, A.column_indexes, A.rows, A.columns, samples, counts.ptr );
   TNLMI_HIP( hipGetLastError() );
   unsigned long long h[ 2 ] = { 0, 0 };
   TNLMI_HIP( hipMemcpyAsync( h, counts.ptr, sizeof( h ), hipMemcpyDeviceToHost, s ) );
   TNLMI_HIP( hipStreamSynchronize( s ) );
   return h[ 0 ] > 0 && 2 * h[ 1 ] >= h[ 0 ];
}

}  // namespace

const tnlmi_spmv_plan* auto_plan( const tnlmi_matrix& A, tnlmi_dtype xt, tnlmi_dtype ct, hipStream_t s )
{
   const tnlmi_segments& seg = *A.segments;
   if( seg.format != TNLMI_CSR || A.matrix_type == TNLMI_SYMMETRIC
       || ( A.value_type != TNLMI_F64 && A.value_type != TNLMI_F32 ) || xt != A.value_type || ct != A.value_type
       || A.columns < kAutoMinColumns || seg.elements < kAutoMinElements )
      return nullptr;
   static const bool disabled = [] {
      const char* e = std::getenv( "TNLMI_AUTO_PLAN" );
      return e && e[ 0 ] == '0';
   }();
   if( disabled )
      return nullptr;
   if( ! seg.auto_plans )
      seg.auto_plans = std::make_shared< AutoPlans >();
   AutoPlans& P = *seg.auto_plans;
   std::lock_guard< std::mutex > lock( P.m );
   for( auto& e : P.entries )
      if( e.values == A.values && e.cols == A.column_indexes && e.value_type == A.value_type && e.rows == A.rows
          && e.columns == A.columns ) {
         e.last_use = ++P.clock;
         return e.plan;
      }
   std::unique_ptr< tnlmi_spmv_plan > built;
   if( columns_spread( A, s ) ) {
      built = std::make_unique< tnlmi_spmv_plan >();
      built->rows = A.rows;
      built->columns = A.columns;
      built->parts = 0;
      built->value_type = A.value_type;
      built->window = std::make_unique< WindowPlan >();
      window_build( *built->window, A, s );
   }
   tnlmi_spmv_plan* plan = built.release();
   if( P.entries.size() >= kAutoEntries ) {
      auto lru = std::min_element( P.entries.begin(), P.entries.end(),
                                   []( const auto& a, const auto& b ) { return a.last_use < b.last_use; } );
      tnlmi_spmv_plan_destroy( lru->plan );
      P.entries.erase( lru );
   }
   P.entries.push_back( { A.values, A.column_indexes, A.value_type, A.rows, A.columns, plan, ++P.clock } );
   return plan;
}

void drop_auto_plans( const tnlmi_segments& seg, const void* values, const int32_t* cols )
{
   if( ! seg.auto_plans )
      return;
   AutoPlans& P = *seg.auto_plans;
   std::lock_guard< std::mutex > lock( P.m );
   for( auto it = P.entries.begin(); it != P.entries.end(); ) {
      if( ( ! values && ! cols ) || it->values == values || it->cols == cols ) {
         tnlmi_spmv_plan_destroy( it->plan );
         it = P.entries.erase( it );
      }
      else
         ++it;
   }
}

}  // namespace tnlmi

extern "C" tnlmi_status tnlmi_matrix_modified( const tnlmi_matrix* A )
{
   return guarded( [ & ] {
      if( ! A || ! A->segments )
         fail( TNLMI_INVALID_ARGUMENT, "matrix_modified: NULL argument" );
      drop_auto_plans( *A->segments, A->values, A->column_indexes );
   } );
}
