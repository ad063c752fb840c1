/* tnlmi_segments_view.h -- the device-side layout of a segments object of
 * libtnlmi.so (formats of tnlmi.h), for code that walks a matrix's rows on
 * the GPU: TNL's SparseMatrix::getRow / forRows / forElements on
 * Devices::MI355X (tnl_amd/include/TNL/MI355X.h, compiled with hipcc) and the
 * library's own kernels.  tnlmi_segments_get_view (tnlmi.h) fills it; the
 * pointers are device pointers owned by the segments object.
 *
 * tnlmi_global_index( v, r, k ) is the reference's getGlobalIndex( r, k ) and
 * tnlmi_segment_size( v, r ) its getSegmentSize( r ) for every format (the
 * reference file is cited per branch); both are host/device functions under a
 * HIP compiler (reading a device pointer on the host is the caller's error).
 */
#ifndef TNLMI_SEGMENTS_VIEW_H
#define TNLMI_SEGMENTS_VIEW_H

#include <stdint.h>

#include "tnlmi.h"

#if defined( __HIPCC__ ) || defined( __HIP__ )
#define TNLMI_HD __host__ __device__
#else
#define TNLMI_HD
#endif

struct tnlmi_segments_view
{
   int32_t format, col_major; /* tnlmi_format; 1: column-major slots */
   int64_t n;                 /* segments (rows) */
   const int32_t* sizes;      /* n requested segment sizes */
   const int32_t* offsets;    /* CSR, AdaptiveCSR: n + 1 */
   /* Ellpack */
   int32_t ell_width, ell_aligned;
   /* SlicedEllpack */
   int32_t slice_size;
   const int32_t* slice_offsets;
   const int32_t* slice_seg_sizes;
   /* BiEllpack; SortedSegments: perm */
   const int32_t* perm;
   const int32_t* group_ptrs;
   const int32_t* inv_perm;
   int32_t warp, logw;
   /* ChunkedEllpack: 4 per slice (size, chunk size, first segment, pointer) */
   const int32_t* ch_slices;
   const int32_t* seg_to_chunk;
   const int32_t* seg_to_slice;
   int32_t chunks_in_slice;
   /* SortedSegments: the embedded CSR's offsets */
   const int32_t* sorted_offsets;
};
typedef struct tnlmi_segments_view tnlmi_segments_view;

/* BiEllpack: the groups row r's strip gives it (detail/BiEllpack.h
 * getActiveGroupsCountDirect) */
static inline TNLMI_HD int tnlmi_biell_groups( const tnlmi_segments_view* v, int32_t rsp )
{
   int32_t bis = 1;
   for( int i = 0; i < v->logw + 1; i++ ) {
      if( rsp < bis )
         return v->logw + 1 - i;
      bis *= 2;
   }
   return 1;
}

/* getGlobalIndex( r, k ): the slot of the k-th element of segment r */
static inline TNLMI_HD int64_t tnlmi_global_index( const tnlmi_segments_view* v, int64_t r, int32_t k )
{
   switch( v->format ) {
      case TNLMI_CSR:
      case TNLMI_ADAPTIVE_CSR:
         return v->offsets[ r ] + k; /* CSRBase.hpp:117-131 */
      case TNLMI_ELLPACK:            /* EllpackBase.hpp:108-117 */
         return v->col_major ? r + (int64_t) v->ell_aligned * k : r * v->ell_width + k;
      case TNLMI_SLICED_ELLPACK: { /* SlicedEllpackBase.hpp:119-146 */
         const int64_t s = r / v->slice_size, in = r % v->slice_size;
         return v->col_major ? v->slice_offsets[ s ] + in + (int64_t) v->slice_size * k
                             : v->slice_offsets[ s ] + in * v->slice_seg_sizes[ s ] + k;
      }
      case TNLMI_BIELLPACK: { /* detail/BiEllpack.h getGlobalIndexDirect */
         const int64_t strip = r / v->warp;
         const int64_t g0 = strip * ( v->logw + 1 );
         const int32_t rsp = v->perm[ r ] - (int32_t) ( strip * v->warp );
         const int groups = tnlmi_biell_groups( v, rsp );
         int32_t local = k, global = v->group_ptrs[ g0 ], height = v->warp;
         for( int g = 0; g < groups; g++ ) {
            const int32_t gsize = v->group_ptrs[ g0 + g + 1 ] - v->group_ptrs[ g0 + g ];
            if( gsize ) {
               const int32_t gw = gsize / height;
               if( local >= gw ) {
                  local -= gw;
                  global += gsize;
               }
               else
                  return v->col_major ? global + rsp + local * height : global + rsp * gw + local;
            }
            height /= 2;
         }
         return -1;
      }
      case TNLMI_CHUNKED_ELLPACK: { /* detail/ChunkedEllpack.h getGlobalIndexDirect */
         const int32_t si = v->seg_to_slice[ r ];
         const int32_t first = r != v->ch_slices[ 4 * si + 2 ] ? v->seg_to_chunk[ r - 1 ] : 0;
         const int32_t off = v->ch_slices[ 4 * si + 3 ], cs = v->ch_slices[ 4 * si + 1 ];
         return v->col_major ? off + ( k % cs ) * v->chunks_in_slice + first + k / cs : off + first * cs + k;
      }
      case TNLMI_SORTED_CSR: /* SortedSegmentsBase.hpp:85-96 */
         return v->sorted_offsets[ v->perm[ r ] ] + k;
   }
   return -1;
}

/* getSegmentSize( r ): the slots a traversal of segment r visits (Ellpack: the
 * width, EllpackBase.hpp; SlicedEllpack: its slice's width,
 * SlicedEllpackBase.hpp; BiEllpack: the widths of the row's groups and
 * ChunkedEllpack: its chunks x the slice's chunk size, getSegmentSizeDirect in
 * detail/BiEllpack.h and detail/ChunkedEllpack.h; the CSR formats: the size) */
static inline TNLMI_HD int32_t tnlmi_segment_size( const tnlmi_segments_view* v, int64_t r )
{
   switch( v->format ) {
      case TNLMI_ELLPACK:
         return v->ell_width;
      case TNLMI_SLICED_ELLPACK:
         return v->slice_seg_sizes[ r / v->slice_size ];
      case TNLMI_BIELLPACK: {
         const int64_t strip = r / v->warp;
         const int64_t g0 = strip * ( v->logw + 1 );
         const int32_t rsp = v->perm[ r ] - (int32_t) ( strip * v->warp );
         const int groups = tnlmi_biell_groups( v, rsp );
         int32_t size = 0, height = v->warp;
         for( int g = 0; g < groups; g++, height /= 2 )
            size += ( v->group_ptrs[ g0 + g + 1 ] - v->group_ptrs[ g0 + g ] ) / height;
         return size;
      }
      case TNLMI_CHUNKED_ELLPACK: {
         const int32_t si = v->seg_to_slice[ r ];
         const int32_t first = r != v->ch_slices[ 4 * si + 2 ] ? v->seg_to_chunk[ r - 1 ] : 0;
         return ( v->seg_to_chunk[ r ] - first ) * v->ch_slices[ 4 * si + 1 ];
      }
      default:
         return v->sizes[ r ];
   }
}

#endif /* TNLMI_SEGMENTS_VIEW_H */
