// segments.hpp -- the device-resident segment formats (TNL's Algorithms::
// Segments) and their global-index functions.
#pragma once

#include <memory>
#include <vector>

#include "common.hpp"

namespace tnlmi {
struct AutoPlans;
}

// The opaque handle of the C ABI.  Every format also keeps the CSR offsets of
// its sizes (used only by assembly and the CSR-ordered index export, never by
// the format's own SpMV kernels).
struct tnlmi_segments
{
   tnlmi_format format = TNLMI_CSR;
   tnlmi_organization org = TNLMI_COLUMN_MAJOR;
   int32_t param = 0;
   int64_t n = 0;
   int64_t storage = 0;
   int64_t elements = 0;
   int32_t max_size = 0;

   tnlmi::DeviceArray< int32_t > sizes;
   tnlmi::DeviceArray< int32_t > offsets;  // exclusive scan of sizes, n+1

   // CSR tiled kernel plan: tile t covers elements [t*tile_nnz, (t+1)*tile_nnz)
   // and owns rows [tile_rows[t], tile_rows[t+1]), the rows whose first
   // element lies there; carry_row / carry_val hold the partial sums of rows
   // continued from an earlier tile (scratch of the SpMV: concurrent products
   // with one segments object on different streams are not allowed).
   tnlmi::DeviceArray< int32_t > tile_rows;
   tnlmi::DeviceArray< int32_t > carry_row;
   tnlmi::DeviceArray< double > carry_val;
   int64_t tiles = 0;
   int32_t tile_nnz = 0;

   // Ellpack (EllpackBase.h)
   int32_t ell_width = 0, ell_aligned = 0;

   // SlicedEllpack (SlicedEllpackBase.h)
   tnlmi::DeviceArray< int32_t > slice_offsets, slice_seg_sizes;
   int64_t slices = 0;

   // BiEllpack (BiEllpackBase.h) and SortedSegments permutations
   tnlmi::DeviceArray< int32_t > perm, group_ptrs, inv_perm;
   int64_t strips = 0;
   int32_t logw = 5;

   // ChunkedEllpack (ChunkedEllpackBase.h)
   tnlmi::DeviceArray< int32_t > ch_slices, seg_to_chunk, seg_to_slice, seg_ptrs, chunks_to_seg;
   int64_t ch_nslices = 0;
   int32_t chunks_in_slice = 256, desired_chunk_size = 16;

   // AdaptiveCSR descriptors per value size log (1,2,4,8,16,32 bytes)
   tnlmi::DeviceArray< int32_t > blocks[ 6 ];
   int64_t nblocks[ 6 ] = { 0, 0, 0, 0, 0, 0 };

   // SortedSegments: CSR offsets of the sorted sizes (the embedded CSR)
   tnlmi::DeviceArray< int32_t > sorted_offsets;

   // SlicedEllpack / BiEllpack / ChunkedEllpack loaded from a file: the `size`
   // field it held (the sum of the sizes the segments were first built from;
   // `elements` is the sum of the sizes as getSegmentSize derives them), written
   // back by save.  -1 otherwise.
   int64_t file_size = -1;

   // Product plans tnlmi_spmv built by itself for matrices on these segments
   // (keyed by their value and column-index arrays): dropped when the library
   // writes such a matrix, or on tnlmi_matrix_modified.
   mutable std::shared_ptr< tnlmi::AutoPlans > auto_plans;
};

namespace tnlmi {

// elements per wavefront tile of the tiled CSR kernel
constexpr int32_t kTileNnz = 512;
// longest row for which the tiled kernel completes rows in their owner's tile
// (one launch, no carries)
constexpr int32_t kWholeRowMax = 32;
// such a matrix of at most kSmallTileElems elements gets tiles of
// kTileNnzSmall: twice the wavefronts to hide a small product's memory latency
constexpr int32_t kTileNnzSmall = 256;
constexpr int64_t kSmallTileElems = int64_t( 1 ) << 22;
inline int32_t csr_tile_size( int32_t max_size, int64_t total )
{
   return max_size <= kWholeRowMax && total <= kSmallTileElems ? kTileNnzSmall : kTileNnz;
}

// Plain-value view of a segments object for kernels.
struct SegView
{
   int32_t format, col_major;
   int64_t n;
   const int32_t* sizes;
   const int32_t* offsets;
   // ellpack
   int32_t ell_width, ell_aligned;
   // sliced
   int32_t slice_size;
   const int32_t* slice_offsets;
   const int32_t* slice_seg_sizes;
   // biellpack / sorted
   const int32_t* perm;
   const int32_t* group_ptrs;
   const int32_t* inv_perm;
   int32_t warp, logw;
   // chunked
   const int32_t* ch_slices;
   const int32_t* seg_to_chunk;
   const int32_t* seg_to_slice;
   int32_t chunks_in_slice;
   // sorted
   const int32_t* sorted_offsets;
};

SegView make_view( const tnlmi_segments& s );

// getGlobalIndex( r, k ) of every format (reference file per branch).
__host__ __device__ inline int64_t global_index( const SegView& v, int64_t r, int32_t k )
{
   switch( v.format ) {
      case TNLMI_CSR:
      case TNLMI_ADAPTIVE_CSR:
         return v.offsets[ r ] + k;  // CSRBase.hpp:117-131
      case TNLMI_ELLPACK:            // EllpackBase.hpp:108-117
         return v.col_major ? r + (int64_t) v.ell_aligned * k : r * v.ell_width + k;
      case TNLMI_SLICED_ELLPACK: {  // SlicedEllpackBase.hpp:119-146
         const int64_t s = r / v.slice_size, in = r % v.slice_size;
         return v.col_major ? v.slice_offsets[ s ] + in + (int64_t) v.slice_size * k
                            : v.slice_offsets[ s ] + in * v.slice_seg_sizes[ s ] + k;
      }
      case TNLMI_BIELLPACK: {  // detail/BiEllpack.h getGlobalIndexDirect
         const int64_t strip = r / v.warp;
         const int64_t g0 = strip * ( v.logw + 1 );
         const int32_t rsp = v.perm[ r ] - (int32_t) ( strip * v.warp );
         int groups = 1;
         {
            int32_t bis = 1;
            for( int i = 0; i < v.logw + 1; i++ ) {
               if( rsp < bis ) {
                  groups = v.logw + 1 - i;
                  break;
               }
               bis *= 2;
            }
         }
         int32_t local = k, global = v.group_ptrs[ g0 ], height = v.warp;
         for( int g = 0; g < groups; g++ ) {
            const int32_t gsize = v.group_ptrs[ g0 + g + 1 ] - v.group_ptrs[ g0 + g ];
            if( gsize ) {
               const int32_t gw = gsize / height;
               if( local >= gw ) {
                  local -= gw;
                  global += gsize;
               }
               else
                  return v.col_major ? global + rsp + local * height : global + rsp * gw + local;
            }
            height /= 2;
         }
         return -1;
      }
      case TNLMI_CHUNKED_ELLPACK: {  // detail/ChunkedEllpack.h getGlobalIndexDirect
         const int32_t si = v.seg_to_slice[ r ];
         const int32_t first = r != v.ch_slices[ 4 * si + 2 ] ? v.seg_to_chunk[ r - 1 ] : 0;
         const int32_t off = v.ch_slices[ 4 * si + 3 ], cs = v.ch_slices[ 4 * si + 1 ];
         return v.col_major ? off + ( k % cs ) * v.chunks_in_slice + first + k / cs : off + first * cs + k;
      }
      case TNLMI_SORTED_CSR:  // SortedSegmentsBase.hpp:85-96
         return v.sorted_offsets[ v.perm[ r ] ] + k;
   }
   return -1;
}

// getSegmentSize( r ) of every format: the slots a traversal of row r visits
// (Ellpack: the width, EllpackBase.hpp; SlicedEllpack: its slice's width,
// SlicedEllpackBase.hpp; BiEllpack: the widths of the row's groups and
// ChunkedEllpack: its chunks x the slice's chunk size, getSegmentSizeDirect in
// detail/BiEllpack.h and detail/ChunkedEllpack.h; the CSR formats: the size)
__host__ __device__ inline int32_t segment_size( const SegView& v, int64_t r )
{
   switch( v.format ) {
      case TNLMI_ELLPACK:
         return v.ell_width;
      case TNLMI_SLICED_ELLPACK:
         return v.slice_seg_sizes[ r / v.slice_size ];
      case TNLMI_BIELLPACK: {
         const int64_t strip = r / v.warp;
         const int64_t g0 = strip * ( v.logw + 1 );
         const int32_t rsp = v.perm[ r ] - (int32_t) ( strip * v.warp );
         int groups = 1;
         for( int i = 0, bis = 1; i < v.logw + 1; i++, bis *= 2 )
            if( rsp < bis ) {
               groups = v.logw + 1 - i;
               break;
            }
         int32_t size = 0;
         for( int g = 0, height = v.warp; g < groups; g++, height /= 2 )
            size += ( v.group_ptrs[ g0 + g + 1 ] - v.group_ptrs[ g0 + g ] ) / height;
         return size;
      }
      case TNLMI_CHUNKED_ELLPACK: {
         const int32_t si = v.seg_to_slice[ r ];
         const int32_t first = r != v.ch_slices[ 4 * si + 2 ] ? v.seg_to_chunk[ r - 1 ] : 0;
         return ( v.seg_to_chunk[ r ] - first ) * v.ch_slices[ 4 * si + 1 ];
      }
      default:
         return v.sizes[ r ];
   }
}

// The index arrays and counters of a segments object as TNL's segments save
// and load them (host copies; the fields of other formats stay empty):
// CSR.hpp:116-130, Ellpack.hpp:105-120, SlicedEllpack.hpp:166-190,
// BiEllpack.hpp:222-240, ChunkedEllpack.hpp:240-275, AdaptiveCSR.hpp:145-160,
// SortedSegments.hpp:242-260.
struct SegmentsFileData
{
   int64_t elements = 0, storage = 0;  // `size` (sum of the sizes) and storageSize
   int32_t ell_width = 0, ell_aligned = 0, ell_count = 0;
   std::vector< int32_t > offsets;  // CSR, AdaptiveCSR; the embedded CSR of SortedSegments
   std::vector< int32_t > slice_offsets, slice_seg_sizes;
   std::vector< int32_t > perm, group_ptrs, inv_perm;
   int32_t ch_nslices = 0, chunks_in_slice = 0, desired_chunk_size = 0;
   std::vector< int32_t > seg_to_chunk, seg_to_slice, chunks_to_seg, seg_ptrs;
   std::vector< int32_t > ch_slices;  // 4 per slice (size, chunkSize, firstSegment, pointer), n slices as TNL sizes it
   std::vector< int32_t > blocks[ 6 ];  // AdaptiveCSR descriptors, 6 int32 each (type, first, size, segments, warp, warps)
};

// Copies the arrays of `seg` to the host in that layout.
void segments_export( const tnlmi_segments& seg, SegmentsFileData& f, hipStream_t s );

// Builds `seg` (format, organization, param and n set) from arrays read from a
// file: checks that every index they define lies inside the storage, derives
// the segment sizes as TNL's getSegmentSize does, and keeps the file's arrays
// (the AdaptiveCSR descriptors are rebuilt for this device).  Throws
// TNLMI_RUNTIME_ERROR on arrays that no TNL segments object can hold.
void segments_import( tnlmi_segments& seg, const SegmentsFileData& f, hipStream_t s );

// Exclusive scan of int32 values into int32 (n+1 outputs, out[n] = total);
// returns the total as int64 (checked by the caller for int32 overflow).
int64_t exclusive_scan_i32( const int32_t* d_in, int32_t* d_out, int64_t n, hipStream_t s );

}  // namespace tnlmi
