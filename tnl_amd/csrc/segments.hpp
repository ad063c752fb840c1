// segments.hpp -- the device-resident segment formats (TNL's Algorithms::
// Segments) and their global-index functions.
#pragma once

#include <memory>
#include <vector>

#include "common.hpp"
#include "tnlmi_segments_view.h"

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

// Plain-value view of a segments object for kernels: the public device-side
// layout (include/tnlmi_segments_view.h), which also holds the one
// implementation of getGlobalIndex / getSegmentSize for every format.
using SegView = tnlmi_segments_view;

SegView make_view( const tnlmi_segments& s );

__host__ __device__ inline int64_t global_index( const SegView& v, int64_t r, int32_t k )
{
   return tnlmi_global_index( &v, r, k );
}

__host__ __device__ inline int32_t segment_size( const SegView& v, int64_t r )
{
   return tnlmi_segment_size( &v, r );
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
