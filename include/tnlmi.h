/*
 * tnlmi.h -- C ABI of the MI355X (gfx950) sparse matrix-vector product and
 * Krylov engine.  Plain C: pointers, sizes and enums only.
 *
 * The reference (TNL, src/TNL/...) exposes this path as C++ templates selected
 * by a Device tag.  Each entry point below names the reference interface it
 * replaces.  A C++ template layer (tnl_amd/include/TNL/...) routes TNL's
 * SparseMatrix / Segments / Solvers::Linear calls for the new device to these
 * functions and turns a non-OK status back into the exception type the
 * reference throws (tnlmi_status values are listed with that type).
 *
 * Memory: every array argument named d_* is device memory on the current HIP
 * device, owned by the caller unless stated otherwise.  Segment objects own
 * the device arrays they build.  Calls taking a stream are asynchronous with
 * respect to the host unless stated otherwise; results are visible after the
 * stream is synchronised (the reference's reduceSegments synchronises before
 * returning, Algorithms/Segments/detail/ReducingOperations_CSR.h:327 -- the C++
 * layer does so too).
 */
#ifndef TNLMI_H
#define TNLMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* tnlmi_stream; /* a hipStream_t; NULL = the null stream */

/* Status codes.  Each names the exception the reference throws in that case. */
typedef enum {
   TNLMI_OK = 0,
   TNLMI_INVALID_ARGUMENT = 1, /* std::invalid_argument (Matrices/SparseMatrixBase.hpp:279-282) */
   TNLMI_RUNTIME_ERROR = 2,    /* std::runtime_error (Segments/detail/ReducingOperations_CSR.h:108-109) */
   TNLMI_OVERFLOW_ERROR = 3,   /* std::overflow_error (Segments/Ellpack.hpp:91-95) */
   TNLMI_BACKEND_ERROR = 4,    /* Exceptions::BackendRuntimeError (Backend/Macros.h:10-39) */
   TNLMI_NOT_IMPLEMENTED = 5,  /* Exceptions::NotImplementedError */
   TNLMI_BAD_ALLOC = 6,        /* Exceptions::BackendBadAlloc */
   TNLMI_LOGIC_ERROR = 7       /* std::logic_error (e.g. GMRES without a matrix, GMRES.hpp:75) */
} tnlmi_status;

/* Message of the last non-OK status returned on the calling thread. */
const char* tnlmi_last_error( void );
/* Library version string and the gfx target it was built for. */
const char* tnlmi_version( void );

typedef enum { TNLMI_F64 = 0, TNLMI_F32 = 1, TNLMI_BF16 = 2 } tnlmi_dtype;

/* Segment formats (the classes of Algorithms/Segments/). */
typedef enum {
   TNLMI_CSR = 0,             /* Segments/CSR.h */
   TNLMI_ADAPTIVE_CSR = 1,    /* Segments/AdaptiveCSR.h */
   TNLMI_ELLPACK = 2,         /* Segments/Ellpack.h (Alignment 32) */
   TNLMI_SLICED_ELLPACK = 3,  /* Segments/SlicedEllpack.h (param = SliceSize) */
   TNLMI_BIELLPACK = 4,       /* Segments/BiEllpack.h (param = WarpSize, 32) */
   TNLMI_CHUNKED_ELLPACK = 5, /* Segments/ChunkedEllpack.h (256 chunks/slice, chunk 16) */
   TNLMI_SORTED_CSR = 6       /* Segments/SortedSegments.h over CSR */
} tnlmi_format;

/* Segments/ElementsOrganization.h:13-19 (same numeric values). */
typedef enum { TNLMI_COLUMN_MAJOR = 0, TNLMI_ROW_MAJOR = 1 } tnlmi_organization;

/* ------------------------------------------------------------ segments */

typedef struct tnlmi_segments tnlmi_segments;

/* Builds the format's index arrays from per-segment sizes (the format
 * converter).  Replaces <Format>::setSegmentsSizes(sizes), e.g.
 * Segments/CSR.hpp:86-102, Ellpack.hpp:72-97, SlicedEllpack.hpp:116-152,
 * BiEllpack.hpp:104-211, ChunkedEllpack.hpp:156-216, AdaptiveCSR.hpp:102-126,
 * SortedSegments.hpp:129-200.  param: SliceSize for SlicedEllpack (power of
 * two, default 32), WarpSize for BiEllpack (32); ignored otherwise (0).
 * d_sizes: n int32 on the device.  Synchronous (returns once built). */
tnlmi_status tnlmi_segments_create( tnlmi_format format, tnlmi_organization organization, int32_t param,
                                    const int32_t* d_sizes, int64_t n, tnlmi_stream stream,
                                    tnlmi_segments** out );
void tnlmi_segments_destroy( tnlmi_segments* segments );

/* getSegmentCount() / getStorageSize() / getElementCount() (the sum of the
 * segment sizes; for segments loaded from a file, the sizes as getSegmentSize
 * derives them from the file's arrays) */
int64_t tnlmi_segments_count( const tnlmi_segments* segments );
int64_t tnlmi_segments_storage_size( const tnlmi_segments* segments );
int64_t tnlmi_segments_element_count( const tnlmi_segments* segments );
tnlmi_format tnlmi_segments_format( const tnlmi_segments* segments );

/* The device-side layout of a segments object (tnlmi_segments_view.h: the
 * format's index arrays as device pointers owned by the object, and the
 * getGlobalIndex / getSegmentSize functions over them), for code that walks
 * the rows of a matrix on the GPU -- the reference's SparseMatrix::getRow,
 * forRows and forElements (Matrices/SparseMatrixBase.h:239-257, 552-606,
 * 858-925) over Algorithms::Segments::forElements / forSegments.  The view is
 * valid while the segments object lives. */
struct tnlmi_segments_view;
tnlmi_status tnlmi_segments_get_view( const tnlmi_segments* segments, struct tnlmi_segments_view* out );

/* Index arrays of a segments object, for inspection (device pointers owned by
 * the object).  which: see tnlmi_segments_array.  Returns
 * TNLMI_INVALID_ARGUMENT if the format has no such array. */
typedef enum {
   TNLMI_ARR_OFFSETS = 0,             /* CSR / AdaptiveCSR / SortedCSR (embedded CSR) offsets, n+1 */
   TNLMI_ARR_SLICE_OFFSETS = 1,       /* SlicedEllpack sliceOffsets, slices+1 */
   TNLMI_ARR_SLICE_SEGMENT_SIZES = 2, /* SlicedEllpack sliceSegmentSizes, slices */
   TNLMI_ARR_PERMUTATION = 3,         /* BiEllpack / Sorted segmentsPermutation, n */
   TNLMI_ARR_GROUP_POINTERS = 4,      /* BiEllpack groupPointers, strips*(log2 W+1)+1 */
   TNLMI_ARR_INVERSE_PERMUTATION = 5, /* Sorted inverseSegmentsPermutation, n */
   TNLMI_ARR_CHUNK_SLICES = 6,        /* ChunkedEllpack slices {size,chunkSize,firstSegment,pointer}, 4*slices */
   TNLMI_ARR_SEGMENT_TO_CHUNK = 7,    /* ChunkedEllpack segmentToChunkMapping, n */
   TNLMI_ARR_SEGMENT_TO_SLICE = 8,    /* ChunkedEllpack segmentToSliceMapping, n */
   TNLMI_ARR_SEGMENT_POINTERS = 9,    /* ChunkedEllpack segmentPointers, n+1 */
   TNLMI_ARR_CHUNKS_TO_SEGMENTS = 10, /* ChunkedEllpack chunksToSegmentsMapping, slices*256 */
   TNLMI_ARR_ADAPTIVE_BLOCKS = 11,    /* AdaptiveCSR descriptors, 6 int32 each (see below) */
   TNLMI_ARR_SIZES = 12,              /* the segment sizes the object was built from, n */
   TNLMI_ARR_TILE_ROWS = 13           /* CSR tile -> first row table of the tiled kernel */
} tnlmi_segments_array;
tnlmi_status tnlmi_segments_get_array( const tnlmi_segments* segments, tnlmi_segments_array which,
                                       const int32_t** d_array, int64_t* length );
/* Ellpack: segmentSize and alignedSize (EllpackBase.h). */
tnlmi_status tnlmi_segments_ellpack_dims( const tnlmi_segments* segments, int32_t* segment_size,
                                          int32_t* aligned_size );

/* AdaptiveCSR descriptors (Segments/detail/CSRAdaptiveKernelBlockDescriptor.h)
 * are exported as 6 int32 {type, firstSegment, blockSize, segmentsInBlock,
 * warpIdx, warpsCount}, type LONG=0 STREAM=1 VECTOR=2, built for value size
 * value_bytes (1..32) with the wave64 device semantics. */
tnlmi_status tnlmi_segments_adaptive_blocks( const tnlmi_segments* segments, int32_t value_bytes,
                                             const int32_t** d_blocks, int64_t* count );

/* For every (segment r, local k < size[r]) write getGlobalIndex(r, k) into
 * d_gidx[csr_offset(r) + k] (CSR-ordered, element-count entries).  Replaces
 * <Format>Base::getGlobalIndex. */
tnlmi_status tnlmi_segments_global_indexes( const tnlmi_segments* segments, int32_t* d_gidx, tnlmi_stream stream );

/* -------------------------------------------------------------- matrix */

/* Matrices/MatrixType.h: GeneralMatrix / SymmetricMatrix. */
typedef enum { TNLMI_GENERAL = 0, TNLMI_SYMMETRIC = 1 } tnlmi_matrix_type;

/* A sparse matrix: Matrices/SparseMatrix.h:41-56.  values / column_indexes
 * have the segments' storage size; column index -1 marks a padding slot
 * (MatrixBase.h:27).  matrix_type TNLMI_SYMMETRIC: the storage holds the lower
 * triangle with the diagonal of a symmetric matrix (SymmetricMatrix); a
 * zero-initialised trailing field means a general matrix. */
typedef struct {
   int64_t rows;
   int64_t columns;
   const tnlmi_segments* segments;
   tnlmi_dtype value_type;
   const void* values;
   const int32_t* column_indexes;
   int32_t matrix_type; /* tnlmi_matrix_type */
} tnlmi_matrix;

/* Writes a CSR-ordered element list into the format storage: padding slots get
 * column -1 and value 0 (SparseMatrix.hpp:316-322), element k of row r goes to
 * getGlobalIndex(r, k) (RowView::setElement).  d_csr_offsets must be the
 * exclusive scan of the segment sizes; column -1 inside a row is kept as
 * padding.  Values are converted from csr_type to value_type (bf16: round to
 * nearest even; f64 goes through f32 first).  Replaces SparseMatrix::setElements / forAllElements-based
 * assembly. */
tnlmi_status tnlmi_matrix_fill( const tnlmi_segments* segments, const int32_t* d_csr_offsets,
                                const int32_t* d_csr_columns, const void* d_csr_values, tnlmi_dtype csr_type,
                                void* d_values, tnlmi_dtype value_type, int32_t* d_column_indexes,
                                tnlmi_stream stream );

/* ------------------------------------------------- cross-format copy */
/* copySparseMatrix( dst, src ) on the device (Matrices/SparseOperations.hpp:
 * 600-690): tnlmi_matrix_row_lengths gives the target's row capacities -- the
 * leading entries of each source row whose column is below the column count
 * (padding slots, column -1, are counted, as the reference counts them) --
 * synchronously into d_lengths (rows int32).  The caller builds the target's
 * segments from them (tnlmi_segments_create) and allocates its storage;
 * tnlmi_matrix_copy clears that storage (value 0, column -1) and copies
 * element k of every row (column and value) through both formats'
 * getGlobalIndex.  Same value type and dimensions (TNLMI_INVALID_ARGUMENT
 * otherwise, and for a row length above a segment size).  Synchronous. */
tnlmi_status tnlmi_matrix_row_lengths( const tnlmi_matrix* A, int32_t* d_lengths, tnlmi_stream stream );
tnlmi_status tnlmi_matrix_copy( const tnlmi_matrix* src, tnlmi_matrix* dst, const int32_t* d_lengths,
                                tnlmi_stream stream );

/* ---------------------------------------------------- element access */
/* SparseMatrixBase::addElement( row, column, value, thisElementMultiplicator )
 * (Matrices/SparseMatrixBase.hpp:166-231; setElement, :152-161, is
 * this_element_multiplicator = 0): writes A's storage.  The row's
 * getSegmentSize slots (Ellpack: the full width; SlicedEllpack: the slice's
 * width) are walked in getGlobalIndex order: a stored `column` becomes
 * multiplicator * old + value (rounded once per operation); otherwise the
 * element takes the first padding slot, or the slot of the first larger
 * column with the slots after it shifted right (a full row loses its last
 * slot, as in the reference).  A symmetric matrix stores (column, row) for
 * row < column.  TNLMI_LOGIC_ERROR "The capacity of the sparse matrix row
 * number r was exceeded." when the row has neither; TNLMI_INVALID_ARGUMENT for
 * a position outside the matrix.  Synchronous. */
tnlmi_status tnlmi_matrix_add_element( tnlmi_matrix* A, int64_t row, int64_t column, double value,
                                       double this_element_multiplicator, tnlmi_stream stream );
/* A list of `count` element operations in one launch: operation i is
 * addElement( d_rows[i], d_columns[i], d_values[i], d_multiplicators[i] )
 * (d_multiplicators NULL: setElement, multiplicator 0), each exactly as
 * tnlmi_matrix_add_element does it.  The operations of one row are applied in
 * list order; rows are independent, so the matrix ends as after calling
 * tnlmi_matrix_add_element for i = 0, 1, ... in turn.  Rows run in parallel
 * (one thread per row, operations sorted by row on the device): the device
 * arrays (int64 positions, FP64 values) are read once, with one host
 * synchronisation for the whole list.  Errors: TNLMI_INVALID_ARGUMENT if a
 * position lies outside the matrix (nothing is written); TNLMI_LOGIC_ERROR
 * "The capacity of the sparse matrix row number r was exceeded." when an
 * operation finds its row full -- that row stops there, other rows' operations
 * are applied, and *failed (may be NULL; -1 otherwise) is the smallest list
 * index of such an operation.  Synchronous.  No reference counterpart as one
 * call (TNL assembles a matrix with addElement / setElement calls or with
 * RowView::setElement inside forAllRows, SparseMatrix.hpp:336-470): it is the
 * device-side assembly of many elements. */
tnlmi_status tnlmi_matrix_add_elements( tnlmi_matrix* A, int64_t count, const int64_t* d_rows,
                                        const int64_t* d_columns, const double* d_values,
                                        const double* d_multiplicators, int64_t* failed, tnlmi_stream stream );
/* Per-row counts over each row's getSegmentSize slots, into d_counts (rows
 * int32), synchronous (SparseMatrixBase.hpp:55-132; padding slots count as 0):
 *   0  getRowCapacity( r ): the segment size (Ellpack: the width)
 *   1  getRowCapacities: the stored (non-padding) slots
 *   2  getCompressedRowLengths: the stored slots with a nonzero value
 *   3  getNonzeroElementsCount of a symmetric matrix, per row: a stored slot
 *      counts 1, plus 1 if it lies off the diagonal inside the matrix
 * (getNonzeroElementsCount of a general matrix is the sum of kind 2.) */
tnlmi_status tnlmi_matrix_row_counts( const tnlmi_matrix* A, int32_t kind, int32_t* d_counts, tnlmi_stream stream );
/* SparseMatrixBase::getElement( row, column ) (SparseMatrixBase.hpp:231-265):
 * *value = the stored element (FP32 / BF16 widened), 0 if the row holds no
 * such column.  Synchronous. */
tnlmi_status tnlmi_matrix_get_element( const tnlmi_matrix* A, int64_t row, int64_t column, double* value,
                                       tnlmi_stream stream );

/* ---------------------------------------------------- MatrixMarket input */
/* A MatrixMarket coordinate file (real / integer / pattern, general /
 * symmetric) read into a host CSR matrix: rows ascending, columns ascending
 * in each row, a repeated position keeps the last value, symmetric files are
 * mirrored, pattern entries are 1.  Replaces MatrixReader< Matrix, Host >::
 * readMtx (Matrices/MatrixReader.hpp:52-66, 158-306) followed by
 * SparseMatrix::setElements of its map; the same header and line checks and
 * messages (TNLMI_RUNTIME_ERROR; indexes below 1 are rejected as well).
 * tnlmi_mtx_parse reads the file's text from memory (the istream overload,
 * MatrixReader.hpp:26-32).  The arrays returned by tnlmi_host_csr_get stay
 * valid until tnlmi_host_csr_destroy. */
typedef struct tnlmi_host_csr tnlmi_host_csr;
tnlmi_status tnlmi_mtx_read( const char* path, tnlmi_host_csr** out );
tnlmi_status tnlmi_mtx_parse( const char* text, int64_t length, tnlmi_host_csr** out );
tnlmi_status tnlmi_host_csr_get( const tnlmi_host_csr* m, int64_t* rows, int64_t* columns, int64_t* nnz,
                                 const int32_t** offsets, const int32_t** columns_out, const double** values );
void tnlmi_host_csr_destroy( tnlmi_host_csr* m );

/* Binary files in TNL's format (File / Object serialization).
 * tnlmi_matrix_save replaces SparseMatrix::save (Matrices/SparseMatrix.hpp:
 * 642-646 with operator<< of SparseMatrixBase.hpp:1149-1159 and the segments'
 * save, e.g. Segments/CSR.hpp:116-121): the type string, rows, columns, the
 * segments' fields and arrays, values, column indexes -- the bytes TNL writes
 * for the same matrix (AdaptiveCSR: its kernel descriptors as built for
 * wave64; ChunkedEllpack: slice entries past the last slice written as zero).
 * FP64 and FP32 values (BF16: TNLMI_NOT_IMPLEMENTED).  Synchronous.
 * tnlmi_matrix_serialization_type: SparseMatrixBase::getSerializationType
 * (SparseMatrixBase.hpp:46-53), written to buffer (NUL-terminated, cut to
 * capacity) with its full length in *length. */
tnlmi_status tnlmi_matrix_save( const tnlmi_matrix* A, const char* path, tnlmi_stream stream );
tnlmi_status tnlmi_matrix_serialization_type( const tnlmi_matrix* A, char* buffer, int64_t capacity,
                                              int64_t* length );
/* SparseMatrix::load (SparseMatrix.hpp:656-694): tnlmi_matrix_file_read parses
 * and checks a file on the host (errors: TNLMI_RUNTIME_ERROR with the messages
 * of Exceptions/FileDeserializationError.h; a type this library cannot hold:
 * TNLMI_NOT_IMPLEMENTED).  Every index the file's segments define must lie in
 * its storage and every column index in [-1, columns).  Then
 * tnlmi_matrix_file_info gives the matrix type and storage size,
 * tnlmi_matrix_file_segments builds the segments (the file's own arrays; the
 * segment sizes as getSegmentSize derives them), tnlmi_matrix_file_upload
 * copies values and column indexes into caller storage of that size.  The
 * caller compares tnlmi_matrix_file_type with its own type where TNL would
 * (object type does not match ...). */
typedef struct tnlmi_matrix_file tnlmi_matrix_file;
tnlmi_status tnlmi_matrix_file_read( const char* path, tnlmi_matrix_file** out );
tnlmi_status tnlmi_matrix_file_info( const tnlmi_matrix_file* f, int64_t* rows, int64_t* columns,
                                     tnlmi_format* format, tnlmi_organization* organization, int32_t* param,
                                     tnlmi_dtype* value_type, int64_t* storage, tnlmi_matrix_type* matrix_type );
const char* tnlmi_matrix_file_type( const tnlmi_matrix_file* f );
tnlmi_status tnlmi_matrix_file_segments( const tnlmi_matrix_file* f, tnlmi_stream stream, tnlmi_segments** out );
tnlmi_status tnlmi_matrix_file_upload( const tnlmi_matrix_file* f, void* d_values, int32_t* d_column_indexes,
                                       tnlmi_stream stream );
void tnlmi_matrix_file_destroy( tnlmi_matrix_file* f );

/* Segments/LaunchConfiguration.h: ThreadsToSegmentsMapping (same values) plus
 * TNLMI_MAP_AUTO, the MI355X-tuned default (tiled nnz-balanced CSR). */
typedef enum {
   TNLMI_MAP_FIXED = 0,
   TNLMI_MAP_WARP = 1,
   TNLMI_MAP_BLOCK = 2,
   TNLMI_MAP_BLOCK_MERGED = 3,
   TNLMI_MAP_DYNAMIC_GROUPING = 4,
   TNLMI_MAP_AUTO = 255
} tnlmi_mapping;
typedef struct {
   int32_t mapping;             /* tnlmi_mapping */
   int32_t threads_per_segment; /* power of two (LaunchConfiguration.h:36-42) */
} tnlmi_launch;

/* y = alpha*A*x + beta*y over rows [begin, end) (end == 0: all rows).
 * Replaces SparseMatrixBase::vectorProduct (Matrices/SparseMatrixBase.hpp:
 * 267-391) with its reduceSegments dispatch.  launch == NULL means the
 * default (TNLMI_MAP_AUTO).  Supported (value, vector, compute) types:
 * (f64,f64,f64), (f32,f32,f32), (f32,f64,f64), (bf16,f32,f32), (bf16,f64,f64).
 * x has `columns` entries of x_type, y has `rows` entries of x_type. */
tnlmi_status tnlmi_spmv( const tnlmi_matrix* A, tnlmi_dtype x_type, const void* d_x, void* d_y,
                         tnlmi_dtype compute_type, double alpha, double beta, int64_t begin, int64_t end,
                         const tnlmi_launch* launch, tnlmi_stream stream );
/* A symmetric matrix (matrix_type TNLMI_SYMMETRIC) in tnlmi_spmv: the
 * reference's symmetric branch (SparseMatrixBase.hpp:298-329): y = beta*y over
 * the whole vector when beta != 1 (beta == 0: y = 0, where the reference's
 * multiplication would keep a NaN already in y), then for every row r in [begin, end)
 * y[r] += alpha * sum_k v_k x[c_k] and, for each stored c_k < r,
 * y[c_k] += alpha * v_k * x[r].  The adds are atomic, so the summation order
 * of a y entry is not fixed (as in the reference's GPU build); compute type =
 * vector type; launch is ignored. */

/* transposedVectorProduct (SparseMatrixBase.hpp:541-588): y = beta*y (whole
 * vector, when beta != 1), then y[c] += alpha * x[r] * v for every stored
 * element (r, c, v) with c in [begin, end) (end == 0: all columns).  x has
 * `rows` entries, y `columns`.  Atomic adds.  A symmetric matrix gives
 * tnlmi_spmv's product (the reference calls vectorProduct).  Types as
 * tnlmi_spmv with compute type = vector type. */
tnlmi_status tnlmi_spmv_transposed( const tnlmi_matrix* A, tnlmi_dtype x_type, const void* d_x, void* d_y,
                                    double alpha, double beta, int64_t begin, int64_t end, tnlmi_stream stream );

/* ------------------------------------------- column-split / windowed plan */
/* An execution plan of y = alpha*A*x + beta*y for a CSR matrix whose x
 * gathers are random (power-law graphs).  parts = 8 or 16: the columns are
 * cut into `parts` ranges (one or two per XCD) and a reordered copy of the CSR
 * arrays is kept per range, so that every XCD gathers from one L2-sized slice
 * of x; per-range partial row sums are added in range order.  parts = 0: the
 * windowed plan -- the columns are cut into windows of 8192 entries, each
 * window's products are formed with x staged in LDS and summed per row block
 * in LDS; every row of at most 256 nonzeros is summed left to right in
 * column order from zero, bit for bit as the host CSR vectorProduct, longer
 * rows in a fixed order (rows longer than a row block's 10000 FP64 / 20000
 * FP32 nonzeros as pieces whose partial sums are added in order).  FP64 or
 * FP32 values, vectors and compute type alike.  Built once (synchronous); the plan
 * owns its copy (the matrix may be freed), so value changes of the matrix
 * need a new plan.  Results equal vectorProduct up to the FP summation order.
 * No reference counterpart (TNL's vectorProduct has no analysis phase); it
 * serves SparseMatrix::vectorProduct of CSR matrices.  Concurrent executions
 * of one plan on different streams are not allowed (the plan holds the
 * partial sums / products). */
typedef struct tnlmi_spmv_plan tnlmi_spmv_plan;
tnlmi_status tnlmi_spmv_plan_create( const tnlmi_matrix* A, int32_t parts, tnlmi_stream stream,
                                     tnlmi_spmv_plan** out );
void tnlmi_spmv_plan_destroy( tnlmi_spmv_plan* plan );
/* Under TNLMI_MAP_AUTO, tnlmi_spmv runs the products of a general CSR matrix
 * with FP64 or FP32 values, vectors and compute type through a windowed plan
 * it builds once and keeps with the segments (20 extra bytes per nonzero),
 * when x is wider than an L2 (at least 2^20 columns, 2^22 stored elements)
 * and the columns of sampled rows mostly lie far from the diagonal: the
 * reference's default vectorProduct call then gets the fastest product of
 * such a matrix.  The plan belongs to the matrix's value and column-index
 * arrays; the library drops it when it writes them (tnlmi_matrix_fill,
 * tnlmi_matrix_add_element, tnlmi_matrix_copy into them).  A caller that
 * writes them itself calls tnlmi_matrix_modified( A ) before the next
 * product.  TNLMI_AUTO_PLAN=0 in the environment turns this off.
 * No reference counterpart (the reference's matrices keep no derived data). */
tnlmi_status tnlmi_matrix_modified( const tnlmi_matrix* A );
tnlmi_status tnlmi_spmv_plan_execute( const tnlmi_spmv_plan* plan, tnlmi_dtype x_type, const void* d_x, void* d_y,
                                      tnlmi_dtype compute_type, double alpha, double beta, int64_t begin, int64_t end,
                                      tnlmi_stream stream );

/* ------------------------------------------------------------- BLAS-1 */
/* (a, b) and lpNorm(a, 2) of FP64 device vectors into *d_result (device
 * memory, asynchronous; deterministic two-pass reduction).  Replace the
 * host-returning (a, b) / lpNorm of Containers/Expressions/
 * ExpressionTemplates.h:372-384, :442-460 (the C++ layer reads the scalar). */
tnlmi_status tnlmi_dot( int64_t n, const double* d_a, const double* d_b, double* d_result, tnlmi_stream stream );
tnlmi_status tnlmi_nrm2( int64_t n, const double* d_a, double* d_result, tnlmi_stream stream );
/* tnlmi_dot without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The same kernels, grid and
 * summation order as tnlmi_dot (bit-identical result); the first pass's
 * partial sums go to a static device array of the library (1024 FP64 =
 * 8 KiB per GPU, held until the process ends) instead of a workspace the
 * library allocates on first use.  Any n >= 0 fits (the first pass has at
 * most 1024 workgroups).  Calls that overlap on one GPU share that array and
 * race: serialise them (one stream, or events). */
tnlmi_status tnlmi_dot_noalloc( int64_t n, const double* d_a, const double* d_b, double* d_result,
                                tnlmi_stream stream );
/* y = alpha*x + beta*y (Containers/detail/VectorAssignment.h expressions) */
tnlmi_status tnlmi_axpby( int64_t n, double alpha, const double* d_x, double beta, double* d_y,
                          tnlmi_stream stream );

/* ------------------------------------------------------- linear solvers */
/* IterativeSolver settings (in) and results (out); defaults of
 * Solvers/IterativeSolver.h:226-235: max 1e9, min 0, convergence 1e-6,
 * divergence DBL_MAX. */
typedef struct {
   int64_t max_iterations;
   int64_t min_iterations;
   double convergence_residue;
   double divergence_residue;
   int64_t iterations; /* out: getIterations() */
   double residue;     /* out: getResidue() (the solver's own residue) */
   int32_t converged;  /* out: the bool LinearSolver::solve returns */
} tnlmi_solver_state;

typedef enum {
   TNLMI_PRECOND_NONE = 0,
   TNLMI_PRECOND_DIAGONAL = 1,
   TNLMI_PRECOND_ILU0 = 2,
   TNLMI_PRECOND_EXPLICIT = 3 /* x = M b with the given matrix as M (e.g. a sparse approximate inverse);
                                 the matrix must outlive the preconditioner.  No reference counterpart. */
} tnlmi_precond_kind;
typedef struct tnlmi_preconditioner tnlmi_preconditioner;

/* Preconditioner::update(matrix) -- Preconditioners/Diagonal.hpp:12-31
 * (d_i = A(i,i)) and ILU0.hpp:12-96 (IKJ incomplete LU on A's pattern,
 * factorised on the host as in the reference; the triangular solves run on the
 * GPU by dependency levels -- the reference's GPU ILU0 throws
 * NotImplementedError, ILU0.h:122-141).  Synchronous. */
tnlmi_status tnlmi_preconditioner_create( tnlmi_precond_kind kind, const tnlmi_matrix* A, tnlmi_stream stream,
                                          tnlmi_preconditioner** out );
void tnlmi_preconditioner_destroy( tnlmi_preconditioner* P );
/* Preconditioner::solve(b, x): x = M^{-1} b (asynchronous). */
tnlmi_status tnlmi_preconditioner_solve( const tnlmi_preconditioner* P, const double* d_b, double* d_x,
                                         tnlmi_stream stream );

/* BiCGStab (Solvers/Linear/BICGStab.hpp:33-127; exact_residue != 0: the
 * bicgstab-exact-residue option, BICGStab.hpp:14-31) and TFQMR
 * (Solvers/Linear/TFQMR.hpp:12-121), left-preconditioned by M (NULL: none),
 * with the reference's residues (||r|| / ||M^-1 b||, the exact preconditioned
 * residual, TFQMR's quasi-residual tau sqrt(it + 1) / ||M^-1 b||), its r*
 * restart and its stopping rule (IterativeSolver.hpp:98-143).  Optional plan
 * of A.  Synchronous. */
tnlmi_status tnlmi_bicgstab_solve( const tnlmi_matrix* A, const tnlmi_spmv_plan* plan, const tnlmi_preconditioner* M,
                                   int32_t exact_residue, const double* d_b, double* d_x, tnlmi_solver_state* state,
                                   tnlmi_stream stream );
/* BICGStabL::solve (Solvers/Linear/BICGStabL.hpp:37-257): BiCGStab(ell),
 * ell in [1, 32] ("bicgstab-ell", default 1), exact_residue as
 * "bicgstab-exact-residue"; one iteration per Bi-CG step.  A NaN initial
 * residue is TNLMI_RUNTIME_ERROR (the reference throws std::runtime_error). */
tnlmi_status tnlmi_bicgstabl_solve( const tnlmi_matrix* A, const tnlmi_spmv_plan* plan,
                                    const tnlmi_preconditioner* M, int32_t ell, int32_t exact_residue,
                                    const double* d_b, double* d_x, tnlmi_solver_state* state, tnlmi_stream stream );
/* IDRs::solve (Solvers/Linear/IDRs.hpp:47-386): IDR(s), s in [1, 32] ("idr-s",
 * default 4), smoothing as "idr-residual-smoothing"; the shadow space drawn as
 * the reference draws it.  A breakdown returns with converged = 0. */
tnlmi_status tnlmi_idrs_solve( const tnlmi_matrix* A, const tnlmi_spmv_plan* plan, const tnlmi_preconditioner* M,
                               int32_t s, int32_t smoothing, const double* d_b, double* d_x,
                               tnlmi_solver_state* state, tnlmi_stream stream );
tnlmi_status tnlmi_tfqmr_solve( const tnlmi_matrix* A, const tnlmi_spmv_plan* plan, const tnlmi_preconditioner* M,
                                const double* d_b, double* d_x, tnlmi_solver_state* state, tnlmi_stream stream );
/* LinearSolver<Matrix>::solve(b, x) -> bool (synchronous; FP64 vectors; any
 * matrix format and value type, FP64 accumulation).  plan: optional
 * column-split plan of A used for its products (NULL: A's own kernels).
 * CG: Solvers/Linear/CG.hpp:10-108.  GMRES: GMRES.hpp:71-806 (restarted,
 * Baker-Jessup-Kolev adaptive restart between restarting_min and _max,
 * restarting_max <= 31).  Jacobi: Jacobi.hpp:65-125. */
tnlmi_status tnlmi_cg_solve( const tnlmi_matrix* A, const tnlmi_spmv_plan* plan, const tnlmi_preconditioner* M,
                             const double* d_b, double* d_x, tnlmi_solver_state* state, tnlmi_stream stream );
/* CG's iteration (CG.hpp:10-108, no or the Diagonal preconditioner) stopped
 * on the residual ||b - A x|| / ||b|| <= convergence_residue instead of the
 * reference's preconditioned residue: the recurrence residual is tracked in
 * every iteration and confirmed by an FP64 b - A x at the stop (a drift
 * above the tolerance continues from x).  state->residue is that true
 * residual.  The stopping rule of tnlmi_mixed_cg_solve, for like-for-like
 * comparisons of FP64 and mixed-precision solves.  column_codes != 0: a
 * column-major Ellpack / SlicedEllpack A with at most 255 distinct offsets
 * column - row is read through 8-bit column codes, as the mixed solver reads
 * its inner matrix (identical results, fewer bytes).  No reference
 * counterpart. */
tnlmi_status tnlmi_cg_solve_true_residual( const tnlmi_matrix* A, const tnlmi_spmv_plan* plan,
                                           const tnlmi_preconditioner* M, int32_t column_codes, const double* d_b,
                                           double* d_x, tnlmi_solver_state* state, tnlmi_stream stream );
typedef enum {
   TNLMI_GMRES_CGS = 0,
   TNLMI_GMRES_CGSR = 1,
   TNLMI_GMRES_MGS = 2,
   TNLMI_GMRES_MGSR = 3,
   TNLMI_GMRES_CWY = 4
} tnlmi_gmres_variant;
tnlmi_status tnlmi_gmres_solve( const tnlmi_matrix* A, const tnlmi_spmv_plan* plan, const tnlmi_preconditioner* M,
                                tnlmi_gmres_variant variant, int32_t restarting_min, int32_t restarting_max,
                                int32_t step_min, int32_t step_max, const double* d_b, double* d_x,
                                tnlmi_solver_state* state, tnlmi_stream stream );
/* Mixed-precision GMRES: tnlmi_gmres_solve's iteration (same variants,
 * adaptive restarting, preconditioner M, stopping rule on the preconditioned
 * residue) whose Arnoldi products use A_lo -- the same matrix with FP32 or
 * BF16 values (FP64 vectors and accumulation) -- while every restart residual
 * M^-1 (b - A x) is computed with A in FP64: iterative refinement with one
 * GMRES cycle per correction.  The Krylov basis, the Hessenberg least squares
 * (Givens rotations, back substitution) and x stay FP64 on the GPU.  When a
 * cycle's estimate meets the rule, the FP64 residue of the updated x decides
 * whether the solve stops.  A holds FP64 values.  Synchronous.  Reduces to
 * GMRES.hpp:71-806 when A_lo = A; no reference counterpart otherwise (TNL has
 * no BF16 type and no mixed-precision solver). */
tnlmi_status tnlmi_gmres_mixed_solve( const tnlmi_matrix* A, const tnlmi_matrix* A_lo, const tnlmi_spmv_plan* plan,
                                      const tnlmi_preconditioner* M, tnlmi_gmres_variant variant,
                                      int32_t restarting_min, int32_t restarting_max, int32_t step_min,
                                      int32_t step_max, const double* d_b, double* d_x, tnlmi_solver_state* state,
                                      tnlmi_stream stream );
tnlmi_status tnlmi_jacobi_solve( const tnlmi_matrix* A, const tnlmi_spmv_plan* plan, double omega,
                                 int64_t residue_period, const double* d_b, double* d_x, tnlmi_solver_state* state,
                                 tnlmi_stream stream );
/* ||A x - b||_2 / ||b||_2 -- Solvers/Linear/Utils/LinearResidueGetter.hpp:11-23
 * (synchronous, result on the host). */
tnlmi_status tnlmi_true_residual( const tnlmi_matrix* A, const double* d_x, const double* d_b, double* residual,
                                  tnlmi_stream stream );

/* Mixed-precision CG: FP64 iterative refinement (r = b - A x, stop on the
 * true residual ||r||/||b|| <= state->convergence_residue, x += d) around a
 * Jacobi-preconditioned inner CG on A_lo -- the same matrix with FP32 or BF16
 * values -- with FP32 vectors and SpMV accumulation, run to the relative
 * tolerance inner_tolerance (at most inner_max_iterations per refinement
 * step).  state->iterations counts refinement steps; *inner_iterations (may be
 * NULL) the inner CG iterations.  Optional plans of A / A_lo.  Synchronous.
 * No reference counterpart (TNL has no BF16 type and no mixed-precision
 * solver); the inner iteration restates CG.hpp:10-108 in FP32. */
tnlmi_status tnlmi_ir_cg_solve( const tnlmi_matrix* A, const tnlmi_matrix* A_lo, const tnlmi_spmv_plan* plan,
                                const tnlmi_spmv_plan* plan_lo, double inner_tolerance,
                                int64_t inner_max_iterations, const double* d_b, double* d_x,
                                tnlmi_solver_state* state, int64_t* inner_iterations, tnlmi_stream stream );

/* Mixed-precision CG with reliable updates: Jacobi-preconditioned CG run in
 * FP32 (FP32 vectors and SpMV accumulation, FP64 dot products) on A_lo -- the
 * same matrix with FP32 or BF16 values -- with x accumulated in FP64.  When
 * the FP32 residual has fallen by `delta` (0 < delta < 1, e.g. 0.1) since the
 * last update, or its estimate reaches the tolerance, x += correction and
 * r = b - A x are recomputed in FP64 (the FP32 residual is replaced, the
 * search direction kept).  Stops when the FP64 true residual ||b-Ax||/||b||
 * is below state->convergence_residue (min/max iterations and divergence as
 * in IterativeSolver.hpp:98-121, checked at the updates); state->iterations
 * counts the FP32 iterations, *updates (may be NULL) the FP64 residual
 * evaluations.  Optional plan of A.  Synchronous.  No reference counterpart
 * (TNL has no BF16 type and no mixed-precision solver); the iteration
 * restates CG.hpp:10-108 in FP32. */
tnlmi_status tnlmi_mixed_cg_solve( const tnlmi_matrix* A, const tnlmi_matrix* A_lo, const tnlmi_spmv_plan* plan,
                                   double delta, const double* d_b, double* d_x, tnlmi_solver_state* state,
                                   int64_t* updates, tnlmi_stream stream );

/* ------------------------------------------- learned sparse approximate inverse */
/* A message-passing network over the graph of a square CSR matrix A (FP64
 * values, nonzero diagonal; an edge per stored entry, column -1 = padding)
 * that predicts M = D^-1 (I + O) on A's own pattern, a left preconditioner
 * (use M through TNLMI_PRECOND_EXPLICIT).  Hidden width 64, 3 layers: edge
 * update E = relu(E We + (H Wu)[col] + (H Wv)[row] + be), node update
 * H = relu(H Wh + mean_row(E) Wg + bh), decoder o = E wo + bo; the 64 x 64
 * layers are BF16 GEMMs on the matrix cores with FP32 accumulation.
 * Parameters: one FP32 vector of tnlmi_spai_parameter_count() entries
 * (encoder Ae[4x64], ae, An[4x64], an; per layer We, Wu, Wv, Wh, Wg [64x64,
 * row-major, y = x W], be, bh; wo[64], bo).  Training minimises the probe
 * loss sum_k ||z_k - M A z_k||^2 / (K n) with Adam.  No reference
 * counterpart (TNL has no learned preconditioner). */
typedef struct tnlmi_spai tnlmi_spai;
int64_t tnlmi_spai_parameter_count( void );
tnlmi_status tnlmi_spai_create( const float* h_parameters, tnlmi_spai** out );
void tnlmi_spai_destroy( tnlmi_spai* net );
/* Host copies of the parameters and of the last computed gradients (either may be NULL). */
tnlmi_status tnlmi_spai_get_parameters( const tnlmi_spai* net, float* h_parameters, float* h_gradients );
/* M's values on A's pattern (storage order of A) into d_M (FP64).  Synchronous. */
tnlmi_status tnlmi_spai_apply( const tnlmi_spai* net, const tnlmi_matrix* A, double* d_M, tnlmi_stream stream );
/* Probe loss for the K probes d_Z (K x n, FP64) and its gradients (kept in
 * the object); d_M (may be NULL) receives M.  Synchronous. */
tnlmi_status tnlmi_spai_loss_grad( tnlmi_spai* net, const tnlmi_matrix* A, const double* d_Z, int32_t K,
                                   double* loss, double* d_M, tnlmi_stream stream );
/* One training step on A: K Rademacher probes drawn from probe_seed, loss,
 * gradients, Adam (beta 0.9 / 0.999).  *loss = the loss before the step. */
tnlmi_status tnlmi_spai_train_step( tnlmi_spai* net, const tnlmi_matrix* A, int32_t K, uint64_t probe_seed,
                                    float learning_rate, double* loss, tnlmi_stream stream );
/* Y[rows x 64] (FP32) = X[rows x 64] W[64 x 64] with BF16 X and W (row-major)
 * on the matrix cores -- the dense layer of the network.  Synchronous. */
tnlmi_status tnlmi_bf16_gemm64( int64_t rows, const uint16_t* d_X, const uint16_t* d_W, float* d_Y,
                                tnlmi_stream stream );

/* ------------------------------------------------------------ batched CG */
/* Many independent SPD systems, the diagonal blocks of one block-diagonal
 * matrix A (FP64 values, any segment format; system k owns rows
 * sys_rows[k] .. sys_rows[k+1], a host array of nsys+1 entries; every
 * stored column of a system's rows lies in its own row range, checked at
 * create: TNLMI_INVALID_ARGUMENT otherwise), solved by one persistent kernel:
 * each workgroup of block_size threads (128/256/512/1024; 0 = 512), with
 * lanes_per_row lanes summing each row (1/2/4/8; 0 = 1: 64, 32, 16 or 8 rows
 * per wavefront), takes the next system from an atomic counter and runs
 * Jacobi-preconditioned CG (CG.hpp:10-108 per system, TNL's stopping rule on
 * the preconditioned residue, default min/divergence settings) to
 * `tolerance`, at most max_iterations.  A_lo (optional): an FP32 copy of A on
 * the same segments for the mixed variant -- FP64 refinement on the true
 * residual ||b-Ax||/||b|| around an FP32 Jacobi-CG run to inner_tolerance.
 * d_iterations / d_residues (device, nsys entries) receive per-system
 * iterations (mixed: inner iterations) and final residues.  use_graph: the
 * launch (counter reset + solve) is captured in a hipGraph once per argument
 * set and replayed.  Asynchronous.  Calls with one object must not overlap.
 * No reference counterpart (TNL solves one system per solver call). */
typedef struct tnlmi_batched_cg tnlmi_batched_cg;
tnlmi_status tnlmi_batched_cg_create( const tnlmi_matrix* A, const tnlmi_matrix* A_lo, const int64_t* sys_rows,
                                      int64_t nsys, int32_t block_size, int32_t lanes_per_row, tnlmi_stream stream,
                                      tnlmi_batched_cg** out );
void tnlmi_batched_cg_destroy( tnlmi_batched_cg* batch );
tnlmi_status tnlmi_batched_cg_solve( tnlmi_batched_cg* batch, int32_t mixed, const double* d_b, double* d_x,
                                     double tolerance, int64_t max_iterations, double inner_tolerance,
                                     int64_t* d_iterations, double* d_residues, int32_t use_graph,
                                     tnlmi_stream stream );

/* ------------------------------------------------- input generators (host) */
/* Seeded host-side generators of the workloads the engine is measured on.
 * Random numbers come from a counter-based generator (splitmix64 of seed,
 * stream and index), so every entry is reproducible independently of the
 * thread count.  No reference counterpart (the reference tests build their
 * matrices inline). */

/* 2D 5-point Laplacian on an nx x ny grid, Dirichlet boundary: diagonal 4,
 * neighbours -1, rows lexicographic (ix fastest), columns ascending.  Pass
 * NULL arrays to get the nonzero count only.  Returns nnz. */
int64_t tnlmi_gen_laplace2d( int64_t nx, int64_t ny, int32_t* offsets, int32_t* columns, double* values );
/* 3D 7-point Laplacian on nx*ny*nz, diagonal 6, neighbours -1. */
int64_t tnlmi_gen_poisson3d( int64_t nx, int64_t ny, int64_t nz, int32_t* offsets, int32_t* columns, double* values );
/* Row lengths ~ Zipf(s) on [1, kmax], scaled to mean `mean` and clipped to
 * [lo, hi]; the scale is bisected and the remaining difference is spread by
 * +-1 in row order so that the sum is exactly round(mean * n). */
void tnlmi_gen_zipf_sizes( int64_t n, double s, int32_t kmax, double mean, int32_t lo, int32_t hi, uint64_t seed,
                           int32_t* sizes );
/* Sorted column indexes drawn uniformly from [0, ncols) without repeats in a
 * row, for CSR offsets `offsets` (n+1 entries). */
void tnlmi_gen_random_columns( int64_t n, int64_t ncols, const int32_t* offsets, uint64_t seed, int32_t* columns );
/* out[i] ~ U[lo, hi) or N(0, 1) for i < n, from (seed, stream). */
/* 2D convection-diffusion, m x m interior grid of the unit square, Dirichlet
 * boundary, 5-point upwind finite differences, rows scaled by h^2/eps
 * (diagonal 4 + h Pe (|vx| + |vy|), neighbours -1 - h Pe max(+-v, 0)),
 * lexicographic rows, sorted columns; 5m^2 - 4m nonzeros.  Velocity: the
 * curl of a stream function of 8 random sine modes (divergence-free, zero
 * normal flow on the boundary) scaled to max |v| = 1; Peclet number
 * Pe ~ U[1, 100].  Returns Pe; with offsets == NULL only returns it. */
double tnlmi_gen_convdiff( int64_t m, uint64_t seed, int32_t* offsets, int32_t* columns, double* values );
void tnlmi_gen_uniform( int64_t n, double lo, double hi, uint64_t seed, uint64_t stream, double* out );
void tnlmi_gen_normal( int64_t n, uint64_t seed, uint64_t stream, double* out );
/* Batch of 2D 5-point variable-coefficient diffusion systems as one
 * block-diagonal CSR matrix: system k on a side s_k grid, s_k =
 * round(exp(U[ln smin, ln smax])), face coefficients ~ LogNormal(0, 1) on all
 * cell faces (Dirichlet boundary faces add to the diagonal), A_ii = sum of the
 * cell's 4 face coefficients, A_ij = -(shared face coefficient).  Pass NULL
 * offsets to get sides[nsys], *rows and *nnz only. */
void tnlmi_gen_batched_diffusion( int64_t nsys, int32_t smin, int32_t smax, uint64_t seed, int32_t* sides,
                                  int64_t* rows, int64_t* nnz, int32_t* offsets, int32_t* columns, double* values );

#ifdef __cplusplus
}
#endif

#endif /* TNLMI_H */
