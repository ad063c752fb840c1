"""tnl_amd -- MI355X (gfx950) sparse matrix-vector product and Krylov solvers.

A Python mirror of the reference's operator interface for this path, over the
C ABI of ``libtnlmi.so`` (include/tnlmi.h):

* ``Segments``      -- Algorithms::Segments::<Format>, built on the GPU from
                       segment sizes (``setSegmentsSizes``).
* ``SparseMatrix``  -- Matrices::SparseMatrix<Real, GPU, int, GeneralMatrix,
                       Segments>: ``vectorProduct(x, y, alpha, beta, begin,
                       end, launch)``.
* generators        -- the seeded workloads the engine is measured on.

Device memory and streams come from PyTorch (plumbing only); every operation
runs in the library's own HIP kernels.  Errors map to the reference's
exception types (see ``_lib``).
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import (InvalidArgument, LogicError, NotImplementedError_, OverflowError_, TnlmiError,  # noqa: F401
                   BackendError, Launch)

# formats (tnlmi_format)
CSR, ADAPTIVE_CSR, ELLPACK, SLICED_ELLPACK, BIELLPACK, CHUNKED_ELLPACK, SORTED_CSR = range(7)
FORMAT_NAMES = {CSR: "CSR", ADAPTIVE_CSR: "AdaptiveCSR", ELLPACK: "Ellpack", SLICED_ELLPACK: "SlicedEllpack",
                BIELLPACK: "BiEllpack", CHUNKED_ELLPACK: "ChunkedEllpack", SORTED_CSR: "SortedCSR"}
# ElementsOrganization
COLUMN_MAJOR, ROW_MAJOR = 0, 1
# dtypes (tnlmi_dtype)
F64, F32, BF16 = 0, 1, 2
# ThreadsToSegmentsMapping (+ AUTO)
FIXED, WARP, BLOCK, BLOCK_MERGED, DYNAMIC_GROUPING, AUTO = 0, 1, 2, 3, 4, 255

# segments arrays (tnlmi_segments_array)
ARR = dict(offsets=0, slice_offsets=1, slice_segment_sizes=2, permutation=3, group_pointers=4,
           inverse_permutation=5, chunk_slices=6, segment_to_chunk=7, segment_to_slice=8, segment_pointers=9,
           chunks_to_segments=10, adaptive_blocks=11, sizes=12, tile_rows=13)


def _torch():
    import torch
    return torch


def dtype_code(t):
    torch = _torch()
    if t == torch.float64:
        return F64
    if t == torch.float32:
        return F32
    if t == torch.bfloat16:
        return BF16
    raise InvalidArgument(_lib.INVALID_ARGUMENT, f"unsupported dtype {t}")


def torch_dtype(code):
    torch = _torch()
    return {F64: torch.float64, F32: torch.float32, BF16: torch.bfloat16}[code]


def _stream(stream=None):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def version():
    return _lib.lib().tnlmi_version().decode()


class Segments:
    """Algorithms::Segments::<Format> on the GPU (setSegmentsSizes on create)."""

    def __init__(self, fmt, sizes, organization=COLUMN_MAJOR, param=0, stream=None):
        torch = _torch()
        sizes = torch.as_tensor(sizes, dtype=torch.int32)
        if not sizes.is_cuda:
            sizes = sizes.cuda()
        sizes = sizes.contiguous()
        self._sizes = sizes
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().tnlmi_segments_create(fmt, organization, param, ctypes.c_void_p(sizes.data_ptr()),
                                                    sizes.numel(), _stream(stream), ctypes.byref(h)))
        self.handle = h
        self.format = fmt
        self.organization = organization

    @classmethod
    def _adopt(cls, handle, fmt, organization):
        """A segments object around a handle the library built (it is destroyed with it)."""
        obj = cls.__new__(cls)
        obj._sizes = None
        obj.handle = handle
        obj.format = fmt
        obj.organization = organization
        return obj

    def __del__(self):
        h = getattr(self, "handle", None)
        try:
            if h is not None and h.value and _lib._lib is not None:
                _lib.lib().tnlmi_segments_destroy(h)
        except (AttributeError, TypeError):  # interpreter shutdown
            pass
        self.handle = None

    def getSegmentCount(self):
        return _lib.lib().tnlmi_segments_count(self.handle)

    def getStorageSize(self):
        return _lib.lib().tnlmi_segments_storage_size(self.handle)

    def getElementCount(self):
        return _lib.lib().tnlmi_segments_element_count(self.handle)

    def array(self, name):
        """Copy of one of the format's index arrays (numpy int32)."""
        p = ctypes.c_void_p()
        n = ctypes.c_int64()
        _lib.check(_lib.lib().tnlmi_segments_get_array(self.handle, ARR[name], ctypes.byref(p), ctypes.byref(n)))
        return _d2h(p.value, n.value, np.int32)

    def ellpack_dims(self):
        w = ctypes.c_int32()
        a = ctypes.c_int32()
        _lib.check(_lib.lib().tnlmi_segments_ellpack_dims(self.handle, ctypes.byref(w), ctypes.byref(a)))
        return w.value, a.value

    def adaptive_blocks(self, value_bytes):
        p = ctypes.c_void_p()
        n = ctypes.c_int64()
        _lib.check(_lib.lib().tnlmi_segments_adaptive_blocks(self.handle, value_bytes, ctypes.byref(p),
                                                             ctypes.byref(n)))
        return _d2h(p.value, 6 * n.value, np.int32).reshape(-1, 6)

    def global_indexes(self):
        torch = _torch()
        out = torch.empty(max(1, self.getElementCount()), dtype=torch.int32, device="cuda")
        _lib.check(_lib.lib().tnlmi_segments_global_indexes(self.handle, ctypes.c_void_p(out.data_ptr()),
                                                            _stream()))
        return out[: self.getElementCount()]


_hip = None


def _d2h(ptr, count, dtype):
    """Synchronous copy of `count` elements at device address `ptr` to numpy."""
    global _hip
    out = np.zeros(count, dtype=dtype)
    if count == 0:
        return out
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.restype = ctypes.c_int
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.restype = ctypes.c_int
    _hip.hipDeviceSynchronize()
    rc = _hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(ptr), ctypes.c_size_t(out.nbytes), 2)
    if rc != 0:
        raise BackendError(_lib.BACKEND_ERROR, f"hipMemcpy failed ({rc})")
    return out


class SparseMatrix:
    """Matrices::SparseMatrix<Real, GPU, int, MatrixType, Segments> over device tensors.
    symmetric=True is SymmetricMatrix: the storage holds the lower triangle and
    the diagonal."""

    def __init__(self, rows, columns, segments, values, column_indexes, symmetric=False):
        self.rows = int(rows)
        self.columns = int(columns)
        self.segments = segments
        self.values = values
        self.column_indexes = column_indexes
        self.symmetric = bool(symmetric)
        self._m = _lib.Matrix(self.rows, self.columns, segments.handle.value, dtype_code(values.dtype),
                              values.data_ptr(), column_indexes.data_ptr(), int(self.symmetric))
        # arrays handed in by the caller: no product plan of earlier arrays at
        # these addresses may serve them
        self.modified()

    def modified(self):
        """Drops what the library derived from the value and column-index
        arrays (the automatic product plan of vectorProduct); call after
        writing those tensors directly (tnlmi_matrix_modified)."""
        _lib.check(_lib.lib().tnlmi_matrix_modified(ctypes.byref(self._m)))

    @classmethod
    def from_csr(cls, rows, columns, offsets, cols, vals, fmt=CSR, value_dtype=None, organization=COLUMN_MAJOR,
                 param=0, capacities=None, symmetric=False):
        """Assemble from CSR arrays (any device): row r holds cols/vals[offsets[r]:offsets[r+1]].

        capacities (optional) are the row capacities (setRowCapacities);
        by default the CSR row lengths.  symmetric=True: a SymmetricMatrix from
        its lower triangle with the diagonal (setElements with the
        SymmetricLower encoding, SparseMatrix.hpp:437-438: an entry above the
        diagonal is an error).
        """
        torch = _torch()
        offsets = torch.as_tensor(offsets, dtype=torch.int32).cuda().contiguous()
        cols = torch.as_tensor(cols, dtype=torch.int32).cuda().contiguous()
        vals = torch.as_tensor(vals).cuda().contiguous()
        if symmetric:
            if rows != columns:
                raise InvalidArgument(_lib.INVALID_ARGUMENT, "a symmetric matrix must be square")
            lens = (offsets[1:] - offsets[:-1]).long()
            row_of = torch.repeat_interleave(torch.arange(rows, device="cuda", dtype=torch.int32), lens)
            if bool((cols > row_of).any()):
                raise LogicError(_lib.LOGIC_ERROR, "Only lower part of the symmetric matrix is expected.")
        if capacities is None:
            capacities = offsets[1:] - offsets[:-1]
        seg = Segments(fmt, capacities, organization, param)
        vdt = value_dtype if value_dtype is not None else vals.dtype
        storage = max(1, seg.getStorageSize())
        values = torch.empty(storage, dtype=vdt, device="cuda")
        colidx = torch.empty(storage, dtype=torch.int32, device="cuda")
        if vals.dtype not in (torch.float64, torch.float32):
            vals = vals.double()
        _lib.check(_lib.lib().tnlmi_matrix_fill(seg.handle, ctypes.c_void_p(offsets.data_ptr()),
                                                ctypes.c_void_p(cols.data_ptr()), ctypes.c_void_p(vals.data_ptr()),
                                                dtype_code(vals.dtype), ctypes.c_void_p(values.data_ptr()),
                                                dtype_code(vdt), ctypes.c_void_p(colidx.data_ptr()), _stream()))
        return cls(rows, columns, seg, values, colidx, symmetric=symmetric)

    def rowLengths(self):
        """The leading entries of each row with a column below getColumns() (padding
        slots, column -1, included): copySparseMatrix's row capacities
        (tnlmi_matrix_row_lengths), an int32 device tensor."""
        torch = _torch()
        out = torch.empty(max(1, self.rows), dtype=torch.int32, device="cuda")
        _lib.check(_lib.lib().tnlmi_matrix_row_lengths(ctypes.byref(self._m), ctypes.c_void_p(out.data_ptr()),
                                                       _stream()))
        return out[:self.rows]

    def convert(self, fmt, param=0, organization=COLUMN_MAJOR):
        """copySparseMatrix into a new matrix of another format (Matrices/SparseOperations.hpp:
        600-690): same rows, values and columns, on the GPU (tnlmi_matrix_copy)."""
        torch = _torch()
        lengths = self.rowLengths()
        seg = Segments(fmt, lengths, organization, param)
        storage = max(1, seg.getStorageSize())
        values = torch.empty(storage, dtype=self.values.dtype, device="cuda")
        colidx = torch.empty(storage, dtype=torch.int32, device="cuda")
        B = SparseMatrix(self.rows, self.columns, seg, values, colidx, symmetric=self.symmetric)
        _lib.check(_lib.lib().tnlmi_matrix_copy(ctypes.byref(self._m), ctypes.byref(B._m),
                                                ctypes.c_void_p(lengths.data_ptr()), _stream()))
        return B

    @classmethod
    def from_mtx(cls, path, fmt=CSR, value_dtype=None, organization=COLUMN_MAJOR, param=0):
        """MatrixReader::readMtx into this format (tnlmi_mtx_read, then assembly on the GPU)."""
        rows, columns, off, cols, vals = read_mtx(path)
        return cls.from_csr(rows, columns, off, cols, vals, fmt=fmt, value_dtype=value_dtype,
                            organization=organization, param=param)

    def getSerializationType(self):
        """SparseMatrixBase::getSerializationType (Matrices/SparseMatrixBase.hpp:46-53)."""
        n = ctypes.c_int64()
        _lib.check(_lib.lib().tnlmi_matrix_serialization_type(ctypes.byref(self._m), None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value + 1)
        _lib.check(_lib.lib().tnlmi_matrix_serialization_type(ctypes.byref(self._m), buf, n.value + 1,
                                                              ctypes.byref(n)))
        return buf.value.decode()

    def save(self, path):
        """SparseMatrix::save: TNL's binary format (tnlmi_matrix_save)."""
        _lib.check(_lib.lib().tnlmi_matrix_save(ctypes.byref(self._m), os.fsencode(path), _stream()))

    @classmethod
    def load(cls, path, expected_type=None):
        """SparseMatrix::load of a TNL binary file (tnlmi_matrix_file_*): the format,
        organization and value type the file names.  expected_type (a
        getSerializationType() string) is checked as TNL's load checks its own
        type (SparseMatrix.hpp:676-682)."""
        torch = _torch()
        L = _lib.lib()
        f = ctypes.c_void_p()
        _lib.check(L.tnlmi_matrix_file_read(os.fsencode(path), ctypes.byref(f)))
        try:
            found = L.tnlmi_matrix_file_type(f).decode()
            if expected_type is not None and found != expected_type:
                raise _lib.TnlmiError(_lib.RUNTIME_ERROR, f"Failed to deserialize an object from the file '{path}': "
                                                          f"object type does not match (expected {expected_type}, "
                                                          f"found {found}).")
            rows, cols, storage = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
            fmt, org, vt, mt = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            param = ctypes.c_int32()
            _lib.check(L.tnlmi_matrix_file_info(f, ctypes.byref(rows), ctypes.byref(cols), ctypes.byref(fmt),
                                                ctypes.byref(org), ctypes.byref(param), ctypes.byref(vt),
                                                ctypes.byref(storage), ctypes.byref(mt)))
            h = ctypes.c_void_p()
            _lib.check(L.tnlmi_matrix_file_segments(f, _stream(), ctypes.byref(h)))
            seg = Segments._adopt(h, fmt.value, org.value)
            n = max(1, storage.value)
            values = torch.empty(n, dtype=torch_dtype(vt.value), device="cuda")
            colidx = torch.empty(n, dtype=torch.int32, device="cuda")
            _lib.check(L.tnlmi_matrix_file_upload(f, ctypes.c_void_p(values.data_ptr()),
                                                  ctypes.c_void_p(colidx.data_ptr()), _stream()))
        finally:
            L.tnlmi_matrix_file_destroy(f)
        return cls(rows.value, cols.value, seg, values, colidx, symmetric=mt.value == 1)

    def transposedVectorProduct(self, x, y, alpha=1.0, beta=0.0, begin=0, end=0, stream=None):
        """y = alpha*A^T*x + beta*y over columns [begin, end) (SparseMatrixBase::transposedVectorProduct)."""
        if x.numel() != self.rows:
            raise InvalidArgument(_lib.INVALID_ARGUMENT, "transposedVectorProduct: size of the input vector does not "
                                                         "match the number of matrix rows")
        if y.numel() != self.columns:
            raise InvalidArgument(_lib.INVALID_ARGUMENT, "transposedVectorProduct: size of the output vector does not "
                                                         "match the number of matrix columns")
        if x.dtype != y.dtype:
            raise InvalidArgument(_lib.INVALID_ARGUMENT, "transposedVectorProduct: x and y must have the same dtype")
        _lib.check(_lib.lib().tnlmi_spmv_transposed(ctypes.byref(self._m), dtype_code(x.dtype),
                                                    ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                                    float(alpha), float(beta), int(begin), int(end), _stream(stream)))
        return y

    def setElement(self, row, column, value, stream=None):
        """SparseMatrixBase::setElement (SparseMatrixBase.hpp:152-161): addElement with multiplicator 0."""
        self.addElement(row, column, value, 0.0, stream)

    def addElement(self, row, column, value, thisElementMultiplicator=1.0, stream=None):
        """SparseMatrixBase::addElement (SparseMatrixBase.hpp:166-231): a stored element becomes
        thisElementMultiplicator * old + value, a new one takes the row's first padding slot or the
        slot of the first larger column (later slots shift right); LogicError when the row is full."""
        _lib.check(_lib.lib().tnlmi_matrix_add_element(ctypes.byref(self._m), int(row), int(column), float(value),
                                                       float(thisElementMultiplicator), _stream(stream)))

    def addElements(self, rows, columns, values, thisElementMultiplicators=None, stream=None):
        """Many addElement operations in one launch (tnlmi_matrix_add_elements): operation i is
        addElement(rows[i], columns[i], values[i], thisElementMultiplicators[i]); None for the
        multiplicators gives setElement (multiplicator 0).  The matrix ends as after the operations
        one by one in list order; LogicError when a row is full (that row stops, other rows' operations
        are applied).  rows / columns / values / multiplicators: sequences, numpy arrays or tensors."""
        torch = _torch()

        def dev(a, dtype):
            t = a if isinstance(a, torch.Tensor) else torch.as_tensor(np.asarray(a))
            return t.to(device="cuda", dtype=dtype).contiguous()
        r, c, v = dev(rows, torch.int64), dev(columns, torch.int64), dev(values, torch.float64)
        if not (r.numel() == c.numel() == v.numel()):
            raise InvalidArgument(_lib.INVALID_ARGUMENT, "addElements: rows, columns and values differ in length")
        m = None
        if thisElementMultiplicators is not None:
            m = dev(thisElementMultiplicators, torch.float64)
            if m.numel() != r.numel():
                raise InvalidArgument(_lib.INVALID_ARGUMENT, "addElements: multiplicators differ in length")
        failed = ctypes.c_int64(-1)
        _lib.check(_lib.lib().tnlmi_matrix_add_elements(
            ctypes.byref(self._m), r.numel(), ctypes.c_void_p(r.data_ptr()), ctypes.c_void_p(c.data_ptr()),
            ctypes.c_void_p(v.data_ptr()), ctypes.c_void_p(m.data_ptr()) if m is not None else None,
            ctypes.byref(failed), _stream(stream)))

    def setElementsBatch(self, rows, columns, values, stream=None):
        """setElement for each (rows[i], columns[i], values[i]) in one launch (addElements with
        multiplicator 0)."""
        self.addElements(rows, columns, values, None, stream)

    def _row_counts(self, kind, stream=None):
        torch = _torch()
        out = torch.empty(max(1, self.rows), dtype=torch.int32, device="cuda")
        _lib.check(_lib.lib().tnlmi_matrix_row_counts(ctypes.byref(self._m), int(kind),
                                                      ctypes.c_void_p(out.data_ptr()), _stream(stream)))
        return out[:self.rows]

    def getRowCapacity(self, row):
        """SparseMatrixBase::getRowCapacity: the row's segment size (Ellpack: the width)."""
        return int(self._row_counts(0)[int(row)])

    def getRowCapacities(self):
        """SparseMatrixBase::getRowCapacities (SparseMatrixBase.hpp:75-92): stored slots per row."""
        return self._row_counts(1)

    def getCompressedRowLengths(self):
        """SparseMatrixBase::getCompressedRowLengths (SparseMatrixBase.hpp:55-73): stored nonzeros per row."""
        return self._row_counts(2)

    def getNonzeroElementsCount(self):
        """SparseMatrixBase::getNonzeroElementsCount (SparseMatrixBase.hpp:102-132); a symmetric
        matrix counts its stored off-diagonal entries twice."""
        return int(self._row_counts(3 if self.symmetric else 2).long().sum())

    def getElement(self, row, column, stream=None):
        """SparseMatrixBase::getElement (SparseMatrixBase.hpp:231-265): 0.0 if not stored."""
        out = ctypes.c_double(0.0)
        _lib.check(_lib.lib().tnlmi_matrix_get_element(ctypes.byref(self._m), int(row), int(column),
                                                       ctypes.byref(out), _stream(stream)))
        return out.value

    def getRows(self):
        return self.rows

    def getColumns(self):
        return self.columns

    def vectorProduct(self, x, y, alpha=1.0, beta=0.0, begin=0, end=0, launch=None, compute_dtype=None,
                      stream=None):
        """y = alpha*A*x + beta*y over rows [begin, end) (SparseMatrixBase::vectorProduct)."""
        if x.numel() != self.columns:
            raise InvalidArgument(_lib.INVALID_ARGUMENT, "vectorProduct: size of the input vector does not match "
                                                         "the number of matrix columns")
        if y.numel() != self.rows:
            raise InvalidArgument(_lib.INVALID_ARGUMENT, "vectorProduct: size of the output vector does not match "
                                                         "the number of matrix rows")
        if x.dtype != y.dtype:
            raise InvalidArgument(_lib.INVALID_ARGUMENT, "vectorProduct: x and y must have the same dtype")
        xt = dtype_code(x.dtype)
        ct = dtype_code(compute_dtype) if compute_dtype is not None else xt
        L = None
        if launch is not None:
            mapping, tps = launch
            L = ctypes.byref(Launch(mapping, tps))
        _lib.check(_lib.lib().tnlmi_spmv(ctypes.byref(self._m), xt, ctypes.c_void_p(x.data_ptr()),
                                         ctypes.c_void_p(y.data_ptr()), ct, float(alpha), float(beta), int(begin),
                                         int(end), L, _stream(stream)))
        return y


class SpmvPlan:
    """Execution plan of a CSR SparseMatrix's vectorProduct (tnlmi_spmv_plan):
    parts = 8 or 16, the XCD column-split plan; parts = 0, the windowed plan
    (x staged in LDS per 8192-column window, rows of at most 256 nonzeros
    summed bit for bit as the host CSR loop)."""

    def __init__(self, A, parts=8, stream=None):
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().tnlmi_spmv_plan_create(ctypes.byref(A._m), parts, _stream(stream), ctypes.byref(h)))
        self.handle = h
        self.rows, self.columns = A.rows, A.columns

    def __del__(self):
        h = getattr(self, "handle", None)
        try:
            if h is not None and h.value and _lib._lib is not None:
                _lib.lib().tnlmi_spmv_plan_destroy(h)
        except (AttributeError, TypeError):  # interpreter shutdown
            pass
        self.handle = None

    def vectorProduct(self, x, y, alpha=1.0, beta=0.0, begin=0, end=0, compute_dtype=None, stream=None):
        if x.numel() != self.columns or y.numel() != self.rows or x.dtype != y.dtype:
            raise InvalidArgument(_lib.INVALID_ARGUMENT, "vectorProduct: vector sizes/dtypes do not match the matrix")
        xt = dtype_code(x.dtype)
        ct = dtype_code(compute_dtype) if compute_dtype is not None else xt
        _lib.check(_lib.lib().tnlmi_spmv_plan_execute(self.handle, xt, ctypes.c_void_p(x.data_ptr()),
                                                      ctypes.c_void_p(y.data_ptr()), ct, float(alpha), float(beta),
                                                      int(begin), int(end), _stream(stream)))
        return y


# ------------------------------------------------------------------ generators
def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _host_csr(h):
    """(rows, columns, offsets, columns, values) numpy copies of a tnlmi_host_csr, destroyed."""
    import numpy as np
    try:
        r, c, z = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        po, pc, pv = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        _lib.check(_lib.lib().tnlmi_host_csr_get(h, ctypes.byref(r), ctypes.byref(c), ctypes.byref(z), ctypes.byref(po),
                                                 ctypes.byref(pc), ctypes.byref(pv)))
        n, nnz = r.value, z.value

        def arr(p, count, ct, dt):
            if count == 0:
                return np.zeros(0, dt)
            return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ct)), shape=(count,)).copy()
        return (n, c.value, arr(po, n + 1, ctypes.c_int32, np.int32), arr(pc, nnz, ctypes.c_int32, np.int32),
                arr(pv, nnz, ctypes.c_double, np.float64))
    finally:
        _lib.lib().tnlmi_host_csr_destroy(h)


def read_mtx(path):
    """MatrixReader::readMtx (MatrixMarket coordinate file) as host CSR arrays:
    (rows, columns, offsets, columns, values)."""
    h = ctypes.c_void_p()
    _lib.check(_lib.lib().tnlmi_mtx_read(os.fsencode(path), ctypes.byref(h)))
    return _host_csr(h)


def parse_mtx(text):
    """read_mtx of a MatrixMarket file's text (str or bytes)."""
    data = text.encode() if isinstance(text, str) else bytes(text)
    h = ctypes.c_void_p()
    _lib.check(_lib.lib().tnlmi_mtx_parse(data, len(data), ctypes.byref(h)))
    return _host_csr(h)


def laplace2d(nx, ny):
    """2D 5-point Laplacian (offsets, columns, values) as numpy arrays."""
    L = _lib.lib()
    nnz = L.tnlmi_gen_laplace2d(nx, ny, None, None, None)
    off = np.zeros(nx * ny + 1, np.int32)
    col = np.zeros(nnz, np.int32)
    val = np.zeros(nnz, np.float64)
    L.tnlmi_gen_laplace2d(nx, ny, _ptr(off), _ptr(col), _ptr(val))
    return off, col, val


def poisson3d(nx, ny, nz):
    L = _lib.lib()
    nnz = L.tnlmi_gen_poisson3d(nx, ny, nz, None, None, None)
    off = np.zeros(nx * ny * nz + 1, np.int32)
    col = np.zeros(nnz, np.int32)
    val = np.zeros(nnz, np.float64)
    L.tnlmi_gen_poisson3d(nx, ny, nz, _ptr(off), _ptr(col), _ptr(val))
    return off, col, val


def convdiff(m, seed=0):
    """2D upwind convection-diffusion on an m x m grid (tnlmi_gen_convdiff):
    (offsets, columns, values, peclet)."""
    L = _lib.lib()
    n = m * m
    nnz = 5 * n - 4 * m
    off = np.zeros(n + 1, np.int32)
    col = np.zeros(nnz, np.int32)
    val = np.zeros(nnz, np.float64)
    pe = L.tnlmi_gen_convdiff(m, seed, _ptr(off), _ptr(col), _ptr(val))
    return off, col, val, pe


def zipf_csr(n, s=2.0, kmax=4096, mean=16.0, seed=0):
    """The power-law matrix: n x n, row lengths ~ Zipf(s) on [1, kmax] scaled to
    mean `mean` (clipped to [1, kmax]), uniform distinct sorted columns, N(0,1)
    values."""
    L = _lib.lib()
    sizes = np.zeros(n, np.int32)
    L.tnlmi_gen_zipf_sizes(n, s, kmax, mean, 1, kmax, seed, _ptr(sizes))
    off = np.zeros(n + 1, np.int32)
    off[1:] = np.cumsum(sizes, dtype=np.int64)
    col = np.zeros(int(off[-1]), np.int32)
    L.tnlmi_gen_random_columns(n, n, _ptr(off), seed, _ptr(col))
    val = np.zeros(int(off[-1]), np.float64)
    L.tnlmi_gen_normal(len(val), seed, 7, _ptr(val))
    return off, col, val


def uniform(n, lo=-1.0, hi=1.0, seed=0, stream=11):
    out = np.zeros(n, np.float64)
    _lib.lib().tnlmi_gen_uniform(n, lo, hi, seed, stream, _ptr(out))
    return out


def normal(n, seed=0, stream=13):
    out = np.zeros(n, np.float64)
    _lib.lib().tnlmi_gen_normal(n, seed, stream, _ptr(out))
    return out
