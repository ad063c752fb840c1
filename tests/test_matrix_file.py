"""SparseMatrix::save / load in TNL's binary format (tnlmi_matrix_save,
tnlmi_matrix_file_*; Matrices/SparseMatrix.hpp:642-694).

The reference's own save and load (oracle/_ref: TNL's SparseMatrix of each
format, built from the same elements as SparseMatrix::setRowCapacities +
setElement) are the oracle: a file TNL writes must load here into the storage
TNL holds, a file written here must load in TNL into that storage, and the
bytes written here must be TNL's (except where TNL writes memory it never
initialised -- ChunkedEllpack slice entries past the last slice -- and the
AdaptiveCSR kernel descriptors, which this library builds for wave64).

CPU tests: file parsing and its error paths (no device call).  GPU tests:
save, load, SpMV of the loaded matrix."""
import ctypes
import os
import struct
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import tnl_amd as tnl  # noqa: E402
from tnl_amd import _lib  # noqa: E402
import oracle_lib as OL  # noqa: E402

# (reference format name, our format, param)
FORMATS = [("CSR", tnl.CSR, 0), ("AdaptiveCSR", tnl.ADAPTIVE_CSR, 0), ("Ellpack", tnl.ELLPACK, 0),
           ("SlicedEllpack32", tnl.SLICED_ELLPACK, 32), ("SlicedEllpack64", tnl.SLICED_ELLPACK, 64),
           ("BiEllpack", tnl.BIELLPACK, 32), ("ChunkedEllpack", tnl.CHUNKED_ELLPACK, 0),
           ("SortedCSR", tnl.SORTED_CSR, 0)]
CASES = [(f, org) for f in FORMATS for org in (tnl.COLUMN_MAJOR, tnl.ROW_MAJOR)
         if org == tnl.COLUMN_MAJOR or f[1] in (tnl.ELLPACK, tnl.SLICED_ELLPACK, tnl.BIELLPACK, tnl.CHUNKED_ELLPACK)]
IDS = [f"{f[0]}-{'cm' if org == tnl.COLUMN_MAJOR else 'rm'}" for f, org in CASES]
BLOCK_TYPE = b"TNL::Containers::Array< TNL::Algorithms::Segments::detail::CSRAdaptiveKernelBlockDescriptor<int> >"
SLICE_TYPE = b"TNL::Containers::Array< TNL::Algorithms::Segments::detail::ChunkedEllpackSliceInfo<int> >"


def ref_or_skip():
    ref = OL.Reference.load()
    if ref is None:
        pytest.skip("oracle/_ref is not built")
    return ref


def matrix_csr(seed, n=300, m=350, maxlen=24, long_rows=((7, 200),)):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(0, maxlen + 1, n).astype(np.int32)
    for r, L in long_rows:
        sizes[r] = L
    off = np.zeros(n + 1, np.int32)
    off[1:] = np.cumsum(sizes)
    cols = np.full(off[-1], -1, np.int32)
    vals = np.zeros(off[-1])
    for r in range(n):
        used = int(sizes[r]) - int(rng.random() < 0.2 and sizes[r] > 0)  # some rows end in padding
        c = np.sort(rng.choice(m, size=used, replace=False))
        cols[off[r]:off[r] + used] = c
        vals[off[r]:off[r] + used] = rng.standard_normal(used)
    return n, m, off, cols, vals


class FileInfo:
    def __init__(self, path):
        L = _lib.lib()
        self.h = ctypes.c_void_p()
        _lib.check(L.tnlmi_matrix_file_read(os.fsencode(path), ctypes.byref(self.h)))
        rows, cols, st = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        fmt, org, vt, mt = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        param = ctypes.c_int32()
        _lib.check(L.tnlmi_matrix_file_info(self.h, ctypes.byref(rows), ctypes.byref(cols), ctypes.byref(fmt),
                                            ctypes.byref(org), ctypes.byref(param), ctypes.byref(vt),
                                            ctypes.byref(st), ctypes.byref(mt)))
        self.symmetric = mt.value == 1
        self.type = L.tnlmi_matrix_file_type(self.h).decode()
        L.tnlmi_matrix_file_destroy(self.h)
        self.rows, self.columns, self.format, self.org = rows.value, cols.value, fmt.value, org.value
        self.param, self.value_type, self.storage = param.value, vt.value, st.value


def read_error(path):
    h = ctypes.c_void_p()
    st = _lib.lib().tnlmi_matrix_file_read(os.fsencode(path), ctypes.byref(h))
    assert st != _lib.OK and not h.value
    return st, _lib.lib().tnlmi_last_error().decode()


# ------------------------------------------------------------------ CPU

@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_read_tnl_file_header(case, tmp_path):
    (name, fmt, param), org = case
    ref = ref_or_skip()
    n, m, off, cols, vals = matrix_csr(1)
    path = tmp_path / "a.tnl"
    ref.save_file(name, off, cols, vals, m, path, col_major=org == tnl.COLUMN_MAJOR)
    info = FileInfo(path)
    want_cols, _ = ref.fill(name, off, cols, vals, m, col_major=org == tnl.COLUMN_MAJOR)
    assert (info.rows, info.columns, info.format, info.param, info.value_type) == (n, m, fmt, param, tnl.F64)
    if fmt in (tnl.ELLPACK, tnl.SLICED_ELLPACK, tnl.BIELLPACK, tnl.CHUNKED_ELLPACK):
        assert info.org == org
    assert info.storage == len(want_cols)
    assert info.type.startswith("Matrices::SparseMatrix< double, ") and info.type.endswith(
        ", [any_device], int, General, [any_allocator], [any_allocator] >")


def test_read_errors(tmp_path):
    ref = ref_or_skip()
    n, m, off, cols, vals = matrix_csr(2, n=40, m=50, long_rows=((7, 45),))
    good = tmp_path / "good.tnl"
    ref.save_file("BiEllpack", off, cols, vals, m, good)
    data = good.read_bytes()
    # missing file
    st, msg = read_error(tmp_path / "none.tnl")
    assert st == _lib.RUNTIME_ERROR and "Unable to open file" in msg
    # wrong magic number
    bad = tmp_path / "magic.tnl"
    bad.write_bytes(b"XNLMN" + data[5:])
    st, msg = read_error(bad)
    assert st == _lib.RUNTIME_ERROR and "wrong magic number - file is not in a TNL-compatible format." in msg
    # truncated anywhere: an error, never a crash
    for cut in (3, 9, 40, len(data) // 3, len(data) // 2, len(data) - 1):
        bad.write_bytes(data[:cut])
        st, msg = read_error(bad)
        assert st == _lib.RUNTIME_ERROR and "Failed to deserialize an object from the file" in msg
    # a column index past the columns
    cpos = data.rfind(b"TNL::Containers::Array< int >")
    first = cpos + len(b"TNL::Containers::Array< int >") + 8
    arr = bytearray(data)
    arr[first:first + 4] = struct.pack("<i", m)
    bad.write_bytes(bytes(arr))
    st, msg = read_error(bad)
    assert st == _lib.RUNTIME_ERROR and "column index" in msg
    # a type this library cannot hold (64-bit indexes)
    t0 = data.index(b"Matrices::SparseMatrix")
    tl = struct.unpack("<i", data[5:9])[0]
    t = data[t0:t0 + tl].replace(b"[any_device], int,", b"[any_device], long int,")
    bad.write_bytes(data[:5] + struct.pack("<i", len(t)) + t + data[t0 + tl:])
    st, msg = read_error(bad)
    assert st == _lib.NOT_IMPLEMENTED and "unsupported object type" in msg


# ------------------------------------------------------------------ GPU

def masked_equal(ours, theirs, fmt):
    """Byte equality of two files, skipping TNL's uninitialised slice entries
    (ChunkedEllpack) and the AdaptiveCSR descriptor arrays."""
    if fmt == tnl.ADAPTIVE_CSR:
        def strip(d):
            out, i = b"", 0
            while True:
                j = d.find(BLOCK_TYPE, i)
                if j < 0:
                    return out + d[i:]
                k = j + len(BLOCK_TYPE)
                count = struct.unpack("<Q", d[k:k + 8])[0]
                out += d[i:k]
                i = k + 8 + 20 * count
        return strip(ours) == strip(theirs)
    if fmt == tnl.CHUNKED_ELLPACK:
        j = ours.find(SLICE_TYPE)
        k = j + len(SLICE_TYPE)
        if ours[:k + 8] != theirs[:k + 8]:
            return False
        count = struct.unpack("<Q", ours[k:k + 8])[0]
        # numberOfSlices: after the type string, rows and columns (size_t), size and storageSize
        p = 9 + struct.unpack("<i", ours[5:9])[0] + 16
        nslices = struct.unpack("<i", ours[p + 8:p + 12])[0]
        used = 16 * nslices
        return (ours[k + 8:k + 8 + used] == theirs[k + 8:k + 8 + used]
                and ours[k + 8 + 16 * count:] == theirs[k + 8 + 16 * count:])
    return ours == theirs


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_save_writes_tnl_bytes(case, tmp_path):
    (name, fmt, param), org = case
    ref = ref_or_skip()
    n, m, off, cols, vals = matrix_csr(3)
    A = tnl.SparseMatrix.from_csr(n, m, off, cols, vals, fmt=fmt, param=param, organization=org)
    ours, theirs = tmp_path / "ours.tnl", tmp_path / "theirs.tnl"
    A.save(ours)
    ref.save_file(name, off, cols, vals, m, theirs, col_major=org == tnl.COLUMN_MAJOR)
    assert masked_equal(ours.read_bytes(), theirs.read_bytes(), fmt)
    assert A.getSerializationType() == FileInfo(theirs).type
    # and TNL loads it into the storage it builds itself
    (r, c), oc, ov = ref.load_file(name, ours, col_major=org == tnl.COLUMN_MAJOR)
    want_c, want_v = ref.fill(name, off, cols, vals, m, col_major=org == tnl.COLUMN_MAJOR)
    assert (r, c) == (n, m)
    assert np.array_equal(oc, want_c) and np.array_equal(ov, want_v)


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_load_tnl_file(case, tmp_path):
    (name, fmt, param), org = case
    ref = ref_or_skip()
    torch = pytest.importorskip("torch")
    n, m, off, cols, vals = matrix_csr(4)
    path = tmp_path / "theirs.tnl"
    ref.save_file(name, off, cols, vals, m, path, col_major=org == tnl.COLUMN_MAJOR)
    A = tnl.SparseMatrix.load(path)
    assert (A.rows, A.columns, A.segments.format) == (n, m, fmt)
    want_c, want_v = ref.fill(name, off, cols, vals, m, col_major=org == tnl.COLUMN_MAJOR)
    assert np.array_equal(A.column_indexes.cpu().numpy()[:len(want_c)], want_c)
    assert np.array_equal(A.values.cpu().numpy()[:len(want_v)], want_v)
    B = tnl.SparseMatrix.from_csr(n, m, off, cols, vals, fmt=fmt, param=param, organization=org)
    assert A.segments.getStorageSize() == B.segments.getStorageSize()
    x = np.random.default_rng(5).standard_normal(m)
    y = torch.zeros(n, dtype=torch.float64, device="cuda")
    A.vectorProduct(torch.from_numpy(x).cuda(), y)
    torch.cuda.synchronize()
    want = ref.spmv(name, off, cols, vals, x, col_major=org == tnl.COLUMN_MAJOR)
    assert np.abs(y.cpu().numpy() - want).max() <= 1e-12 * np.abs(want).max()
    # the loaded matrix saves to the same bytes, and loads again
    again = tmp_path / "again.tnl"
    A.save(again)
    assert masked_equal(again.read_bytes(), path.read_bytes(), fmt)
    assert A.getSerializationType() == FileInfo(path).type
    with pytest.raises(tnl.TnlmiError, match="object type does not match"):
        tnl.SparseMatrix.load(path, expected_type=A.getSerializationType().replace("double", "float"))


@pytest.mark.gpu
def test_round_trip_f32_and_errors(tmp_path):
    torch = pytest.importorskip("torch")
    n, m, off, cols, vals = matrix_csr(6, n=500, m=400)
    for fmt, param in [(tnl.CSR, 0), (tnl.BIELLPACK, 64), (tnl.SLICED_ELLPACK, 16)]:
        A = tnl.SparseMatrix.from_csr(n, m, off, cols, vals, fmt=fmt, param=param, value_dtype=torch.float32)
        p1, p2 = tmp_path / "1.tnl", tmp_path / "2.tnl"
        A.save(p1)
        B = tnl.SparseMatrix.load(p1)
        assert B.values.dtype == torch.float32
        assert torch.equal(A.values, B.values) and torch.equal(A.column_indexes, B.column_indexes)
        B.save(p2)
        assert p1.read_bytes() == p2.read_bytes()
        assert "SparseMatrix< float, " in B.getSerializationType()
    # empty matrix
    E = tnl.SparseMatrix.from_csr(0, 5, np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0))
    E.save(p1)
    F = tnl.SparseMatrix.load(p1)
    assert (F.rows, F.columns) == (0, 5)
    # bf16 has no TNL serialization type
    H = tnl.SparseMatrix.from_csr(n, m, off, cols, vals, value_dtype=torch.bfloat16)
    with pytest.raises(tnl.NotImplementedError_):
        H.save(p1)
    # a BiEllpack permutation that leaves its strip is refused before any kernel runs
    A = tnl.SparseMatrix.from_csr(n, m, off, cols, vals, fmt=tnl.BIELLPACK, param=32)
    A.save(p1)
    data = bytearray(p1.read_bytes())
    ppos = data.index(b"TNL::Containers::Array< int >") + len(b"TNL::Containers::Array< int >") + 8
    data[ppos:ppos + 4] = struct.pack("<i", 1000)
    p2.write_bytes(bytes(data))
    with pytest.raises(tnl.TnlmiError, match="permutation"):
        tnl.SparseMatrix.load(p2)
