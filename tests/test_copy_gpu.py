"""copySparseMatrix on the GPU (tnlmi_matrix_row_lengths / tnlmi_matrix_copy,
SparseMatrix.convert): converting a matrix to another format gives exactly the
storage TNL's own copySparseMatrix builds from the same source (oracle/_ref
ref_copy).  As in the reference, a row's length is its leading slots with a
column below the column count over the whole segment (getSegmentSize), so
the padding slots of a padded source format (column -1) are copied as slots
of the target; a chain of conversions back to CSR keeps every row's entries
in order, followed only by such padding slots."""
import ctypes
import os
import sys

import numpy as np
import pytest
import torch

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import tnl_amd as tnl  # noqa: E402
from tnl_amd import _lib  # noqa: E402
import oracle_lib as OL  # noqa: E402
from test_spmv_gpu import random_csr  # noqa: E402

pytestmark = pytest.mark.gpu

TARGETS = [(tnl.CSR, 0, "CSR"), (tnl.ADAPTIVE_CSR, 0, "AdaptiveCSR"), (tnl.ELLPACK, 0, "Ellpack"),
           (tnl.SLICED_ELLPACK, 32, "SlicedEllpack32"), (tnl.SLICED_ELLPACK, 64, "SlicedEllpack64"),
           (tnl.BIELLPACK, 32, "BiEllpack"), (tnl.CHUNKED_ELLPACK, 0, "ChunkedEllpack"),
           (tnl.SORTED_CSR, 0, "SortedCSR")]


@pytest.mark.parametrize("org", [tnl.COLUMN_MAJOR, tnl.ROW_MAJOR], ids=["cm", "rm"])
@pytest.mark.parametrize("fmt,param,name", TARGETS, ids=[t[2] for t in TARGETS])
def test_convert_matches_tnl_storage(fmt, param, name, org):
    ref = OL.Reference.load()
    if ref is None:
        pytest.skip("oracle/_ref is not built")
    rng = np.random.default_rng(17)
    n, m = 700, 900
    off, cols, vals = random_csr(n, m, 40, rng, pad=0.3, long_rows=[(5, 600), (699, 300)])
    A = tnl.SparseMatrix.from_csr(n, m, off, cols, vals, fmt=tnl.BIELLPACK, param=32)  # a padded source format
    # every slot of a BiEllpack segment (entries and padding, column -1) is below the column count:
    # the copy's row lengths are the segment sizes, TNL's getRowCapacity
    seg_sizes = ref.row_counts("BiEllpack", off, cols, vals, m)[0]
    assert np.array_equal(A.rowLengths().cpu().numpy(), seg_sizes)
    assert np.all(seg_sizes >= np.diff(off))
    B = A.convert(fmt, param, org)
    want_c, want_v = ref.copy("BiEllpack", name, off, cols, vals, m, src_col_major=True,
                              dst_col_major=org == tnl.COLUMN_MAJOR)
    assert np.array_equal(B.column_indexes.cpu().numpy()[:len(want_c)], want_c)
    assert np.array_equal(B.values.cpu().numpy()[:len(want_v)], want_v)
    x = rng.standard_normal(m)
    y = torch.zeros(n, dtype=torch.float64, device="cuda")
    B.vectorProduct(torch.from_numpy(x).cuda(), y)
    torch.cuda.synchronize()
    want = OL.Oracle().spmv(off, cols, vals, x)  # FP64 within 1e-12 (kernels split long rows)
    assert np.abs(y.cpu().numpy() - want).max() <= 1e-12 * np.abs(want).max()


def test_convert_chain_back_to_csr():
    ref = OL.Reference.load()
    rng = np.random.default_rng(18)
    n, m = 1000, 1000
    off, cols, vals = random_csr(n, m, 30, rng, pad=0.2, long_rows=[(0, 900)])
    A = tnl.SparseMatrix.from_csr(n, m, off, cols, vals)
    chain = A
    hops = [(tnl.BIELLPACK, 32, tnl.ROW_MAJOR), (tnl.CHUNKED_ELLPACK, 0, tnl.COLUMN_MAJOR),
            (tnl.SLICED_ELLPACK, 32, tnl.ROW_MAJOR), (tnl.SORTED_CSR, 0, tnl.COLUMN_MAJOR),
            (tnl.ELLPACK, 0, tnl.COLUMN_MAJOR), (tnl.CSR, 0, tnl.COLUMN_MAJOR)]
    for i, (fmt, param, org) in enumerate(hops):
        chain = chain.convert(fmt, param, org)
        if i == 0 and ref is not None:  # the first hop, from CSR arrays, exactly as TNL's copySparseMatrix
            want_c, want_v = ref.copy("CSR", "BiEllpack", off, cols, vals, m, dst_col_major=False)
            assert np.array_equal(chain.column_indexes.cpu().numpy()[:len(want_c)], want_c)
            assert np.array_equal(chain.values.cpu().numpy()[:len(want_v)], want_v)
    # every row: its entries in their order, then only padding slots (column -1, value 0)
    off2 = chain.segments.array("offsets")
    c2, v2 = chain.column_indexes.cpu().numpy(), chain.values.cpu().numpy()
    for r in range(n):
        a, b = off[r], off[r + 1]
        real = cols[a:b] >= 0
        k = int(real.sum())
        row_c, row_v = c2[off2[r]:off2[r + 1]], v2[off2[r]:off2[r + 1]]
        assert len(row_c) >= b - a
        assert np.array_equal(row_c[:k], cols[a:b][real]) and np.array_equal(row_v[:k], vals[a:b][real])
        assert np.all(row_c[k:] == -1) and np.all(row_v[k:] == 0.0)
    x = rng.standard_normal(m)
    y = torch.zeros(n, dtype=torch.float64, device="cuda")
    chain.vectorProduct(torch.from_numpy(x).cuda(), y)
    torch.cuda.synchronize()
    want = OL.Oracle().spmv(off, cols, vals, x)
    assert np.abs(y.cpu().numpy() - want).max() <= 1e-12 * np.abs(want).max()


def test_convert_bf16_and_errors():
    rng = np.random.default_rng(19)
    off, cols, vals = random_csr(200, 300, 20, rng)
    A = tnl.SparseMatrix.from_csr(200, 300, off, cols, vals, value_dtype=torch.bfloat16)
    B = A.convert(tnl.ELLPACK)
    C = B.convert(tnl.CSR)
    # every Ellpack row has the full width: the copy back to CSR keeps its padding slots, as TNL's
    # copySparseMatrix does (the BF16 values are exact in FP64, so TNL's FP64 chain is the expectation)
    ref = OL.Reference.load()
    if ref is not None:
        vb = A.values[:len(vals)].double().cpu().numpy()
        want_c, want_v = ref.copy("Ellpack", "CSR", off, cols, vb, 300)
        assert np.array_equal(C.column_indexes.cpu().numpy()[:len(want_c)], want_c)
        assert np.array_equal(C.values[:len(want_v)].double().cpu().numpy(), want_v)
    # a target whose segments are shorter than the rows is refused, nothing written past them
    L = A.rowLengths()
    short = L.clone()
    r = int(torch.argmax(L))
    assert int(L[r]) > 0
    short[r] = int(L[r]) - 1
    seg = tnl.Segments(tnl.CSR, short)
    vals_t = torch.empty(max(1, seg.getStorageSize()), dtype=torch.bfloat16, device="cuda")
    cols_t = torch.empty(max(1, seg.getStorageSize()), dtype=torch.int32, device="cuda")
    D = tnl.SparseMatrix(200, 300, seg, vals_t, cols_t)
    with pytest.raises(tnl.TnlmiError):
        _lib.check(_lib.lib().tnlmi_matrix_copy(ctypes.byref(A._m), ctypes.byref(D._m),
                                                ctypes.c_void_p(L.data_ptr()), None))
    # different value types
    E = tnl.SparseMatrix.from_csr(200, 300, off, cols, vals)
    with pytest.raises(tnl.TnlmiError):
        _lib.check(_lib.lib().tnlmi_matrix_copy(ctypes.byref(E._m), ctypes.byref(B._m),
                                                ctypes.c_void_p(L.data_ptr()), None))
