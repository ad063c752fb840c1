"""copySparseMatrix on the GPU (tnlmi_matrix_row_lengths / tnlmi_matrix_copy,
SparseMatrix.convert): converting a matrix to another format gives exactly the
storage TNL builds for that format from the same elements (oracle/_ref fill:
SparseMatrix::setRowCapacities + setElement), and a chain of conversions back
to CSR gives the original arrays."""
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
    A = tnl.SparseMatrix.from_csr(n, m, off, cols, vals, fmt=tnl.BIELLPACK, param=32)  # any source format
    assert np.array_equal(A.rowLengths().cpu().numpy(), np.diff(off))
    B = A.convert(fmt, param, org)
    want_c, want_v = ref.fill(name, off, cols, vals, m, col_major=org == tnl.COLUMN_MAJOR)
    assert np.array_equal(B.column_indexes.cpu().numpy()[:len(want_c)], want_c)
    assert np.array_equal(B.values.cpu().numpy()[:len(want_v)], want_v)
    x = rng.standard_normal(m)
    y = torch.zeros(n, dtype=torch.float64, device="cuda")
    B.vectorProduct(torch.from_numpy(x).cuda(), y)
    torch.cuda.synchronize()
    want = OL.Oracle().spmv(off, cols, vals, x)  # FP64 within 1e-12 (kernels split long rows)
    assert np.abs(y.cpu().numpy() - want).max() <= 1e-12 * np.abs(want).max()


def test_convert_chain_back_to_csr():
    rng = np.random.default_rng(18)
    n, m = 1000, 1000
    off, cols, vals = random_csr(n, m, 30, rng, pad=0.2, long_rows=[(0, 900)])
    A = tnl.SparseMatrix.from_csr(n, m, off, cols, vals)
    chain = A
    for fmt, param, org in [(tnl.BIELLPACK, 32, tnl.ROW_MAJOR), (tnl.CHUNKED_ELLPACK, 0, tnl.COLUMN_MAJOR),
                            (tnl.SLICED_ELLPACK, 32, tnl.ROW_MAJOR), (tnl.SORTED_CSR, 0, tnl.COLUMN_MAJOR),
                            (tnl.ELLPACK, 0, tnl.COLUMN_MAJOR), (tnl.CSR, 0, tnl.COLUMN_MAJOR)]:
        chain = chain.convert(fmt, param, org)
    assert np.array_equal(chain.segments.array("offsets"), off)
    assert np.array_equal(chain.column_indexes.cpu().numpy()[:len(cols)], cols)
    assert np.array_equal(chain.values.cpu().numpy()[:len(vals)], vals)


def test_convert_bf16_and_errors():
    rng = np.random.default_rng(19)
    off, cols, vals = random_csr(200, 300, 20, rng)
    A = tnl.SparseMatrix.from_csr(200, 300, off, cols, vals, value_dtype=torch.bfloat16)
    B = A.convert(tnl.ELLPACK)
    C = B.convert(tnl.CSR)
    assert np.array_equal(C.values[:len(vals)].float().cpu().numpy(),
                          A.values[:len(vals)].float().cpu().numpy())
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
