"""Format converters on the GPU (setSegmentsSizes of every format), bit-exact
against the pinned oracle (which agrees bit for bit with TNL's own converters,
tests/test_oracle.py), plus the global-index map and the assembled storage."""
import numpy as np
import pytest

import oracle_lib as OL

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
tnl = pytest.importorskip("tnl_amd")

RNG = np.random.default_rng(2024)
SIZE_CASES = [np.array([0], np.int32), np.array([5]), np.array([1, 2, 1, 1]), np.arange(31) % 7,
              np.arange(32) % 5, np.arange(33) % 9, np.full(64, 3), RNG.integers(0, 40, 1000),
              np.concatenate([RNG.integers(0, 20, 900), [700, 0, 3000]]), RNG.integers(0, 3, 5001),
              np.zeros(100, np.int64), RNG.integers(0, 100, 70000)]


@pytest.mark.parametrize("k", range(len(SIZE_CASES)))
def test_converters_bitexact(orc, k):
    sizes = OL.i32(SIZE_CASES[k])
    off = orc.csr_offsets(sizes)
    s = tnl.Segments(tnl.CSR, sizes)
    assert np.array_equal(s.array("offsets"), off)
    assert s.getStorageSize() == off[-1] and s.getElementCount() == off[-1]
    for org, cm in [(tnl.COLUMN_MAJOR, True), (tnl.ROW_MAJOR, False)]:
        rc, w, a = orc.ellpack_dims(sizes, cm)
        e = tnl.Segments(tnl.ELLPACK, sizes, org)
        assert e.ellpack_dims() == (w, a) and e.getStorageSize() == w * a
    for S in (32, 64):
        so, ss = orc.sliced(sizes, S)
        e = tnl.Segments(tnl.SLICED_ELLPACK, sizes, param=S)
        assert np.array_equal(e.array("slice_offsets"), so) and np.array_equal(e.array("slice_segment_sizes"), ss)
        assert e.getStorageSize() == so[-1]
    perm, gp = orc.biellpack(sizes)
    e = tnl.Segments(tnl.BIELLPACK, sizes, param=32)
    assert np.array_equal(e.array("permutation"), perm) and np.array_equal(e.array("group_pointers"), gp)
    c = orc.chunked(sizes)
    e = tnl.Segments(tnl.CHUNKED_ELLPACK, sizes)
    assert e.getStorageSize() == c["storage"]
    assert np.array_equal(e.array("chunk_slices"), c["slices"])
    for key, name in [("seg_to_chunk", "segment_to_chunk"), ("seg_to_slice", "segment_to_slice"),
                      ("seg_ptrs", "segment_pointers"), ("chunks_to_seg", "chunks_to_segments")]:
        assert np.array_equal(e.array(name), c[key]), key
    e = tnl.Segments(tnl.ADAPTIVE_CSR, sizes)
    for vs in (4, 8):
        assert np.array_equal(e.adaptive_blocks(vs), orc.adaptive_blocks(off, vs, host_warp=32, device_warp=64))
    ss, inv, perm = orc.sorted_segments(sizes)
    e = tnl.Segments(tnl.SORTED_CSR, sizes)
    assert np.array_equal(e.array("inverse_permutation"), inv) and np.array_equal(e.array("permutation"), perm)
    assert np.array_equal(e.array("offsets"), orc.csr_offsets(ss))


@pytest.mark.parametrize("fmt,name,param", [(tnl.ELLPACK, "Ellpack", 0), (tnl.SLICED_ELLPACK, "SlicedEllpack", 32),
                                            (tnl.SLICED_ELLPACK, "SlicedEllpack", 64), (tnl.BIELLPACK, "BiEllpack", 32),
                                            (tnl.CHUNKED_ELLPACK, "ChunkedEllpack", 0), (tnl.SORTED_CSR, "SortedCSR", 0),
                                            (tnl.CSR, "CSR", 0)])
@pytest.mark.parametrize("org", [tnl.COLUMN_MAJOR, tnl.ROW_MAJOR], ids=["cm", "rm"])
def test_global_indexes_and_fill(orc, fmt, name, param, org):
    for n, maxlen in [(1, 3), (45, 12), (2000, 50)]:
        sizes = RNG.integers(0, maxlen + 1, n).astype(np.int32)
        kw = dict(slice_size=param) if name == "SlicedEllpack" else {}
        want = orc.gidx(name, sizes, col_major=(org == tnl.COLUMN_MAJOR), **kw)
        seg = tnl.Segments(fmt, sizes, org, param)
        got = seg.global_indexes().cpu().numpy()
        assert np.array_equal(got, want), (name, n)
        # assembled storage = scatter of the CSR element list through the map
        off = orc.csr_offsets(sizes)
        cols = RNG.integers(0, 1000, off[-1]).astype(np.int32)
        vals = RNG.standard_normal(off[-1])
        A = tnl.SparseMatrix.from_csr(n, 1000, off, cols, vals, fmt=fmt, param=param, organization=org)
        st = seg.getStorageSize()
        ec = np.full(st, -1, np.int32)
        ev = np.zeros(st)
        ec[want] = cols
        ev[want] = vals
        assert np.array_equal(A.column_indexes[:st].cpu().numpy(), ec)
        assert np.array_equal(A.values[:st].cpu().numpy(), ev)
