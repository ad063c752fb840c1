"""MatrixMarket input (tnlmi_mtx_read / tnlmi_mtx_parse): the arrays equal
those of TNL's own MatrixReader::readMtx + SparseMatrix::setElements
(oracle/_ref, Matrices/MatrixReader.hpp:52-306), bit for bit, and bad files
raise with the reference's messages."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import tnl_amd as tnl  # noqa: E402

H = "%%MatrixMarket matrix coordinate"

FILES = {
    # duplicates (the last value wins), comments, unsorted entries, an empty row
    "general_dups": f"{H} real general\n% a comment\n%another\n5 6 7\n1 1 2.5\n3 6 -1\n2 2 7\n1 1 3.5\n"
                    "3 1 1e-3\n5 2 -0.25\n3 3 4\n",
    "symmetric": f"{H} real symmetric\n4 4 6\n1 1 4\n2 1 -1\n2 2 4\n3 2 -1.5\n4 1 0.125\n4 4 2\n",
    "pattern_general": f"{H} pattern general\n3 5 4\n3 5\n1 2\n1 1\n2 4\n",
    "pattern_symmetric": f"{H} pattern symmetric\n3 3 3\n1 1\n3 1\n2 3\n",
    "integer": f"{H} integer general\n2 3 3\n1 3 7\n2 1 -2\n1 1 5\n",
    "coordinates_keyword": f"{H}s real general\n2 2 2\n2 2 1.5\n1 2 -3\n",
    "extra_spaces": f"{H}   real   general\n  3   3   2  \n 1   3   9.75 \n3 1    -8\n",
    "empty_matrix": f"{H} real general\n3 2 0\n",
}

BAD = {
    "array": "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",
    "complex": f"{H} complex general\n2 2 1\n1 1 1 0\n",
    "hermitian": f"{H} real hermitian\n2 2 1\n1 1 1\n",
    "no_banner": "matrix coordinate real general\n2 2 1\n1 1 1\n",
    "not_matrix": "%%MatrixMarket vector coordinate real general\n2 2 1\n1 1 1\n",
    "short_banner": "%%MatrixMarket matrix coordinate real\n2 2 1\n1 1 1\n",
    "dims_tokens": f"{H} real general\n2 2\n1 1 1\n",
    "dims_zero": f"{H} real general\n0 2 1\n1 1 1\n",
    "out_of_range": f"{H} real general\n2 2 1\n3 1 1\n",
    "entry_tokens": f"{H} real general\n2 2 1\n1 1\n",
    "pattern_with_value": f"{H} pattern general\n2 2 1\n1 1 1\n",
}


def _write(tmp_path, name, text):
    p = tmp_path / f"{name}.mtx"
    p.write_text(text)
    return str(p)


def test_known_answer():
    rows, cols, off, col, val = tnl.parse_mtx(FILES["general_dups"])
    assert (rows, cols) == (5, 6)
    assert off.tolist() == [0, 1, 2, 5, 5, 6]
    assert col.tolist() == [0, 1, 0, 2, 5, 1]
    assert val.tolist() == [3.5, 7.0, 1e-3, 4.0, -1.0, -0.25]
    rows, cols, off, col, val = tnl.parse_mtx(FILES["pattern_symmetric"])
    assert off.tolist() == [0, 2, 3, 5] and col.tolist() == [0, 2, 2, 0, 1] and val.tolist() == [1.0] * 5


@pytest.mark.parametrize("name", sorted(FILES))
def test_matches_tnl_reader(ref, tmp_path, name):
    path = _write(tmp_path, name, FILES[name])
    got = tnl.read_mtx(path)
    want = ref.read_mtx(path)
    assert got[:2] == want[:2]
    for g, w in zip(got[2:], want[2:]):
        assert g.dtype == w.dtype and np.array_equal(g, w), name
    # the in-memory overload reads the same
    again = tnl.parse_mtx(FILES[name])
    assert all(np.array_equal(a, b) for a, b in zip(again[2:], got[2:]))


def test_matches_tnl_reader_random(ref, tmp_path):
    rng = np.random.default_rng(3)
    for trial in range(4):
        n, m = int(rng.integers(1, 300)), int(rng.integers(1, 300))
        k = int(rng.integers(0, 2000))
        sym = trial % 2 == 1
        if sym:
            m = n
        r = rng.integers(1, n + 1, k)
        c = rng.integers(1, m + 1, k)
        if sym:
            r, c = np.maximum(r, c), np.minimum(r, c)
        v = rng.standard_normal(k)
        lines = [f"{H} real {'symmetric' if sym else 'general'}", f"{n} {m} {k}"]
        lines += [f"{a} {b} {x!r}" for a, b, x in zip(r, c, v)]
        path = _write(tmp_path, f"random{trial}", "\n".join(lines) + "\n")
        got, want = tnl.read_mtx(path), ref.read_mtx(path)
        assert got[:2] == want[:2]
        assert all(np.array_equal(g, w) for g, w in zip(got[2:], want[2:])), trial


@pytest.mark.parametrize("name", sorted(BAD))
def test_bad_files_raise_like_tnl(ref, tmp_path, name):
    path = _write(tmp_path, name, BAD[name])
    with pytest.raises(RuntimeError) as want:
        ref.read_mtx(path)
    with pytest.raises(tnl.TnlmiError) as got:
        tnl.read_mtx(path)
    assert str(got.value) == str(want.value)


def test_missing_file_and_bad_indexes(tmp_path):
    with pytest.raises(tnl.TnlmiError, match="not able to open"):
        tnl.read_mtx(str(tmp_path / "none.mtx"))
    with pytest.raises(tnl.TnlmiError, match="below the first row"):
        tnl.parse_mtx(f"{H} real general\n2 2 1\n0 1 1\n")


@pytest.mark.gpu
@pytest.mark.parametrize("fmt,param", [(tnl.CSR, 0), (tnl.ELLPACK, 0), (tnl.SLICED_ELLPACK, 32),
                                       (tnl.BIELLPACK, 32), (tnl.CHUNKED_ELLPACK, 0)])
def test_from_mtx_spmv(orc, tmp_path, fmt, param):
    import torch
    path = _write(tmp_path, "sym", FILES["symmetric"])
    A = tnl.SparseMatrix.from_mtx(path, fmt=fmt, param=param)
    rows, cols, off, col, val = tnl.read_mtx(path)
    x = np.arange(1.0, cols + 1)
    y = torch.zeros(rows, dtype=torch.float64, device="cuda")
    A.vectorProduct(torch.from_numpy(x).cuda(), y)
    torch.cuda.synchronize()
    assert np.array_equal(y.cpu().numpy(), orc.spmv(off, col, val, x))
