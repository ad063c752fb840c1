"""Symmetric-storage SpMV and transposedVectorProduct
(SparseMatrixBase.hpp:298-329, :541-588; tnlmi_spmv with matrix_type
TNLMI_SYMMETRIC, tnlmi_spmv_transposed).

CPU: the oracle restatements (orc_spmv_symmetric, orc_spmv_transposed) against
the dense closed form (A = L + L^T - D; A^T x) and against TNL's own
SymmetricMatrix / transposedVectorProduct run with one thread (same order:
bit-exact).  GPU: every format and organisation against the oracle within a
relative 1e-12 (FP64; the GPU adds with atomics, so the order of a y entry's
terms is free, as in the reference's GPU build), CG on symmetric storage,
save/load of symmetric matrices, and the paths that refuse symmetric storage."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import tnl_amd as tnl  # noqa: E402
import oracle_lib as OL  # noqa: E402

REF_FORMATS = ["CSR", "AdaptiveCSR", "Ellpack", "SlicedEllpack32", "SlicedEllpack64", "BiEllpack", "ChunkedEllpack",
               "SortedCSR"]
GPU_FORMATS = [("CSR", tnl.CSR, 0), ("AdaptiveCSR", tnl.ADAPTIVE_CSR, 0), ("Ellpack", tnl.ELLPACK, 0),
               ("SlicedEllpack32", tnl.SLICED_ELLPACK, 32), ("BiEllpack", tnl.BIELLPACK, 32),
               ("ChunkedEllpack", tnl.CHUNKED_ELLPACK, 0), ("SortedCSR", tnl.SORTED_CSR, 0)]
CASES = [(f, org) for f in GPU_FORMATS for org in (tnl.COLUMN_MAJOR, tnl.ROW_MAJOR)
         if org == tnl.COLUMN_MAJOR or f[1] in (tnl.ELLPACK, tnl.SLICED_ELLPACK, tnl.BIELLPACK, tnl.CHUNKED_ELLPACK)]
IDS = [f"{f[0]}-{'cm' if org == tnl.COLUMN_MAJOR else 'rm'}" for f, org in CASES]


def lower_csr(n, maxlen, seed, pad=0.2):
    """Random lower triangle with the diagonal; some rows end in a padding slot (column -1)."""
    rng = np.random.default_rng(seed)
    off = [0]
    cols, vals = [], []
    for r in range(n):
        k = int(rng.integers(0, min(maxlen, r) + 1))
        c = sorted(rng.choice(r, size=k, replace=False).tolist()) if k else []
        c = c + [r]
        if rng.random() < pad:
            c = c + [-1]
        cols += c
        vals += [float(v) if cc >= 0 else 0.0 for v, cc in zip(rng.standard_normal(len(c)), c)]
        off.append(len(cols))
    return np.array(off, np.int32), np.array(cols, np.int32), np.array(vals)


def dense(off, cols, vals, n, m):
    A = np.zeros((n, m))
    for r in range(n):
        for g in range(off[r], off[r + 1]):
            if cols[g] >= 0:
                A[r, cols[g]] += vals[g]
    return A


def general_csr(n, m, maxlen, seed):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(0, maxlen + 1, n)
    off = np.zeros(n + 1, np.int32)
    off[1:] = np.cumsum(sizes)
    cols = np.concatenate([np.sort(rng.choice(m, size=s, replace=False)) for s in sizes]).astype(np.int32)
    return off, cols, rng.standard_normal(off[-1])


@pytest.fixture(scope="module")
def orc():
    return OL.Oracle()


# ------------------------------------------------------------------ CPU

def test_oracle_symmetric_closed_form(orc):
    n = 80
    off, cols, vals = lower_csr(n, 12, 1)
    L = dense(off, cols, vals, n, n)
    A = L + L.T - np.diag(np.diag(L))
    rng = np.random.default_rng(2)
    x, y0 = rng.standard_normal(n), rng.standard_normal(n)
    for alpha, beta in [(1.0, 0.0), (-0.5, 2.0), (2.0, 1.0)]:
        y = orc.spmv_symmetric(off, cols, vals, x, y0, alpha, beta)
        np.testing.assert_allclose(y, beta * y0 + alpha * A @ x, rtol=1e-13, atol=1e-13)
    # rows [b, e): only those rows' stored entries (and their mirrors) take part
    b, e = 20, 50
    mask = np.zeros((n, n))
    mask[b:e, :] = 1
    Lbe = L * mask
    y = orc.spmv_symmetric(off, cols, vals, x, y0, 1.5, 0.25, b, e)
    want = 0.25 * y0 + 1.5 * (Lbe @ x + (Lbe - np.diag(np.diag(Lbe))).T @ x)
    np.testing.assert_allclose(y, want, rtol=1e-13, atol=1e-13)


def test_oracle_transposed_closed_form(orc):
    n, m = 70, 90
    off, cols, vals = general_csr(n, m, 10, 3)
    A = dense(off, cols, vals, n, m)
    rng = np.random.default_rng(4)
    x, y0 = rng.standard_normal(n), rng.standard_normal(m)
    y = orc.spmv_transposed(off, cols, vals, m, x, y0, -1.25, 0.5)
    np.testing.assert_allclose(y, 0.5 * y0 - 1.25 * A.T @ x, rtol=1e-13, atol=1e-13)
    y = orc.spmv_transposed(off, cols, vals, m, x, y0, 1.0, 1.0, 10, 40)  # columns [10, 40) only
    want = y0.copy()
    want[10:40] += (A.T @ x)[10:40]
    np.testing.assert_allclose(y, want, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("fmt", REF_FORMATS)
def test_oracle_vs_tnl_symmetric_and_transposed(orc, fmt):
    ref = OL.Reference.load()
    if ref is None:
        pytest.skip("oracle/_ref is not built")
    ref.set_threads(1)  # the reference's atomic adds in one thread's order: bit-exact
    try:
        n = 300
        off, cols, vals = lower_csr(n, 9, 5)
        rng = np.random.default_rng(6)
        x, y0 = rng.standard_normal(n), rng.standard_normal(n)
        for alpha, beta, b, e in [(1.0, 0.0, 0, 0), (0.5, -2.0, 0, 0), (1.0, 1.0, 100, 250)]:
            for cm in (True, False):
                want = ref.spmv_symmetric(fmt, off, cols, vals, x, y0, alpha, beta, b, e, col_major=cm)
                got = orc.spmv_symmetric(off, cols, vals, x, y0, alpha, beta, b, e)
                if fmt == "SortedCSR":  # rows visited in the sorted order: the mirrored adds land in another order
                    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)
                else:
                    assert np.array_equal(got, want), (fmt, cm, alpha, beta, b, e)
        goff, gcols, gvals = general_csr(n, 200, 8, 7)
        for alpha, beta, b, e in [(1.0, 0.0, 0, 0), (2.0, 0.5, 30, 120)]:
            want = ref.spmv_transposed(fmt, goff, gcols, gvals, 200, x, y0[:200], alpha, beta, b, e)
            got = orc.spmv_transposed(goff, gcols, gvals, 200, x, y0[:200], alpha, beta, b, e)
            if fmt == "SortedCSR":
                np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)
            else:
                assert np.array_equal(got, want), (fmt, alpha, beta, b, e)
    finally:
        ref.set_threads(os.cpu_count() or 1)


# ------------------------------------------------------------------ GPU

def close(got, want, tol):
    assert np.abs(got - want).max() <= tol * max(1.0, np.abs(want).max())


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_symmetric_spmv(orc, case):
    torch = pytest.importorskip("torch")
    (name, fmt, param), org = case
    n = 2500
    off, cols, vals = lower_csr(n, 14, 11)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=fmt, param=param, organization=org, symmetric=True)
    assert A.symmetric and "Symmetric" in A.getSerializationType()
    rng = np.random.default_rng(12)
    x, y0 = rng.standard_normal(n), rng.standard_normal(n)
    xd = torch.from_numpy(x).cuda()
    for alpha, beta, b, e in [(1.0, 0.0, 0, 0), (-0.75, 1.5, 0, 0), (1.0, 1.0, 0, 0), (2.0, 0.5, 300, 1700),
                              (1.0, 3.0, 900, 900)]:
        yd = torch.from_numpy(y0.copy()).cuda()
        A.vectorProduct(xd, yd, alpha, beta, b, e)
        torch.cuda.synchronize()
        close(yd.cpu().numpy(), orc.spmv_symmetric(off, cols, vals, x, y0, alpha, beta, b, e), 1e-12)
    # transposedVectorProduct of a symmetric matrix is its vectorProduct
    yd = torch.zeros(n, dtype=torch.float64, device="cuda")
    A.transposedVectorProduct(xd, yd)
    torch.cuda.synchronize()
    close(yd.cpu().numpy(), orc.spmv_symmetric(off, cols, vals, x), 1e-12)


@pytest.mark.gpu
def test_symmetric_spmv_f32_bf16(orc):
    torch = pytest.importorskip("torch")
    n = 3000
    off, cols, vals = lower_csr(n, 10, 13)
    x = np.random.default_rng(14).standard_normal(n)
    for vdt, xdt, tol in [(torch.float32, torch.float32, 2e-5), (torch.float32, torch.float64, 1e-12),
                          (torch.bfloat16, torch.float32, 2e-5), (torch.bfloat16, torch.float64, 1e-12)]:
        A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals, value_dtype=vdt, symmetric=True)
        stored = A.values[:len(vals)].double().cpu().numpy()  # the rounded values the matrix holds
        y = torch.zeros(n, dtype=xdt, device="cuda")
        xd = torch.from_numpy(x).to(xdt).cuda()
        A.vectorProduct(xd, y)
        torch.cuda.synchronize()
        want = orc.spmv_symmetric(off, cols, stored, xd.double().cpu().numpy())
        close(y.double().cpu().numpy(), want, tol)
    with pytest.raises(tnl.NotImplementedError_):
        A.vectorProduct(xd.float(), torch.zeros(n, dtype=torch.float32, device="cuda"), compute_dtype=torch.float64)


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_transposed_spmv(orc, case):
    torch = pytest.importorskip("torch")
    (name, fmt, param), org = case
    n, m = 2200, 1900
    off, cols, vals = general_csr(n, m, 20, 15)
    A = tnl.SparseMatrix.from_csr(n, m, off, cols, vals, fmt=fmt, param=param, organization=org)
    rng = np.random.default_rng(16)
    x, y0 = rng.standard_normal(n), rng.standard_normal(m)
    for alpha, beta, b, e in [(1.0, 0.0, 0, 0), (0.5, -1.0, 0, 0), (1.0, 1.0, 250, 1300)]:
        yd = torch.from_numpy(y0.copy()).cuda()
        A.transposedVectorProduct(torch.from_numpy(x).cuda(), yd, alpha, beta, b, e)
        torch.cuda.synchronize()
        close(yd.cpu().numpy(), orc.spmv_transposed(off, cols, vals, m, x, y0, alpha, beta, b, e), 1e-12)


@pytest.mark.gpu
def test_cg_on_symmetric_storage():
    torch = pytest.importorskip("torch")
    from tnl_amd import solvers as S
    n1 = 24
    off, cols, vals = tnl.poisson3d(n1, n1, n1)
    n = len(off) - 1
    keep = np.concatenate([cols[off[r]:off[r + 1]] <= r for r in range(n)])
    lens = np.array([np.count_nonzero(cols[off[r]:off[r + 1]] <= r) for r in range(n)])
    loff = np.zeros(n + 1, np.int32)
    loff[1:] = np.cumsum(lens)
    b = torch.from_numpy(np.random.default_rng(17).standard_normal(n)).cuda()
    results = []
    for fmt, sym, (o, c, v) in [(tnl.ELLPACK, False, (off, cols, vals)), (tnl.ELLPACK, True, (loff, cols[keep], vals[keep])),
                                (tnl.CSR, True, (loff, cols[keep], vals[keep]))]:
        A = tnl.SparseMatrix.from_csr(n, n, o, c, v, fmt=fmt, symmetric=sym)
        for pre in (None, S.Diagonal):
            cg = S.CG()
            cg.setMatrix(A)
            if pre is not None:
                cg.setPreconditioner(pre(A))
            cg.setConvergenceResidue(1e-10)
            x = torch.zeros(n, dtype=torch.float64, device="cuda")
            assert cg.solve(b, x)
            results.append((sym, pre, cg.getIterations(), x.cpu().numpy()))
    full = {pre: (it, x) for sym, pre, it, x in results if not sym}
    for sym, pre, it, x in results:
        assert abs(it - full[pre][0]) <= 2
        np.testing.assert_allclose(x, full[pre][1], rtol=1e-7, atol=1e-8)


@pytest.mark.gpu
def test_symmetric_save_load_and_refusals(tmp_path):
    torch = pytest.importorskip("torch")
    from tnl_amd import solvers as S
    ref = OL.Reference.load()
    n = 400
    off, cols, vals = lower_csr(n, 8, 18)
    for fmt, name in [(tnl.CSR, "CSR"), (tnl.ELLPACK, "Ellpack"), (tnl.BIELLPACK, "BiEllpack")]:
        A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=fmt, symmetric=True)
        p = tmp_path / f"{name}.tnl"
        A.save(p)
        if ref is not None:
            q = tmp_path / f"{name}.ref.tnl"
            ref.save_symmetric_file(name, off, cols, vals, n, q)
            assert p.read_bytes() == q.read_bytes()
        B = tnl.SparseMatrix.load(p)
        assert B.symmetric and B.getSerializationType() == A.getSerializationType()
        assert torch.equal(A.values, B.values) and torch.equal(A.column_indexes, B.column_indexes)
        C = A.convert(tnl.SLICED_ELLPACK, 32)
        assert C.symmetric
    # an entry above the diagonal
    bad_cols = cols.copy()
    r = 5
    bad_cols[off[r]] = r + 1
    with pytest.raises(tnl.LogicError, match="Only lower part"):
        tnl.SparseMatrix.from_csr(n, n, off, bad_cols, vals, symmetric=True)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals, symmetric=True)
    with pytest.raises(tnl.NotImplementedError_):
        S.ILU0(A)
    with pytest.raises(tnl.NotImplementedError_):
        tnl.SpmvPlan(A, 0)
