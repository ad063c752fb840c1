"""Krylov solvers on the GPU against the pinned oracle (which agrees with TNL's
own CG / GMRES / Jacobi / ILU0 on the host, tests/test_oracle.py).

FP64 throughout.  The GPU reductions sum in a different order than the host
loops, so iteration counts may differ by one near the stopping threshold; the
solutions agree to the tolerance stated per test and reach the same true
residual."""
import numpy as np
import pytest

import oracle_lib as OL

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
tnl = pytest.importorskip("tnl_amd")
from tnl_amd import solvers as S  # noqa: E402


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def nonsym(n, seed):
    rng = np.random.default_rng(seed)
    off = np.zeros(n + 1, np.int32)
    cols, vals = [], []
    for i in range(n):
        row = {i: 4.0}
        if i > 0:
            row[i - 1] = -1.3
        if i < n - 1:
            row[i + 1] = -0.7
        j = int(rng.integers(0, n))
        if j != i:
            row[j] = row.get(j, 0.0) + 0.2 * rng.standard_normal()
        for c in sorted(row):
            cols.append(c)
            vals.append(row[c])
        off[i + 1] = len(cols)
    return off, np.array(cols, np.int32), np.array(vals)


FORMATS = [(tnl.CSR, 0), (tnl.ELLPACK, 0), (tnl.SLICED_ELLPACK, 64), (tnl.BIELLPACK, 32), (tnl.CHUNKED_ELLPACK, 0),
           (tnl.ADAPTIVE_CSR, 0), (tnl.SORTED_CSR, 0)]


@pytest.mark.parametrize("fmt,param", FORMATS, ids=[tnl.FORMAT_NAMES[f] for f, _ in FORMATS])
@pytest.mark.parametrize("pre", [0, 1], ids=["none", "diagonal"])
def test_cg(orc, fmt, param, pre):
    off, cols, vals = tnl.laplace2d(40, 31)
    n = len(off) - 1
    b = np.random.default_rng(1).standard_normal(n)
    xo, so = orc.cg(off, cols, vals, b, precond=pre, tol=1e-10)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=fmt, param=param)
    cg = S.CG()
    cg.setMatrix(A)
    if pre:
        cg.setPreconditioner(S.Diagonal(A))
    cg.setConvergenceResidue(1e-10)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert cg.solve(dev(b), x)
    assert abs(cg.getIterations() - so.iterations) <= 1
    np.testing.assert_allclose(x.cpu().numpy(), xo, rtol=1e-7, atol=1e-9)
    assert S.true_residual(A, x, dev(b)) < 1e-9
    assert abs(S.true_residual(A, x, dev(b)) - orc.true_residual(off, cols, vals, x.cpu().numpy(), b)) < 1e-12


def test_cg_split_plan_and_stopping(orc):
    off, cols, vals = tnl.laplace2d(30, 30)
    n = len(off) - 1
    b = np.random.default_rng(2).standard_normal(n)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals)
    P = tnl.SpmvPlan(A, 8)
    cg = S.CG()
    cg.setMatrix(A, plan=P)
    cg.setConvergenceResidue(1e-10)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert cg.solve(dev(b), x)
    xo, so = orc.cg(off, cols, vals, b, tol=1e-10)
    assert abs(cg.getIterations() - so.iterations) <= 1
    np.testing.assert_allclose(x.cpu().numpy(), xo, rtol=1e-7, atol=1e-9)
    # IterativeSolver stopping: max iterations, divergence, NaN (IterativeSolver.hpp:98-143)
    cg = S.CG()
    cg.setMatrix(A)
    cg.setConvergenceResidue(1e-14)
    cg.setMaxIterations(5)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert not cg.solve(dev(b), x)
    xo, so = orc.cg(off, cols, vals, b, tol=1e-14, maxit=5)
    assert cg.getIterations() == so.iterations and not so.converged
    cg.setMaxIterations(10**9)
    cg.setDivergenceResidue(1e-3)
    assert not cg.solve(dev(b), torch.zeros(n, dtype=torch.float64, device="cuda"))
    bn = b.copy()
    bn[3] = np.nan
    assert not cg.solve(dev(bn), torch.zeros(n, dtype=torch.float64, device="cuda"))
    # zero right-hand side: x = 0, converged (CG.hpp:31-35)
    x = dev(np.ones(n))
    assert cg.solve(dev(np.zeros(n)), x) and float(x.abs().max()) == 0.0


@pytest.mark.parametrize("variant", ["CGS", "CGSR", "MGS", "MGSR", "CWY"])
@pytest.mark.parametrize("pre", [0, 1, 2], ids=["none", "diagonal", "ilu0"])
def test_gmres(orc, variant, pre):
    off, cols, vals = nonsym(500, 7)
    n = 500
    b = np.random.default_rng(8).standard_normal(n)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals)
    for rmin, rmax in [(10, 10), (5, 20)]:
        xo, so = orc.gmres(off, cols, vals, b, precond=pre, variant=S.GMRES_VARIANTS[variant], rmin=rmin, rmax=rmax,
                           tol=1e-10)
        g = S.GMRES()
        g.setMatrix(A)
        assert g.setup(variant=variant, restarting_min=rmin, restarting_max=rmax)
        if pre == 1:
            g.setPreconditioner(S.Diagonal(A))
        elif pre == 2:
            g.setPreconditioner(S.ILU0(A))
        g.setConvergenceResidue(1e-10)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        assert g.solve(dev(b), x) and so.converged
        assert abs(g.getIterations() - so.iterations) <= 1, (g.getIterations(), so.iterations)
        np.testing.assert_allclose(x.cpu().numpy(), xo, rtol=1e-7, atol=1e-9)


@pytest.mark.parametrize("lo", ["fp32", "bf16"])
@pytest.mark.parametrize("variant", ["CGSR", "MGSR", "CWY"])
def test_gmres_mixed_precision(orc, lo, variant):
    """Mixed-precision GMRES (Arnoldi on FP32/BF16 values, FP64 restart residuals)
    reaches the FP64 solution: the true residual ||b - A x|| / ||b|| of the FP64
    matrix below 1e-10 (the preconditioned residue below the tolerance), x within
    1e-8 of the oracle's FP64 GMRES; with the inner matrix equal to A it is TNL's
    GMRES (same iterations as the FP64 solver)."""
    off, cols, vals, _ = tnl.convdiff(48, seed=3)
    n = len(off) - 1
    b = np.random.default_rng(12).standard_normal(n)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals)
    dt = torch.float32 if lo == "fp32" else torch.bfloat16
    A_lo = tnl.SparseMatrix(A.rows, A.columns, A.segments, A.values.to(dt), A.column_indexes)
    xo, so = orc.gmres(off, cols, vals, b, precond=1, variant=S.GMRES_VARIANTS[variant], rmin=30, rmax=30,
                       tol=1e-11, maxit=100000)
    assert so.converged
    for inner in (A_lo, A):
        g = S.MixedGMRES()
        g.setMatrix(A)
        g.setInnerMatrix(inner)
        g.setPreconditioner(S.Diagonal(A))
        assert g.setup(variant=variant, restarting_min=30, restarting_max=30)
        g.setConvergenceResidue(1e-11)
        g.setMaxIterations(100000)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        assert g.solve(dev(b), x)
        assert g.getResidue() < 1e-11
        xs = x.cpu().numpy()
        tr = np.linalg.norm(orc.spmv(off, cols, vals, xs) - b) / np.linalg.norm(b)
        assert tr < 1e-10, tr
        np.testing.assert_allclose(xs, xo, rtol=1e-8, atol=1e-8 * np.abs(xo).max())
        if inner is A:
            ref = S.GMRES()
            ref.setMatrix(A)
            ref.setPreconditioner(S.Diagonal(A))
            ref.setup(variant=variant, restarting_min=30, restarting_max=30)
            ref.setConvergenceResidue(1e-11)
            ref.setMaxIterations(100000)
            x2 = torch.zeros(n, dtype=torch.float64, device="cuda")
            assert ref.solve(dev(b), x2)
            # a cycle whose estimate meets the rule is confirmed by the FP64 residue of its
            # iterate: the same iterations, or at most one more cycle when the two straddle it
            assert ref.getIterations() <= g.getIterations() <= ref.getIterations() + 30
            np.testing.assert_allclose(x2.cpu().numpy(), xs, rtol=1e-9, atol=1e-9 * np.abs(xs).max())


def test_gmres_errors():
    g = S.GMRES()
    with pytest.raises(tnl.LogicError):
        g.solve(dev(np.ones(3)), dev(np.zeros(3)))
    off, cols, vals = nonsym(50, 1)
    A = tnl.SparseMatrix.from_csr(50, 50, off, cols, vals)
    g.setMatrix(A)
    assert not g.setup(variant="XYZ")
    g.setup(variant="MGS", restarting_min=20, restarting_max=10)
    assert not g.solve(dev(np.ones(50)), dev(np.zeros(50)))  # wrong restarting parameters -> false


def test_gmres_cwy_is_the_cpp_default_and_matches_tnl(orc):
    # CWY (compact-WY Householder) is GMRES's class default in the reference
    # (GMRES.h:212); check it against TNL's own Host GMRES where that is built
    off, cols, vals = nonsym(300, 11)
    b = np.random.default_rng(12).standard_normal(300)
    A = tnl.SparseMatrix.from_csr(300, 300, off, cols, vals)
    g = S.GMRES()
    g.setMatrix(A)
    g.setup(variant="CWY", restarting_min=30, restarting_max=30)
    g.setConvergenceResidue(1e-11)
    x = torch.zeros(300, dtype=torch.float64, device="cuda")
    assert g.solve(dev(b), x)
    assert S.true_residual(A, x, dev(b)) < 1e-9
    xo, so = orc.gmres(off, cols, vals, b, variant=4, rmin=30, rmax=30, tol=1e-11)
    assert abs(g.getIterations() - so.iterations) <= 1
    np.testing.assert_allclose(x.cpu().numpy(), xo, rtol=1e-7, atol=1e-9)


def test_jacobi(orc):
    off, cols, vals = nonsym(400, 3)
    b = np.random.default_rng(4).standard_normal(400)
    xo, so = orc.jacobi(off, cols, vals, b, tol=1e-9)
    A = tnl.SparseMatrix.from_csr(400, 400, off, cols, vals)
    j = S.Jacobi()
    j.setMatrix(A)
    j.setConvergenceResidue(1e-9)
    x = torch.zeros(400, dtype=torch.float64, device="cuda")
    assert j.solve(dev(b), x)
    assert j.getIterations() == so.iterations
    np.testing.assert_allclose(x.cpu().numpy(), xo, rtol=1e-10, atol=1e-12)


def test_preconditioners(orc, ):
    off, cols, vals = tnl.laplace2d(23, 19)
    n = len(off) - 1
    b = np.random.default_rng(9).standard_normal(n)
    for fmt, param in FORMATS:
        A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=fmt, param=param)
        L, U = orc.ilu0(off, cols, vals)
        got = S.ILU0(A).solve(dev(b), torch.zeros(n, dtype=torch.float64, device="cuda")).cpu().numpy()
        np.testing.assert_allclose(got, orc.ilu0_solve(L, U, b), rtol=1e-12, atol=1e-13)
        got = S.Diagonal(A).solve(dev(b), torch.zeros(n, dtype=torch.float64, device="cuda")).cpu().numpy()
        np.testing.assert_array_equal(got, b / 4.0)


def _block_csr(blocks):
    """Block-diagonal CSR arrays from a list of (off, cols, vals) blocks."""
    offs, cs, vs, base, nnz = [np.zeros(1, np.int32)], [], [], 0, 0
    for off, cols, vals in blocks:
        offs.append(off[1:] + nnz)
        cs.append(cols + base)
        vs.append(vals)
        base += len(off) - 1
        nnz += int(off[-1])
    return np.concatenate(offs).astype(np.int32), np.concatenate(cs).astype(np.int32), np.concatenate(vs)


def test_ilu0_apply_narrow_and_wide_levels(orc):
    """The triangular solves run consecutive levels of at most 2,048 rows in one
    one-workgroup launch (at most 16,384 levels per launch; levels of at most
    1,024 rows with at most 3 entries per row through the LDS ring) and wider
    levels one launch each; all of them, in any order, match the oracle's ILU(0)
    solve."""
    rng = np.random.default_rng(31)
    # 5,000 independent rows (one wide level 0), then a 60 x 60 Laplacian (narrow levels)
    d = 5000
    diag = (np.arange(d + 1, dtype=np.int32), np.arange(d, dtype=np.int32), 2.0 + rng.random(d))
    lap = tnl.laplace2d(60, 60)
    # rows with more than 4 lower entries (entries past the ones loaded ahead of their level):
    # a random lower-heavy pattern plus a dominant diagonal
    m = 3000
    rows = []
    for r in range(m):
        lo = rng.choice(r, size=min(r, 9), replace=False) if r else np.zeros(0, np.int64)
        hi = rng.choice(np.arange(r + 1, m), size=min(m - r - 1, 3), replace=False) if r < m - 1 else np.zeros(0)
        rows.append(np.sort(np.concatenate([lo, [r], hi]).astype(np.int32)))
    roff = np.zeros(m + 1, np.int32)
    roff[1:] = np.cumsum([len(c) for c in rows])
    rcols = np.concatenate(rows)
    rvals = np.where(rcols == np.repeat(np.arange(m), np.diff(roff)), 20.0, rng.standard_normal(len(rcols)) * 0.5)
    # a tridiagonal matrix of 40,000 rows: 40,000 levels of one row each, so the
    # run is cut into three launches
    t = 40000
    toff = np.zeros(t + 1, np.int32)
    toff[1:] = np.cumsum(np.minimum(np.arange(t) + 2, 3) - (np.arange(t) == t - 1))
    tcols = np.concatenate([np.arange(max(r - 1, 0), min(r + 2, t), dtype=np.int32) for r in range(t)])
    tvals = np.where(tcols == np.repeat(np.arange(t), np.diff(toff)), 2.5, -1.0 + 0.1 * rng.random(len(tcols)))
    tri = (toff, tcols, tvals)
    for blocks in ([diag, lap], [lap, diag], [(roff, rcols, rvals)], [diag, (roff, rcols, rvals), lap],
                   [tnl.laplace2d(300, 300)], [tri], [diag, tri]):
        off, cols, vals = _block_csr(blocks)
        n = len(off) - 1
        b = rng.standard_normal(n)
        A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals)
        L, U = orc.ilu0(off, cols, vals)
        want = orc.ilu0_solve(L, U, b)
        P = S.ILU0(A)
        for _ in range(2):  # the same apply twice (no state kept between applies)
            got = P.solve(dev(b), torch.zeros(n, dtype=torch.float64, device="cuda")).cpu().numpy()
            np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-13 * np.abs(want).max())


def test_ilu0_apply_tiny_matrices(orc):
    """ILU(0) on matrices of 1, 2, 3 and 5 rows (one ring run of one or a few
    levels of a single row; the stages past the run's last level load nothing
    to solve) against the oracle's solve."""
    rng = np.random.default_rng(41)
    for n in (1, 2, 3, 5):
        rows = [np.arange(max(r - 1, 0), min(r + 2, n), dtype=np.int32) for r in range(n)]
        off = np.zeros(n + 1, np.int32)
        off[1:] = np.cumsum([len(c) for c in rows])
        cols = np.concatenate(rows)
        vals = np.where(cols == np.repeat(np.arange(n), np.diff(off)), 3.0 + rng.random(len(cols)),
                        -1.0 + 0.2 * rng.random(len(cols)))
        b = rng.standard_normal(n)
        A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals)
        L, U = orc.ilu0(off, cols, vals)
        want = orc.ilu0_solve(L, U, b)
        got = S.ILU0(A).solve(dev(b), torch.zeros(n, dtype=torch.float64, device="cuda")).cpu().numpy()
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=0)


def test_ilu0_flow_and_ring_solves_agree(orc, monkeypatch):
    """The sync-free triangular solve (rows waiting on per-position stamps)
    and the level-by-level ring solve give the oracle's ILU(0) solve; with a
    spin limit of one poll the sync-free solve gives up on waiting rows and
    returns NaN instead of hanging."""
    off, cols, vals, _ = tnl.convdiff(200, seed=5)
    n = len(off) - 1
    b = np.random.default_rng(43).standard_normal(n)
    L, U = orc.ilu0(off, cols, vals)
    want = orc.ilu0_solve(L, U, b)
    got = {}
    for flow in ("1", "0"):
        monkeypatch.setenv("TNLMI_ILU_FLOW", flow)
        A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals)
        P = S.ILU0(A)
        got[flow] = P.solve(dev(b), torch.zeros(n, dtype=torch.float64, device="cuda")).cpu().numpy()
        np.testing.assert_allclose(got[flow], want, rtol=1e-12, atol=1e-13 * np.abs(want).max())
    np.testing.assert_array_equal(got["1"], got["0"])
    monkeypatch.setenv("TNLMI_ILU_FLOW", "1")
    P = S.ILU0(tnl.SparseMatrix.from_csr(n, n, off, cols, vals))
    monkeypatch.setenv("TNLMI_ILU_SPIN_CAP", "1")
    bad = P.solve(dev(b), torch.zeros(n, dtype=torch.float64, device="cuda")).cpu().numpy()
    assert np.isnan(bad).any()
    monkeypatch.delenv("TNLMI_ILU_SPIN_CAP")
    again = P.solve(dev(b), torch.zeros(n, dtype=torch.float64, device="cuda")).cpu().numpy()
    np.testing.assert_array_equal(again, got["1"])


def test_blas1(orc):
    rng = np.random.default_rng(5)
    for n in [1, 1000, 3_000_001]:
        a, c = rng.standard_normal(n), rng.standard_normal(n)
        assert abs(S.dot(dev(a), dev(c)) - float(np.dot(a, c))) <= 1e-12 * (np.abs(a) @ np.abs(c))
        assert abs(S.l2Norm(dev(a)) - float(np.linalg.norm(a))) <= 1e-13 * np.linalg.norm(a)


def test_solver_example_5x5():
    # Documentation/Examples/Solvers/Linear/IterativeLinearSolverExample.cpp
    off = np.array([0, 2, 5, 8, 11, 13], np.int32)
    cols = np.array([0, 1, 0, 1, 2, 1, 2, 3, 2, 3, 4, 3, 4], np.int32)
    vals = np.array([2.5, -1, -1, 2.5, -1, -1, 2.5, -1, -1, 2.5, -1, -1, 2.5])
    A = tnl.SparseMatrix.from_csr(5, 5, off, cols, vals)
    b = torch.zeros(5, dtype=torch.float64, device="cuda")
    A.vectorProduct(dev(np.ones(5)), b)
    for solver in (S.CG(), S.GMRES()):
        solver.setMatrix(A)
        solver.setConvergenceResidue(1e-12)
        x = torch.zeros(5, dtype=torch.float64, device="cuda")
        assert solver.solve(b, x)
        np.testing.assert_allclose(x.cpu().numpy(), np.ones(5), rtol=1e-10)


@pytest.mark.parametrize("vt", ["f32", "bf16"])
def test_mixed_precision_cg(orc, vt):
    # 3D Poisson 24^3: values 6 and -1 are exact in BF16, so both inner
    # matrices equal A; the FP64 refinement must reach the same true residual
    # as plain FP64 CG
    off, cols, vals = tnl.poisson3d(24, 24, 24)
    n = len(off) - 1
    b = np.random.default_rng(11).standard_normal(n)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals)
    Alo = tnl.SparseMatrix.from_csr(n, n, off, cols, vals,
                                    value_dtype=torch.float32 if vt == "f32" else torch.bfloat16)
    s = S.MixedCG()
    s.setMatrix(A)
    s.setInnerMatrix(Alo)
    s.setConvergenceResidue(1e-10)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert s.solve(dev(b), x)
    r = orc.true_residual(off, cols, vals, x.cpu().numpy(), b)
    assert r <= 1e-10 and abs(r - s.getResidue()) < 1e-12
    xo, _ = orc.cg(off, cols, vals, b, precond=1, tol=1e-12)
    np.testing.assert_allclose(x.cpu().numpy(), xo, rtol=1e-8, atol=1e-9)
    assert 1 <= s.getIterations() <= 8 and s.getInnerIterations() > 0


def test_mixed_precision_bf16_rounded_values(orc):
    # values that BF16 rounds: the refinement still converges to the FP64 solution
    off, cols, vals = tnl.laplace2d(30, 30)
    vals = vals * 1.0137  # 4.0548 / -1.0137 are not BF16 numbers
    n = len(off) - 1
    b = np.random.default_rng(12).standard_normal(n)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals)
    Ab = tnl.SparseMatrix.from_csr(n, n, off, cols, vals, value_dtype=torch.bfloat16)
    s = S.MixedCG()
    s.setMatrix(A)
    s.setInnerMatrix(Ab)
    s.setConvergenceResidue(1e-10)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert s.solve(dev(b), x)
    assert orc.true_residual(off, cols, vals, x.cpu().numpy(), b) <= 1e-10


def test_cg_iteration_limit_matches_tnl(orc):
    # the device-side stopping rule counts iterations as IterativeSolver::nextIteration does
    off, cols, vals = tnl.laplace2d(32, 32)
    n = len(off) - 1
    b = np.random.default_rng(5).standard_normal(n)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals)
    for maxit, minit in ((5, 0), (13, 0), (40, 30)):
        xo, so = orc.cg(off, cols, vals, b, precond=1, tol=1e-3, maxit=maxit, minit=minit)
        cg = S.CG()
        cg.setMatrix(A)
        cg.setPreconditioner(S.Diagonal(A))
        cg.setConvergenceResidue(1e-3)
        cg.setMaxIterations(maxit)
        cg.setMinIterations(minit)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        ok = cg.solve(dev(b), x)
        assert ok == bool(so.converged) and cg.getIterations() == so.iterations, (maxit, minit)
        assert abs(cg.getResidue() - so.residue) <= 1e-9 * max(1.0, so.residue)
        np.testing.assert_allclose(x.cpu().numpy(), xo, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("vt", ["f32", "bf16"])
@pytest.mark.parametrize("fmt,param", [(tnl.CSR, 0), (tnl.ELLPACK, 0), (tnl.SLICED_ELLPACK, 32),
                                       (tnl.BIELLPACK, 32)])
def test_reliable_mixed_cg(orc, vt, fmt, param):
    # FP32 CG with reliable updates: the FP64 true residual reaches 1e-10 in
    # about as many iterations as FP64 CG (the Krylov process is not restarted)
    off, cols, vals = tnl.poisson3d(20, 18, 16)
    n = len(off) - 1
    b = np.random.default_rng(21).standard_normal(n)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=fmt, param=param)
    Alo = tnl.SparseMatrix(A.rows, A.columns, A.segments,
                           A.values.to(torch.float32 if vt == "f32" else torch.bfloat16), A.column_indexes)
    s = S.ReliableMixedCG()
    s.setMatrix(A)
    s.setInnerMatrix(Alo)
    s.setConvergenceResidue(1e-10)
    s.setMaxIterations(3000)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert s.solve(dev(b), x)
    r = orc.true_residual(off, cols, vals, x.cpu().numpy(), b)
    assert r <= 1e-10 and abs(r - s.getResidue()) < 1e-13
    xo, so = orc.cg(off, cols, vals, b, precond=1, tol=1e-12)
    np.testing.assert_allclose(x.cpu().numpy(), xo, rtol=1e-8, atol=1e-9)
    assert s.getIterations() <= 1.25 * so.iterations + 10, (s.getIterations(), so.iterations)
    assert 2 <= s.getUpdates() <= 40


@pytest.mark.parametrize("fmt,param", [(tnl.CSR, 0), (tnl.ELLPACK, 0), (tnl.SLICED_ELLPACK, 32)])
@pytest.mark.parametrize("pre", [False, True], ids=["none", "jacobi"])
def test_true_residual_cg(orc, fmt, param, pre):
    # CG's iteration stopped on ||b - A x|| / ||b|| (the mixed solvers' rule):
    # the true residual meets the tolerance, x is the oracle's CG solution,
    # and it stops no later than an oracle CG run to the same true residual
    # needs (the oracle stops on the preconditioned residue, pinned to TNL's
    # CG; its iterate count at a tolerance tight enough to reach 1e-10 true
    # bounds this one); 8-bit column codes give the identical solve
    off, cols, vals = tnl.poisson3d(20, 18, 16)
    n = len(off) - 1
    b = np.random.default_rng(31).standard_normal(n)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=fmt, param=param)
    out = []
    for codes in (False, True):
        s = S.TrueResidualCG(column_codes=codes)
        s.setMatrix(A)
        if pre:
            s.setPreconditioner(S.Diagonal(A))
        s.setConvergenceResidue(1e-10)
        s.setMaxIterations(5000)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        assert s.solve(dev(b), x)
        xs = x.cpu().numpy()
        r = orc.true_residual(off, cols, vals, xs, b)
        assert r <= 1e-10 and abs(r - s.getResidue()) <= 1e-13
        out.append((xs, s.getIterations()))
    assert np.array_equal(out[0][0], out[1][0]) and out[0][1] == out[1][1]
    xo, so = orc.cg(off, cols, vals, b, precond=1 if pre else 0, tol=1e-12)
    np.testing.assert_allclose(out[0][0], xo, rtol=1e-7, atol=1e-9)
    assert out[0][1] <= so.iterations
    # the recurrence residual stops the solve: a looser tolerance stops earlier
    s = S.TrueResidualCG()
    s.setMatrix(A)
    s.setConvergenceResidue(1e-4)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert s.solve(dev(b), x)
    assert s.getIterations() < out[0][1] and orc.true_residual(off, cols, vals, x.cpu().numpy(), b) <= 1e-4


@pytest.mark.parametrize("fmt,param", [(tnl.ELLPACK, 0), (tnl.SLICED_ELLPACK, 32)])
def test_mixed_cg_column_codes_identical(monkeypatch, fmt, param):
    # the inner SpMV of an Ellpack / SlicedEllpack stencil matrix reads 8-bit
    # column codes (column = row + dictionary[code]); the columns it decodes
    # are the stored ones, so the solve is bit-identical to int32 columns
    off, cols, vals = tnl.poisson3d(20, 18, 16)
    n = len(off) - 1
    b = np.random.default_rng(23).standard_normal(n)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=fmt, param=param)
    Alo = tnl.SparseMatrix(A.rows, A.columns, A.segments, A.values.to(torch.bfloat16), A.column_indexes)
    out = []
    for codes in ("1", "0"):
        monkeypatch.setenv("TNLMI_COLUMN_CODES", codes)
        s = S.ReliableMixedCG()
        s.setMatrix(A)
        s.setInnerMatrix(Alo)
        s.setConvergenceResidue(1e-10)
        s.setMaxIterations(3000)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        assert s.solve(dev(b), x)
        out.append((x.cpu().numpy(), s.getIterations(), s.getUpdates()))
    assert np.array_equal(out[0][0], out[1][0]) and out[0][1:] == out[1][1:]


def _cg_solve(A, b, inner=None, its=3000):
    if inner is None:
        s = S.CG()
        s.setMatrix(A)
        s.setPreconditioner(S.Diagonal(A))
    else:
        s = S.ReliableMixedCG()
        s.setMatrix(A)
        s.setInnerMatrix(inner)
    s.setConvergenceResidue(1e-10)
    s.setMaxIterations(its)
    x = torch.zeros(A.rows, dtype=torch.float64, device="cuda")
    assert s.solve(dev(b), x)
    return x.cpu().numpy(), s.getIterations()


@pytest.mark.parametrize("vt", [torch.float64, torch.float32, torch.bfloat16], ids=["f64", "f32", "bf16"])
def test_cg_csr_stage_copy_matches_ellpack(orc, monkeypatch, vt):
    # a CSR matrix whose rows are short and about equally long runs the fused
    # Ellpack kernels on a column-major copy of its slots made by the solve:
    # the result equals the same matrix stored as Ellpack bit for bit; with
    # the copy switched off (the one-launch CSR kernel) it agrees to the
    # solver's tolerance
    off, cols, vals = tnl.poisson3d(19, 18, 16)  # 5,472 rows: not a multiple of 64
    n = len(off) - 1
    b = np.random.default_rng(31).standard_normal(n)
    mats = {}
    for fmt in (tnl.CSR, tnl.ELLPACK):
        A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=fmt)
        lo = None if vt == torch.float64 else tnl.SparseMatrix(A.rows, A.columns, A.segments, A.values.to(vt),
                                                                A.column_indexes)
        mats[fmt] = (A, lo)
    x_csr, it_csr = _cg_solve(*mats[tnl.CSR][:1], b, inner=mats[tnl.CSR][1])
    x_ell, it_ell = _cg_solve(*mats[tnl.ELLPACK][:1], b, inner=mats[tnl.ELLPACK][1])
    assert it_csr == it_ell and np.array_equal(x_csr, x_ell)
    monkeypatch.setenv("TNLMI_CG_CSR_COPY", "0")
    x_one, it_one = _cg_solve(*mats[tnl.CSR][:1], b, inner=mats[tnl.CSR][1])
    assert abs(it_one - it_csr) <= 2
    np.testing.assert_allclose(x_one, x_csr, rtol=1e-7, atol=1e-9)
    assert orc.true_residual(off, cols, vals, x_one, b) <= 1e-9


def test_cg_csr_uneven_rows_one_launch_kernel(orc):
    # rows of 1..~20 entries: an Ellpack copy would be much larger than the
    # CSR arrays, so the solve keeps the one-launch CSR kernel
    import scipy.sparse as sp
    n = 4000
    R = sp.random(n, n, density=3.0 / n, random_state=np.random.RandomState(5), format="csr")
    R.data[:] = -np.abs(R.data)
    M = (R + R.T).tocsr()
    M = (M + sp.diags(np.asarray(abs(M).sum(axis=1)).ravel() + 1.0)).tocsr()
    M.sort_indices()
    lens = np.diff(M.indptr)
    assert lens.max() <= 32 and lens.max() * n > 1.2 * M.nnz
    off, cols, vals = M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data.astype(np.float64)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals)
    b = np.random.default_rng(32).standard_normal(n)
    x, its = _cg_solve(A, b)
    xo, so = orc.cg(off, cols, vals, b, precond=1, tol=1e-10)
    assert abs(its - so.iterations) <= 1
    np.testing.assert_allclose(x, xo, rtol=1e-7, atol=1e-9)


def test_mixed_cg_many_offsets_keeps_int32_columns(orc):
    # more than 255 distinct offsets column - row: no codes, the int32 kernel
    import scipy.sparse as sp
    n = 3000
    R = sp.random(n, n, density=4.0 / n, random_state=np.random.RandomState(3), format="csr")
    R.data[:] = -np.abs(R.data)
    M = (R + R.T).tocsr()
    M = (M + sp.diags(np.asarray(abs(M).sum(axis=1)).ravel() + 1.0)).tocsr()
    M.sort_indices()
    assert len(np.unique(M.nonzero()[1] - M.nonzero()[0])) > 255
    off, cols, vals = M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data.astype(np.float64)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=tnl.ELLPACK)
    Alo = tnl.SparseMatrix(A.rows, A.columns, A.segments, A.values.to(torch.float32), A.column_indexes)
    b = np.random.default_rng(24).standard_normal(n)
    s = S.ReliableMixedCG()
    s.setMatrix(A)
    s.setInnerMatrix(Alo)
    s.setConvergenceResidue(1e-10)
    s.setMaxIterations(3000)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert s.solve(dev(b), x)
    assert orc.true_residual(off, cols, vals, x.cpu().numpy(), b) <= 1e-10


def test_reliable_mixed_cg_rounded_bf16(orc):
    off, cols, vals = tnl.laplace2d(40, 40)
    vals = vals * 1.0137  # not BF16 numbers: the inner matrix differs from A
    n = len(off) - 1
    b = np.random.default_rng(22).standard_normal(n)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals)
    Ab = tnl.SparseMatrix.from_csr(n, n, off, cols, vals, value_dtype=torch.bfloat16)
    s = S.ReliableMixedCG()
    s.setMatrix(A)
    s.setInnerMatrix(Ab)
    s.setConvergenceResidue(1e-10)
    s.setMaxIterations(5000)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert s.solve(dev(b), x)
    assert orc.true_residual(off, cols, vals, x.cpu().numpy(), b) <= 1e-10


def test_reliable_mixed_cg_errors():
    off, cols, vals = tnl.laplace2d(8, 8)
    n = len(off) - 1
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals)
    s = S.ReliableMixedCG()
    s.setMatrix(A)
    with pytest.raises(tnl._lib.LogicError):
        s.solve(dev(np.ones(n)), torch.zeros(n, dtype=torch.float64, device="cuda"))
    s.setInnerMatrix(A)  # FP64 values: not an inner matrix
    with pytest.raises(ValueError):
        s.solve(dev(np.ones(n)), torch.zeros(n, dtype=torch.float64, device="cuda"))
    s.setInnerMatrix(tnl.SparseMatrix.from_csr(n, n, off, cols, vals, value_dtype=torch.float32))
    s.setDelta(1.5)
    with pytest.raises(ValueError):
        s.solve(dev(np.ones(n)), torch.zeros(n, dtype=torch.float64, device="cuda"))
    # zero right-hand side: x = 0, converged without iterations
    s.setDelta(0.1)
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    assert s.solve(torch.zeros(n, dtype=torch.float64, device="cuda"), x)
    assert s.getIterations() == 0 and float(x.abs().max()) == 0.0


@pytest.mark.parametrize("fmt,param", [(tnl.ELLPACK, 0), (tnl.SLICED_ELLPACK, 32), (tnl.SLICED_ELLPACK, 2),
                                       (tnl.CSR, 0)])
def test_fused_cg_ragged_rows(orc, fmt, param):
    # 1073 rows: the multi-row SpMV kernels end inside a row group
    off, cols, vals = tnl.laplace2d(37, 29)
    n = len(off) - 1
    b = np.random.default_rng(23).standard_normal(n)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=fmt, param=param)
    cg = S.CG()
    cg.setMatrix(A)
    cg.setPreconditioner(S.Diagonal(A))
    cg.setConvergenceResidue(1e-10)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert cg.solve(dev(b), x)
    xo, so = orc.cg(off, cols, vals, b, precond=1, tol=1e-10)
    assert abs(cg.getIterations() - so.iterations) <= 1
    np.testing.assert_allclose(x.cpu().numpy(), xo, rtol=1e-7, atol=1e-9)
    m = S.ReliableMixedCG()
    m.setMatrix(A)
    m.setInnerMatrix(tnl.SparseMatrix(A.rows, A.columns, A.segments, A.values.float(), A.column_indexes))
    m.setConvergenceResidue(1e-10)
    m.setMaxIterations(3000)
    x.zero_()
    assert m.solve(dev(b), x)
    assert orc.true_residual(off, cols, vals, x.cpu().numpy(), b) <= 1e-10


# ------------------------------------------------------- BiCGStab / TFQMR
@pytest.mark.parametrize("solver", ["bicgstab", "bicgstab_exact", "tfqmr"])
@pytest.mark.parametrize("pre", [0, 1, 2], ids=["none", "diag", "ilu0"])
@pytest.mark.parametrize("fmt", [tnl.CSR, tnl.ELLPACK], ids=["csr", "ell"])
def test_bicgstab_tfqmr_match_oracle(orc, solver, pre, fmt):
    # BICGStab.hpp:33-127 / TFQMR.hpp:12-121 on the GPU: the oracle's iterations
    # (+-1: reduction order) and solution, and a true residual at the tolerance
    off, cols, vals = nonsym(600, 21)
    n = len(off) - 1
    b = np.random.default_rng(22).standard_normal(n)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=fmt)
    if solver == "tfqmr":
        s = S.TFQMR()
        xo, so = orc.tfqmr(off, cols, vals, b, precond=pre, tol=1e-10)
    else:
        s = S.BICGStab()
        s.setup(exact_residue=solver == "bicgstab_exact")
        xo, so = orc.bicgstab(off, cols, vals, b, precond=pre, exact=solver == "bicgstab_exact", tol=1e-10)
    s.setMatrix(A)
    if pre:
        s.setPreconditioner(S.Diagonal(A) if pre == 1 else S.ILU0(A))
    s.setConvergenceResidue(1e-10)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert s.solve(dev(b), x) and so.converged
    assert abs(s.getIterations() - so.iterations) <= 1, (s.getIterations(), so.iterations)
    np.testing.assert_allclose(x.cpu().numpy(), xo, rtol=1e-7, atol=1e-8)
    assert orc.true_residual(off, cols, vals, x.cpu().numpy(), b) < 1e-8


@pytest.mark.parametrize("solver", ["bicgstab", "tfqmr"])
def test_bicgstab_tfqmr_limits_and_zero_rhs(orc, solver):
    off, cols, vals = nonsym(300, 23)
    n = len(off) - 1
    b = np.random.default_rng(24).standard_normal(n)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals)
    for maxit, minit in ((3, 0), (8, 0), (12, 10)):
        s = S.TFQMR() if solver == "tfqmr" else S.BICGStab()
        s.setMatrix(A)
        s.setConvergenceResidue(1e-3 if minit else 1e-14)
        s.setMaxIterations(maxit)
        s.setMinIterations(minit)
        fn = orc.tfqmr if solver == "tfqmr" else orc.bicgstab
        xo, so = fn(off, cols, vals, b, tol=1e-3 if minit else 1e-14, maxit=maxit, minit=minit)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        ok = s.solve(dev(b), x)
        assert ok == bool(so.converged) and s.getIterations() == so.iterations, (maxit, minit)
        np.testing.assert_allclose(x.cpu().numpy(), xo, rtol=1e-9, atol=1e-11)
    s = S.TFQMR() if solver == "tfqmr" else S.BICGStab()
    s.setMatrix(A)
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    assert s.solve(torch.zeros(n, dtype=torch.float64, device="cuda"), x)
    assert float(x.abs().max()) == 0.0 and s.getIterations() == 0


def test_bicgstab_tfqmr_convection_diffusion():
    # the nonsymmetric alternatives to GMRES on the convection-diffusion class
    off, cols, vals, _ = tnl.convdiff(64, seed=3)
    n = len(off) - 1
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals)
    b = torch.from_numpy(np.random.default_rng(25).standard_normal(n)).cuda()
    for s in (S.BICGStab(), S.TFQMR()):
        s.setMatrix(A)
        s.setPreconditioner(S.ILU0(A))
        s.setConvergenceResidue(1e-10)
        s.setMaxIterations(5000)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        assert s.solve(b, x), type(s).__name__
        assert S.true_residual(A, x, b) < 1e-8


def scaled_nonsym(n, seed):
    off, cols, vals = nonsym(n, seed)
    scale = 1.0 + (np.arange(n) % 5)
    return off, cols, vals * np.repeat(scale, np.diff(off))


@pytest.mark.parametrize("fmt", [tnl.CSR, tnl.ELLPACK], ids=["csr", "ellpack"])
@pytest.mark.parametrize("pre", [0, 1, 2], ids=["none", "diag", "ilu0"])
@pytest.mark.parametrize("variant", ["bicgstabl1", "bicgstabl2", "bicgstabl4exact", "idr1", "idr4", "idr4smooth",
                                     "idr8smooth"])
def test_bicgstabl_idrs_match_oracle(orc, variant, pre, fmt):
    # BICGStabL.hpp:37-257 / IDRs.hpp:47-386 on the GPU: the oracle's iterations
    # (+-1: the GPU sums inner products in another order) and solution
    off, cols, vals = scaled_nonsym(700, 31)
    n = len(off) - 1
    b = np.random.default_rng(32).standard_normal(n)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=fmt)
    if variant.startswith("bicgstabl"):
        ell, exact = int(variant[9]), variant.endswith("exact")
        s = S.BICGStabL()
        s.setup(ell=ell, exact_residue=exact)
        xo, so = orc.bicgstabl(off, cols, vals, b, precond=pre, ell=ell, exact=exact, tol=1e-10)
    else:
        sd, sm = int(variant[3]), variant.endswith("smooth")
        s = S.IDRs()
        s.setShadowSpaceDimension(sd)
        s.setResidualSmoothing(sm)
        xo, so = orc.idrs(off, cols, vals, b, precond=pre, s=sd, smoothing=sm, tol=1e-10)
    s.setMatrix(A)
    if pre:
        s.setPreconditioner(S.Diagonal(A) if pre == 1 else S.ILU0(A))
    s.setConvergenceResidue(1e-10)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert s.solve(dev(b), x) and so.converged
    tol_it = 1 if variant.startswith("idr") else int(variant[9])  # BiCGStab(ell) stops every ell steps
    assert abs(s.getIterations() - so.iterations) <= tol_it, (s.getIterations(), so.iterations)
    np.testing.assert_allclose(x.cpu().numpy(), xo, rtol=1e-7, atol=1e-8)
    assert orc.true_residual(off, cols, vals, x.cpu().numpy(), b) < 1e-8


@pytest.mark.parametrize("solver", ["bicgstabl", "idrs"])
def test_bicgstabl_idrs_limits_zero_rhs_nan(orc, solver):
    off, cols, vals = scaled_nonsym(300, 33)
    n = len(off) - 1
    b = np.random.default_rng(34).standard_normal(n)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals)

    def make():
        s = S.BICGStabL() if solver == "bicgstabl" else S.IDRs()
        if solver == "bicgstabl":
            s.setup(ell=2)
        s.setMatrix(A)
        return s
    fn = (lambda **kw: orc.bicgstabl(off, cols, vals, b, ell=2, **kw)) if solver == "bicgstabl" else \
        (lambda **kw: orc.idrs(off, cols, vals, b, **kw))
    for maxit in (3, 9):
        s = make()
        s.setConvergenceResidue(1e-14)
        s.setMaxIterations(maxit)
        xo, so = fn(tol=1e-14, maxit=maxit)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        ok = s.solve(dev(b), x)
        assert ok == bool(so.converged) and s.getIterations() == so.iterations, maxit
        np.testing.assert_allclose(x.cpu().numpy(), xo, rtol=1e-7, atol=1e-10)
    s = make()
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    assert s.solve(torch.zeros(n, dtype=torch.float64, device="cuda"), x)
    assert float(x.abs().max()) == 0.0 and s.getIterations() == 0
    bn = b.copy()
    bn[7] = np.nan
    with pytest.raises(tnl.TnlmiError, match="initial residue is NAN"):
        make().solve(dev(bn), torch.zeros(n, dtype=torch.float64, device="cuda"))
