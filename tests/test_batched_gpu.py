"""Batched CG (one persistent launch over many independent SPD systems)
against the oracle's per-system Jacobi-CG (CG.hpp:10-108 with the Diagonal
preconditioner, pinned to TNL's own CG in tests/test_oracle.py)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
tnl = pytest.importorskip("tnl_amd")
from tnl_amd import batched as B  # noqa: E402


def split(off, cols, vals, sys_rows, k):
    r0, r1 = int(sys_rows[k]), int(sys_rows[k + 1])
    o = off[r0:r1 + 1] - off[r0]
    c = cols[off[r0]:off[r1]] - r0
    v = vals[off[r0]:off[r1]]
    return o, c, v


@pytest.mark.parametrize("fmt,param", [(tnl.CSR, 0), (tnl.SLICED_ELLPACK, 64), (tnl.ELLPACK, 0), (tnl.BIELLPACK, 32)])
@pytest.mark.parametrize("block", [256, 1024])
@pytest.mark.parametrize("lanes", [1, 4])
def test_batched_cg_matches_per_system_cg(orc, fmt, param, block, lanes):
    sides, off, cols, vals, sys_rows = B.diffusion_batch(40, 4, 24, seed=3)
    n = int(sys_rows[-1])
    b = tnl.normal(n, seed=4)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=fmt, param=param)
    bc = B.BatchedCG(A, sys_rows, block_size=block, lanes_per_row=lanes)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    bd = torch.from_numpy(b).cuda()
    its, res = bc.solve(bd, x, 1e-10, use_graph=False)
    torch.cuda.synchronize()
    its, res, xs = its.cpu().numpy(), res.cpu().numpy(), x.cpu().numpy()
    for k in range(len(sides)):
        o, c, v = split(off, cols, vals, sys_rows, k)
        r0, r1 = int(sys_rows[k]), int(sys_rows[k + 1])
        xo, so = orc.cg(o, c, v, b[r0:r1], precond=1, tol=1e-10)
        assert so.converged and res[k] < 1e-10
        assert abs(int(its[k]) - so.iterations) <= 1, (k, its[k], so.iterations)
        np.testing.assert_allclose(xs[r0:r1], xo, rtol=1e-7, atol=1e-9)
        assert orc.true_residual(o, c, v, xs[r0:r1], b[r0:r1]) < 1e-8


def test_batched_graph_replays_and_mixed(orc):
    sides, off, cols, vals, sys_rows = B.diffusion_batch(60, 4, 30, seed=5)
    n = int(sys_rows[-1])
    b = torch.from_numpy(tnl.normal(n, seed=6)).cuda()
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals)
    A32 = tnl.SparseMatrix(A.rows, A.columns, A.segments, A.values.float(), A.column_indexes)
    bc = B.BatchedCG(A, sys_rows, A_lo=A32, block_size=512)
    x_eager = torch.zeros(n, dtype=torch.float64, device="cuda")
    bc.solve(b, x_eager, 1e-10, use_graph=False)
    torch.cuda.synchronize()
    ref_its = bc.iters.clone()
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    for _ in range(3):  # replays of one captured launch, eager work in between
        x.zero_()
        bc.solve(b, x, 1e-10, use_graph=True)
        torch.cuda.synchronize()
        assert torch.equal(x, x_eager) and torch.equal(bc.iters, ref_its)
    # mixed: FP32 inner CG, FP64 refinement on the true residual
    xm = torch.zeros(n, dtype=torch.float64, device="cuda")
    bc.solve(b, xm, 1e-10, mixed=True, inner_tolerance=1e-4, use_graph=True)
    torch.cuda.synchronize()
    xs, bh = xm.cpu().numpy(), b.cpu().numpy()
    res = bc.res.cpu().numpy()
    for k in range(len(sides)):
        o, c, v = split(off, cols, vals, sys_rows, k)
        r0, r1 = int(sys_rows[k]), int(sys_rows[k + 1])
        tr = orc.true_residual(o, c, v, xs[r0:r1], bh[r0:r1])
        assert tr < 1e-10 and abs(tr - res[k]) < 1e-12


def test_autotune_picks_a_candidate(tmp_path, monkeypatch):
    monkeypatch.setenv("TNL_AMD_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    sides, off, cols, vals, sys_rows = B.diffusion_batch(64, 8, 32, seed=7)
    best = B.autotune(off, cols, vals, sys_rows, sample=32, blocks=[256, 512], lanes=[1, 4])
    assert best["block"] in (256, 512) and best["format"] in ("CSR", "SlicedEllpack", "Ellpack")
    assert best["rows_per_wavefront"] in (64, 16)
    assert len(best["trials"]) == 16
    again = B.autotune(off, cols, vals, sys_rows, sample=32, blocks=[256, 512], lanes=[1, 4])
    assert again == best  # served from the cache
    other = B.autotune(off, cols, vals, sys_rows, sample=32, blocks=[256], lanes=[1, 4])
    assert len(other["trials"]) == 8  # other candidate lists are timed again, not served from the cache


@pytest.mark.parametrize("lanes", [2, 8])
def test_batched_mixed_rows_per_wavefront(orc, lanes):
    sides, off, cols, vals, sys_rows = B.diffusion_batch(24, 4, 20, seed=9)
    n = int(sys_rows[-1])
    bh = tnl.normal(n, seed=10)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=tnl.SLICED_ELLPACK, param=32)
    A32 = tnl.SparseMatrix(A.rows, A.columns, A.segments, A.values.float(), A.column_indexes)
    bc = B.BatchedCG(A, sys_rows, A_lo=A32, block_size=256, lanes_per_row=lanes)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    bc.solve(torch.from_numpy(bh).cuda(), x, 1e-10, mixed=True, inner_tolerance=1e-4, use_graph=False)
    torch.cuda.synchronize()
    xs = x.cpu().numpy()
    for k in range(len(sides)):
        o, c, v = split(off, cols, vals, sys_rows, k)
        r0, r1 = int(sys_rows[k]), int(sys_rows[k + 1])
        assert orc.true_residual(o, c, v, xs[r0:r1], bh[r0:r1]) < 1e-10


def test_batched_rejects_columns_outside_their_system():
    # the systems must be the diagonal blocks: a column of system 1 stored in
    # a row of system 0 is refused at create (the kernel would read vectors a
    # concurrent workgroup updates)
    sides, off, cols, vals, sys_rows = B.diffusion_batch(4, 4, 6, seed=1)
    n = int(sys_rows[-1])
    bad = cols.copy()
    bad[0] = int(sys_rows[1])  # row 0 belongs to system 0
    A = tnl.SparseMatrix.from_csr(n, n, off, bad, vals)
    with pytest.raises(tnl.InvalidArgument, match="outside its system"):
        B.BatchedCG(A, sys_rows)
    with pytest.raises(tnl.InvalidArgument, match="lanes per row"):
        B.BatchedCG(tnl.SparseMatrix.from_csr(n, n, off, cols, vals), sys_rows, lanes_per_row=3)
