"""Learned sparse approximate inverse on the GPU against the CPU oracle
(oracle/spai_oracle.py, whose gradients are pinned by finite differences in
tests/test_spai_oracle.py): the MFMA dense layer, the forward pass, the probe
loss and every parameter gradient; then what training gains over the Jacobi
inverse the untrained network starts from, and GMRES with the fallback."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
tnl = pytest.importorskip("tnl_amd")
from tnl_amd import solvers as S  # noqa: E402
from tnl_amd import spai  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import spai_oracle as SO  # noqa: E402


def bf16_bits(a):
    a32 = np.asarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return (u >> 16).astype(np.uint16)


def live_params(seed):
    # a network whose every layer matters: random biases and decoder
    rng = np.random.default_rng(seed)
    p = SO.unflatten(spai.init_parameters(seed).astype(np.float64))
    p["wo"] = rng.standard_normal(64) * 0.05
    p["bo"] = np.array([0.02])
    for l in range(3):
        p[f"be{l}"] = rng.standard_normal(64) * 0.1
        p[f"bh{l}"] = rng.standard_normal(64) * 0.1
    return SO.flatten(p).astype(np.float32)


def test_layout_matches_oracle():
    assert [(n, s) for n, s in spai.layout()] == [(n, s) for n, s in SO.param_shapes()]
    assert len(spai.init_parameters(0)) == sum(int(np.prod(s)) for _, s in SO.param_shapes())


@pytest.mark.parametrize("rows", [1, 64, 173, 4096])
def test_bf16_gemm64(rows):
    rng = np.random.default_rng(rows)
    X = rng.standard_normal((rows, 64))
    W = rng.standard_normal((64, 64))
    Xb, Wb = bf16_bits(X), bf16_bits(W)
    Y = spai.bf16_gemm64(torch.from_numpy(Xb.view(np.int16)).cuda(), torch.from_numpy(Wb.view(np.int16)).cuda())
    ref = SO.bf16(X) @ SO.bf16(W)
    got = Y.cpu().numpy()
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()


def instance(m, seed):
    off, cols, vals, _ = tnl.convdiff(m, seed=seed)
    n = len(off) - 1
    return off, cols, vals, tnl.SparseMatrix.from_csr(n, n, off, cols, vals)


def test_forward_matches_oracle():
    off, cols, vals, A = instance(14, 4)
    flat = live_params(1)
    net = spai.Network(flat)
    M = net.inverse(A).values.cpu().numpy()
    Mo, _ = SO.forward(SO.unflatten(flat.astype(np.float64)), SO.graph(off, cols, vals), bf=True)
    # BF16 activations: an FP32 accumulation that lands on the other side of a
    # rounding boundary moves one value by one BF16 ulp
    assert np.abs(M - Mo).max() <= 2e-2 * np.abs(Mo).max()
    assert np.abs(M - Mo).mean() <= 1e-3 * np.abs(Mo).mean()


def test_untrained_network_is_jacobi():
    off, cols, vals, A = instance(16, 5)
    net = spai.Network(spai.init_parameters(3))
    M = net.inverse(A).values.cpu().numpy()
    g = SO.graph(off, cols, vals)
    assert np.array_equal(M, np.where(g["isd"], 1.0 / g["diag"][g["row"]], 0.0))


def test_gradients_match_oracle():
    off, cols, vals, A = instance(14, 6)
    n = len(off) - 1
    flat = live_params(2)
    net = spai.Network(flat)
    Z = np.random.default_rng(9).integers(0, 2, (3, n)) * 2.0 - 1.0
    L, _ = net.loss_grad(A, torch.from_numpy(Z).cuda())
    g = net.gradients().astype(np.float64)
    Lo, go, _ = SO.loss_and_grad(flat.astype(np.float64), off, cols, vals, Z, bf=True)
    assert abs(L - Lo) <= 1e-3 * Lo
    k = 0
    for name, shape in SO.param_shapes():
        size = int(np.prod(shape))
        a, b = g[k:k + size], go[k:k + size]
        k += size
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        assert err <= 3e-2, (name, err, np.linalg.norm(b))


def test_training_gains_over_jacobi():
    # a few hundred Adam steps on small instances: the probe loss and the
    # GMRES(30) iteration count on an unseen instance both fall below those
    # of the untrained network (= the Jacobi inverse)
    net = spai.Network(seed=0)
    A0 = spai.convdiff_matrix(40, 777)[0]
    Z = torch.from_numpy(np.random.default_rng(1).integers(0, 2, (8, A0.rows)) * 2.0 - 1.0).cuda()
    L_jacobi, _ = net.loss_grad(A0, Z)
    hist = spai.train(net, sizes=(32, 40), instances=300, steps_per_instance=1, probes=4, lr=3e-3, seed=50)
    L_trained, _ = net.loss_grad(A0, Z)
    assert np.isfinite(hist).all() and L_trained < 0.85 * L_jacobi, (L_jacobi, L_trained)
    b = torch.from_numpy(np.random.default_rng(2).standard_normal(A0.rows)).cuda()
    M = net.inverse(A0)
    its = {}
    for name, pre in (("jacobi", S.Diagonal(A0)), ("learned", S.Explicit(M))):
        x = torch.zeros(A0.rows, dtype=torch.float64, device="cuda")
        ok, it, r = spai.gmres_solve(A0, b, x, pre, tolerance=1e-10)
        assert ok and r <= 1e-10
        its[name] = it
    assert its["learned"] < its["jacobi"], its


def test_fallback_on_stall():
    off, cols, vals, A = instance(24, 8)
    n = len(off) - 1
    b = torch.from_numpy(np.random.default_rng(3).standard_normal(n)).cuda()
    # a useless "inverse": M = 0 (GMRES on M A x = M b sees a zero right-hand side)
    bad = tnl.SparseMatrix(A.rows, A.columns, A.segments, torch.zeros_like(A.values), A.column_indexes)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    ok, its, r, used = spai.gmres_with_fallback(A, b, x, S.Explicit(bad), S.ILU0(A), tolerance=1e-10,
                                                max_iterations=20_000, stall_cycles=3)
    assert ok and r <= 1e-10 and used
    assert S.true_residual(A, x, b) <= 1e-10


def test_errors():
    off, cols, vals = tnl.laplace2d(6, 6)
    n = len(off) - 1
    net = spai.Network(seed=0)
    E = tnl.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=tnl.ELLPACK)
    with pytest.raises(ValueError):
        net.inverse(E)  # not CSR
    F = tnl.SparseMatrix.from_csr(n, n, off, cols, vals, value_dtype=torch.float32)
    with pytest.raises(ValueError):
        net.inverse(F)  # not FP64
