"""The learned-SPAI oracle (oracle/spai_oracle.py) pinned on the CPU: its
analytic gradients against central finite differences of its own loss (FP64,
no BF16 rounding), the loss against a dense evaluation of ||(I - M A) Z^T||,
and the zero decoder giving exactly the Jacobi inverse."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import spai_oracle as SO  # noqa: E402

tnl = pytest.importorskip("tnl_amd")


def small(m=6, seed=3):
    off, cols, vals, _ = tnl.convdiff(m, seed=seed)
    return off, cols, vals


def rademacher(k, n, seed):
    return np.random.default_rng(seed).integers(0, 2, (k, n)) * 2.0 - 1.0


@pytest.mark.parametrize("hidden,layers", [(8, 2), (16, 3)])
def test_gradients_match_finite_differences(hidden, layers):
    off, cols, vals = small()
    n = len(off) - 1
    Z = rademacher(3, n, 1)
    flat = SO.init_params(7, hidden, layers)
    rng = np.random.default_rng(2)
    p = SO.unflatten(flat, hidden, layers)
    p["wo"] = rng.standard_normal(hidden) * 0.3  # a live decoder so every layer gets a gradient
    p["bo"] = np.array([0.05])
    for l in range(layers):
        p[f"be{l}"] = rng.standard_normal(hidden) * 0.1
        p[f"bh{l}"] = rng.standard_normal(hidden) * 0.1
    flat = SO.flatten(p, hidden, layers)
    L, g, _ = SO.loss_and_grad(flat, off, cols, vals, Z, bf=False, hidden=hidden, layers=layers)
    idx = rng.choice(len(flat), 60, replace=False)
    # every parameter group is probed at least once
    names = [nm for nm, shape in SO.param_shapes(hidden, layers) for _ in range(int(np.prod(shape)))]
    extra = [names.index(nm) for nm, _ in SO.param_shapes(hidden, layers)]
    h = 1e-6
    worst = 0.0
    for i in list(idx) + extra:
        fp, fm = flat.copy(), flat.copy()
        fp[i] += h
        fm[i] -= h
        Lp = SO.loss_and_grad(fp, off, cols, vals, Z, bf=False, hidden=hidden, layers=layers)[0]
        Lm = SO.loss_and_grad(fm, off, cols, vals, Z, bf=False, hidden=hidden, layers=layers)[0]
        fd = (Lp - Lm) / (2 * h)
        err = abs(fd - g[i]) / max(1e-4 * abs(L), abs(fd), abs(g[i]))
        worst = max(worst, err)
        assert err < 2e-4, (names[i], i, fd, g[i])
    assert worst < 2e-4


def test_loss_matches_dense_evaluation():
    off, cols, vals = small(5, 4)
    n = len(off) - 1
    Z = rademacher(2, n, 3)
    flat = SO.init_params(1, 8, 2)
    p = SO.unflatten(flat, 8, 2)
    p["wo"] = np.linspace(-0.2, 0.3, 8)
    flat = SO.flatten(p, 8, 2)
    L, _, M = SO.loss_and_grad(flat, off, cols, vals, Z, bf=False, hidden=8, layers=2)
    A = np.zeros((n, n))
    Md = np.zeros((n, n))
    row = np.repeat(np.arange(n), np.diff(off))
    A[row, cols] = vals
    Md[row, cols] = M
    R = (np.eye(n) - Md @ A) @ Z.T
    assert abs(L - (R ** 2).sum() / (Z.shape[0] * n)) < 1e-12 * max(1.0, L)


def test_zero_decoder_is_jacobi():
    off, cols, vals = small(7, 5)
    flat = SO.init_params(0)
    g = SO.graph(off, cols, vals)
    M, _ = SO.forward(SO.unflatten(flat), g, bf=True)
    expect = np.where(g["isd"], 1.0 / g["diag"][g["row"]], 0.0)
    assert np.array_equal(M, expect)


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 65504.0, 1e-3])
    r = SO.bf16(x)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even -> 1
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125
    assert abs(r[3] + 3.140625) < 1e-12
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8)
