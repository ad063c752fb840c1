"""CPU oracle of the learned sparse-approximate-inverse network (test
infrastructure only: imported by tests/ to check the GPU kernels; the product
never calls it).

No reference counterpart: TNL has no learned preconditioner.  This file is
the specification the HIP kernels in tnl_amd/csrc/spai.hip implement, in
FP64 numpy.  With `bf16=True` every value the GPU stores in BF16 (the
activations and the GEMM operands) is rounded to BF16 (round to nearest
even) at the same point, so the GPU results match to FP32-accumulation
rounding.  Its analytic gradients are pinned by central finite differences
in tests/test_spai_oracle.py (FP64, bf16=False).

The network, on the graph of A (an edge per nonzero (i, j), CSR order):
  edge input   x_e = [a_ij/a_ii, a_ij/a_jj, [i == j], |a_ij|/a_ii]
  node input   x_i = [sum_{j!=i} |a_ij|/a_ii, sum_j a_ij/a_ii, deg_i/5, 1]
  encoder      E = relu(x_e Ae + ae),  H = relu(x_i An + an)
  layer l      P = H Wu,  Q = H Wv
               E = relu(E We + P[col] + Q[row] + be)
               G = mean over the row's edges of E
               H = relu(H Wh + G Wg + bh)
  decoder      o_e = E wo + bo
  output       M_ij = (delta_ij + o_ij) / a_ii        (M = D^-1 (I + O))
Probe loss over K Rademacher vectors z_k (rows of Z):
  y = A z,  r = z - M y,  L = sum_k ||r_k||^2 / (K n)
"""
import numpy as np

HIDDEN = 64
LAYERS = 3
N_EDGE_IN = 4
N_NODE_IN = 4
LAYER_MATS = ("We", "Wu", "Wv", "Wh", "Wg")


def bf16(a):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    a32 = np.asarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def param_shapes(hidden=HIDDEN, layers=LAYERS):
    """(name, shape) in the order of the flat parameter vector."""
    out = [("Ae", (N_EDGE_IN, hidden)), ("ae", (hidden,)), ("An", (N_NODE_IN, hidden)), ("an", (hidden,))]
    for l in range(layers):
        for w in LAYER_MATS:
            out.append((f"{w}{l}", (hidden, hidden)))
        out.append((f"be{l}", (hidden,)))
        out.append((f"bh{l}", (hidden,)))
    out += [("wo", (hidden,)), ("bo", (1,))]
    return out


def unflatten(flat, hidden=HIDDEN, layers=LAYERS):
    p, k = {}, 0
    for name, shape in param_shapes(hidden, layers):
        size = int(np.prod(shape))
        p[name] = np.asarray(flat[k:k + size], np.float64).reshape(shape)
        k += size
    assert k == len(flat)
    return p


def flatten(p, hidden=HIDDEN, layers=LAYERS):
    return np.concatenate([np.asarray(p[name], np.float64).ravel() for name, _ in param_shapes(hidden, layers)])


def graph(off, cols, vals):
    """Per-edge row, diagonal entries, degrees and the raw inputs."""
    n = len(off) - 1
    deg = np.diff(off).astype(np.int64)
    row = np.repeat(np.arange(n), deg)
    diag = np.zeros(n)
    isd = cols == row
    diag[row[isd]] = vals[isd]
    xe = np.stack([vals / diag[row], vals / diag[cols], isd.astype(float), np.abs(vals) / diag[row]], 1)
    off_abs = np.bincount(row, weights=np.where(isd, 0.0, np.abs(vals)), minlength=n) / diag
    rsum = np.bincount(row, weights=vals, minlength=n) / diag
    xn = np.stack([off_abs, rsum, deg / 5.0, np.ones(n)], 1)
    return dict(n=n, row=row, col=np.asarray(cols, np.int64), diag=diag, deg=deg, xe=xe, xn=xn, isd=isd)


def forward(p, g, bf=True, layers=LAYERS):
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    relu = lambda a: np.maximum(a, 0.0)
    n, row, col, deg = g["n"], g["row"], g["col"], g["deg"]
    cache = {}
    E = r(relu(g["xe"] @ p["Ae"] + p["ae"]))
    H = r(relu(g["xn"] @ p["An"] + p["an"]))
    cache["E0"], cache["H0"] = E, H
    for l in range(layers):
        W = {w: r(p[f"{w}{l}"]) for w in LAYER_MATS}
        P = H @ W["Wu"]
        Q = H @ W["Wv"]
        S = E @ W["We"] + P[col] + Q[row] + p[f"be{l}"]
        E = r(relu(S))
        G = np.zeros((n, E.shape[1]))
        np.add.at(G, row, E)
        G = r(G / deg[:, None])
        T = H @ W["Wh"] + G @ W["Wg"] + p[f"bh{l}"]
        H = r(relu(T))
        cache[f"E{l + 1}"], cache[f"H{l + 1}"], cache[f"G{l + 1}"] = E, H, G
    o = E @ p["wo"] + p["bo"][0]
    M = (g["isd"].astype(float) + o) / g["diag"][row]
    return M, cache


def probe_loss(off, cols, vals, Mvals, Z):
    """L and dL/dM (per edge) for the probes Z (K x n)."""
    n = len(off) - 1
    deg = np.diff(off)
    row = np.repeat(np.arange(n), deg)
    K = Z.shape[0]
    L = 0.0
    dM = np.zeros(len(cols))
    for z in Z:
        y = np.bincount(row, weights=vals * z[cols], minlength=n)
        w = np.bincount(row, weights=Mvals * y[cols], minlength=n)
        res = z - w
        L += res @ res
        dM += -2.0 * res[row] * y[cols]
    return L / (K * n), dM / (K * n)


def backward(p, g, cache, dM, bf=True, layers=LAYERS):
    """Gradients of the loss wrt every parameter, given dL/dM per edge."""
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    n, row, col, deg = g["n"], g["row"], g["col"], g["deg"]
    grads = {}
    do = dM / g["diag"][row]
    EL = cache[f"E{layers}"]
    grads["wo"] = EL.T @ do
    grads["bo"] = np.array([do.sum()])
    dE = np.outer(do, p["wo"])
    dH = np.zeros((n, dE.shape[1]))
    for l in reversed(range(layers)):
        W = {w: r(p[f"{w}{l}"]) for w in LAYER_MATS}
        Eprev, Hprev = cache[f"E{l}"], cache[f"H{l}"]
        E, H, G = cache[f"E{l + 1}"], cache[f"H{l + 1}"], cache[f"G{l + 1}"]
        # node update
        dT = r(dH * (H > 0))
        grads[f"Wh{l}"] = Hprev.T @ dT
        grads[f"Wg{l}"] = G.T @ dT
        grads[f"bh{l}"] = dT.sum(0)
        dHp = dT @ W["Wh"].T
        dG = dT @ W["Wg"].T
        dE = dE + (dG / deg[:, None])[row]
        # edge update
        dS = r(dE * (E > 0))
        grads[f"We{l}"] = Eprev.T @ dS
        grads[f"be{l}"] = dS.sum(0)
        dEp = dS @ W["We"].T
        dP = np.zeros((n, dS.shape[1]))
        np.add.at(dP, col, dS)
        dQ = np.zeros((n, dS.shape[1]))
        np.add.at(dQ, row, dS)
        dP, dQ = r(dP), r(dQ)
        grads[f"Wu{l}"] = Hprev.T @ dP
        grads[f"Wv{l}"] = Hprev.T @ dQ
        dHp = dHp + dP @ W["Wu"].T + dQ @ W["Wv"].T
        dE, dH = dEp, dHp
    dE0 = dE * (cache["E0"] > 0)
    dH0 = dH * (cache["H0"] > 0)
    grads["Ae"] = g["xe"].T @ dE0
    grads["ae"] = dE0.sum(0)
    grads["An"] = g["xn"].T @ dH0
    grads["an"] = dH0.sum(0)
    return grads


def loss_and_grad(flat, off, cols, vals, Z, bf=True, hidden=HIDDEN, layers=LAYERS):
    p = unflatten(flat, hidden, layers)
    g = graph(off, cols, vals)
    M, cache = forward(p, g, bf, layers)
    L, dM = probe_loss(off, cols, vals, M, Z)
    grads = backward(p, g, cache, dM, bf, layers)
    return L, flatten(grads, hidden, layers), M


def init_params(seed, hidden=HIDDEN, layers=LAYERS):
    """He-normal weights, zero biases, zero decoder (M starts as Jacobi).
    The GPU library draws its own initial weights; tests pass these in."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, shape in param_shapes(hidden, layers):
        if len(shape) == 2:
            p[name] = rng.standard_normal(shape) * np.sqrt(2.0 / shape[0])
        else:
            p[name] = np.zeros(shape)
    return flatten(p, hidden, layers)
