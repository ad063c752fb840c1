"""Learned sparse approximate inverse (tnlmi_spai_*): a message-passing
network over the graph of A that predicts M = D^-1 (I + O) on A's pattern, a
left preconditioner for GMRES, trained on the GPU by the library's own
forward and backward kernels with the probe loss ||(I - M A) z||^2.

Also: the workload generator of its benchmark (2D convection-diffusion,
tnl_amd.convdiff), the preconditioner wrapper, and GMRES with fallback to
Jacobi / ILU(0) when the learned preconditioner stalls."""
import ctypes
import time

import numpy as np

from . import _lib
from . import SparseMatrix, _stream, _ptr, convdiff
from . import solvers

HIDDEN = 64
LAYERS = 3
_MATS = ("We", "Wu", "Wv", "Wh", "Wg")


def layout(hidden=HIDDEN, layers=LAYERS):
    """(name, shape) of every parameter, in the order of the flat vector."""
    out = [("Ae", (4, hidden)), ("ae", (hidden,)), ("An", (4, hidden)), ("an", (hidden,))]
    for l in range(layers):
        out += [(f"{w}{l}", (hidden, hidden)) for w in _MATS]
        out += [(f"be{l}", (hidden,)), (f"bh{l}", (hidden,))]
    return out + [("wo", (hidden,)), ("bo", (1,))]


def init_parameters(seed=0):
    """He-normal weight matrices, zero biases and a zero decoder, so that the
    untrained network outputs the Jacobi inverse D^-1."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in layout():
        if len(shape) == 2:
            parts.append((rng.standard_normal(shape) * np.sqrt(2.0 / shape[0])).ravel())
        else:
            parts.append(np.zeros(int(np.prod(shape))))
    flat = np.concatenate(parts).astype(np.float32)
    assert len(flat) == _lib.lib().tnlmi_spai_parameter_count()
    return flat


class Network:
    """The network on the GPU (parameters, gradients, Adam state)."""

    def __init__(self, parameters=None, seed=0):
        p = np.ascontiguousarray(parameters if parameters is not None else init_parameters(seed), np.float32)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().tnlmi_spai_create(_ptr(p), ctypes.byref(h)))
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        try:
            if h is not None and h.value and _lib._lib is not None:
                _lib.lib().tnlmi_spai_destroy(h)
        except (AttributeError, TypeError):
            pass
        self.handle = None

    def parameters(self):
        p = np.zeros(_lib.lib().tnlmi_spai_parameter_count(), np.float32)
        _lib.check(_lib.lib().tnlmi_spai_get_parameters(self.handle, _ptr(p), None))
        return p

    def gradients(self):
        g = np.zeros(_lib.lib().tnlmi_spai_parameter_count(), np.float32)
        _lib.check(_lib.lib().tnlmi_spai_get_parameters(self.handle, None, _ptr(g)))
        return g

    def inverse(self, A):
        """M on A's pattern: a SparseMatrix sharing A's segments and column indexes."""
        import torch
        M = torch.empty(A.values.numel(), dtype=torch.float64, device="cuda")
        _lib.check(_lib.lib().tnlmi_spai_apply(self.handle, ctypes.byref(A._m), ctypes.c_void_p(M.data_ptr()),
                                               _stream()))
        return SparseMatrix(A.rows, A.columns, A.segments, M, A.column_indexes)

    def loss_grad(self, A, Z):
        """(loss, M values) for the probes Z (K x n device FP64); gradients kept."""
        import torch
        Z = Z.contiguous()
        M = torch.empty(A.values.numel(), dtype=torch.float64, device="cuda")
        L = ctypes.c_double()
        _lib.check(_lib.lib().tnlmi_spai_loss_grad(self.handle, ctypes.byref(A._m), ctypes.c_void_p(Z.data_ptr()),
                                                   int(Z.shape[0]), ctypes.byref(L), ctypes.c_void_p(M.data_ptr()),
                                                   _stream()))
        return L.value, M

    def train_step(self, A, probes=4, seed=0, lr=1e-3):
        L = ctypes.c_double()
        _lib.check(_lib.lib().tnlmi_spai_train_step(self.handle, ctypes.byref(A._m), int(probes), int(seed),
                                                    float(lr), ctypes.byref(L), _stream()))
        return L.value


def bf16_gemm64(X, W):
    """Y = X W for BF16 X [rows x 64] and W [64 x 64] (FP32 result)."""
    import torch
    Y = torch.empty(X.shape[0], 64, dtype=torch.float32, device="cuda")
    _lib.check(_lib.lib().tnlmi_bf16_gemm64(X.shape[0], ctypes.c_void_p(X.contiguous().data_ptr()),
                                            ctypes.c_void_p(W.contiguous().data_ptr()),
                                            ctypes.c_void_p(Y.data_ptr()), _stream()))
    return Y


def convdiff_matrix(m, seed):
    off, cols, vals, pe = convdiff(m, seed)
    return SparseMatrix.from_csr(len(off) - 1, len(off) - 1, off, cols, vals), pe


def train(net, sizes=(256,), instances=2000, steps_per_instance=1, probes=4, lr=1e-3, seed=1000, log=None):
    """Trains on `instances` seeded convection-diffusion matrices (sides drawn
    from `sizes`), each generated on the host and used for
    steps_per_instance Adam steps; returns the loss history."""
    hist = []
    t0 = time.perf_counter()
    for k in range(instances):
        m = sizes[k % len(sizes)]
        A, _ = convdiff_matrix(m, seed + k)
        for j in range(steps_per_instance):
            hist.append(net.train_step(A, probes, seed=(seed + k) * 131 + j, lr=lr))
        if log and (k % max(1, instances // 20) == 0 or k == instances - 1):
            log(f"spai train {k + 1}/{instances}: loss {np.mean(hist[-steps_per_instance * 20:]):.4f} "
                f"({time.perf_counter() - t0:.1f}s)")
    return hist


def gmres_solve(A, b, x, preconditioner, tolerance=1e-10, restart=30, max_iterations=100_000, variant="MGSR",
                stall_cycles=20, true_residual_check=True, log=None):
    """Left-preconditioned GMRES(restart) until the FP64 true residual
    ||b - A x|| / ||b|| <= tolerance (LinearResidueGetter.hpp:11-23).  GMRES
    stops on its preconditioned residue, so the solve is resumed from x with
    a proportionally tighter residue target while the true residual is above
    the tolerance.  Stops early (returns False) when the true residual falls
    by less than 1 % over stall_cycles restart cycles (or over one
    resumed solve).  log, when given, is called with a progress line after
    every stall_cycles restart cycles.
    Returns (converged, iterations, true residual)."""
    s = solvers.GMRES()
    s.setMatrix(A)
    s.setPreconditioner(preconditioner)
    s.setup(variant, restart, restart)
    target = tolerance
    its = 0
    chunk = restart * stall_cycles
    prev = solvers.true_residual(A, x, b)
    while its < max_iterations:
        s.setConvergenceResidue(target)
        s.setMaxIterations(min(chunk, max_iterations - its))
        ok = s.solve(b, x)
        its += s.getIterations()
        r = solvers.true_residual(A, x, b)
        if log is not None:
            log(f"gmres: {its} iterations, true residual {r:.3e}")
        if r <= tolerance:
            return True, its, r
        if not true_residual_check or not (r < 0.99 * prev):  # stalled (or NaN)
            return False, its, r
        if ok:
            target = target * max(0.05, min(0.5, 0.5 * tolerance / r))
        prev = r
    return False, its, solvers.true_residual(A, x, b)


def gmres_with_fallback(A, b, x, preconditioner, fallback, tolerance=1e-10, restart=30, max_iterations=100_000,
                        stall_cycles=20, variant="MGSR"):
    """GMRES with the (learned) preconditioner; when it stalls -- the true
    residual falls by less than 1 % over stall_cycles restart cycles -- the
    solve continues from the current x with the fallback preconditioner
    (Jacobi or ILU(0)).  Returns (converged, iterations, true residual,
    used_fallback)."""
    ok, its, r = gmres_solve(A, b, x, preconditioner, tolerance, restart, max_iterations, variant, stall_cycles)
    if ok:
        return True, its, r, False
    ok2, its2, r2 = gmres_solve(A, b, x, fallback, tolerance, restart, max(1, max_iterations - its), variant,
                                stall_cycles)
    return ok2, its + its2, r2, True
