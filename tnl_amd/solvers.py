"""Solvers::Linear mirror: CG, GMRES, BICGStab, TFQMR, Jacobi and the Diagonal / ILU0
preconditioners of the reference (Solvers/Linear/*.h), running on the GPU
through the C ABI.  Method names and defaults follow the reference
(IterativeSolver.h:226-235, GMRES.hpp:18-69, Jacobi.h:121-123)."""
import ctypes
import sys

from . import _lib
from . import _stream, InvalidArgument

NONE, DIAGONAL, ILU0_KIND = 0, 1, 2
GMRES_VARIANTS = {"CGS": 0, "CGSR": 1, "MGS": 2, "MGSR": 3, "CWY": 4}


def _f64(t):
    import torch
    if t.dtype != torch.float64 or not t.is_cuda:
        raise InvalidArgument(_lib.INVALID_ARGUMENT, "solvers work on FP64 device vectors")
    return ctypes.c_void_p(t.data_ptr())


class Preconditioner:
    kind = NONE

    def __init__(self, matrix=None):
        self.handle = None
        if matrix is not None:
            self.update(matrix)

    def update(self, matrix):
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().tnlmi_preconditioner_create(self.kind, ctypes.byref(matrix._m), _stream(),
                                                          ctypes.byref(h)))
        self.handle = h
        self.n = matrix.rows

    def solve(self, b, x):
        _lib.check(_lib.lib().tnlmi_preconditioner_solve(self.handle, _f64(b), _f64(x), _stream()))
        return x

    def __del__(self):
        h = getattr(self, "handle", None)
        try:
            if h is not None and h.value and _lib._lib is not None:
                _lib.lib().tnlmi_preconditioner_destroy(h)
        except (AttributeError, TypeError):
            pass
        self.handle = None


class Diagonal(Preconditioner):
    """Preconditioners::Diagonal (Jacobi)."""
    kind = DIAGONAL


class ILU0(Preconditioner):
    """Preconditioners::ILU0."""
    kind = ILU0_KIND


class Explicit(Preconditioner):
    """x = M b with an explicit sparse M (e.g. a learned sparse approximate
    inverse, tnl_amd.spai); M must stay alive while the preconditioner is used."""
    kind = 3

    def update(self, matrix):
        self.M = matrix
        return super().update(matrix)


class IterativeSolver:
    def __init__(self):
        self.max_iterations = 1_000_000_000
        self.min_iterations = 0
        self.convergence_residue = 1e-6
        self.divergence_residue = sys.float_info.max
        self.iterations = 0
        self.residue = 0.0
        self.matrix = None
        self.plan = None
        self.preconditioner = None

    def setMatrix(self, matrix, plan=None):
        self.matrix = matrix
        self.plan = plan

    def setPreconditioner(self, p):
        self.preconditioner = p

    def setMaxIterations(self, v):
        self.max_iterations = int(v)

    def setMinIterations(self, v):
        self.min_iterations = int(v)

    def setConvergenceResidue(self, v):
        self.convergence_residue = float(v)

    def setDivergenceResidue(self, v):
        self.divergence_residue = float(v)

    def getIterations(self):
        return self.iterations

    def getResidue(self):
        return self.residue

    def _state(self):
        return _lib.SolverState(self.max_iterations, self.min_iterations, self.convergence_residue,
                                self.divergence_residue, 0, 0.0, 0)

    def _finish(self, st):
        self.iterations = st.iterations
        self.residue = st.residue
        return bool(st.converged)

    def _handles(self):
        if self.matrix is None:
            raise _lib.LogicError(_lib.LOGIC_ERROR, "No matrix was set. Call setMatrix() before solve().")
        plan = self.plan.handle if self.plan is not None else None
        pre = self.preconditioner.handle if self.preconditioner is not None else None
        return ctypes.byref(self.matrix._m), plan, pre


class CG(IterativeSolver):
    def solve(self, b, x):
        m, plan, pre = self._handles()
        st = self._state()
        _lib.check(_lib.lib().tnlmi_cg_solve(m, plan, pre, _f64(b), _f64(x), ctypes.byref(st), _stream()))
        return self._finish(st)


class TrueResidualCG(IterativeSolver):
    """CG's iteration (no or the Diagonal preconditioner) stopped on the true
    residual ||b - A x|| / ||b|| -- the stopping rule of the mixed-precision
    solvers -- instead of the reference's preconditioned residue
    (tnlmi_cg_solve_true_residual).  column_codes: read a column-major
    Ellpack / SlicedEllpack matrix through 8-bit column codes, as the mixed
    solver reads its inner matrix."""

    def __init__(self, column_codes=False):
        super().__init__()
        self.column_codes = bool(column_codes)

    def solve(self, b, x):
        m, plan, pre = self._handles()
        st = self._state()
        _lib.check(_lib.lib().tnlmi_cg_solve_true_residual(m, plan, pre, int(self.column_codes), _f64(b), _f64(x),
                                                           ctypes.byref(st), _stream()))
        return self._finish(st)


class GMRES(IterativeSolver):
    def __init__(self):
        super().__init__()
        self.variant = "MGSR"  # gmres-variant default of configSetup (GMRES.hpp:23)
        self.restarting_min = 10
        self.restarting_max = 10
        self.restarting_step_min = 3
        self.restarting_step_max = 3

    def setup(self, variant=None, restarting_min=None, restarting_max=None, step_min=None, step_max=None):
        if variant is not None:
            if variant not in GMRES_VARIANTS:
                return False
            self.variant = variant
        if restarting_min is not None:
            self.restarting_min = restarting_min
        if restarting_max is not None:
            self.restarting_max = restarting_max
        if step_min is not None:
            self.restarting_step_min = step_min
        if step_max is not None:
            self.restarting_step_max = step_max
        return True

    def solve(self, b, x):
        m, plan, pre = self._handles()
        st = self._state()
        _lib.check(_lib.lib().tnlmi_gmres_solve(m, plan, pre, GMRES_VARIANTS[self.variant], self.restarting_min,
                                                self.restarting_max, self.restarting_step_min,
                                                self.restarting_step_max, _f64(b), _f64(x), ctypes.byref(st),
                                                _stream()))
        return self._finish(st)


class MixedGMRES(GMRES):
    """Mixed-precision GMRES (tnlmi_gmres_mixed_solve): GMRES's iteration and
    options (variant, adaptive restarting, preconditioner, stopping rule) with
    the Arnoldi products on the inner matrix -- the same matrix with FP32 or
    BF16 values -- and every restart residual M^-1 (b - A x) in FP64 with the
    matrix of setMatrix: iterative refinement with one GMRES cycle per
    correction.  Basis, Hessenberg least squares and x stay FP64.  No
    reference counterpart."""

    def __init__(self):
        super().__init__()
        self.matrix_lo = None

    def setInnerMatrix(self, matrix_lo):
        self.matrix_lo = matrix_lo

    def solve(self, b, x):
        if self.matrix_lo is None:
            raise _lib.LogicError(_lib.LOGIC_ERROR, "mixed GMRES: no inner matrix was set")
        m, plan, pre = self._handles()
        st = self._state()
        _lib.check(_lib.lib().tnlmi_gmres_mixed_solve(m, ctypes.byref(self.matrix_lo._m), plan, pre,
                                                      GMRES_VARIANTS[self.variant], self.restarting_min,
                                                      self.restarting_max, self.restarting_step_min,
                                                      self.restarting_step_max, _f64(b), _f64(x), ctypes.byref(st),
                                                      _stream()))
        return self._finish(st)


class BICGStab(IterativeSolver):
    """Solvers::Linear::BICGStab (BICGStab.hpp:33-127); setup(exact_residue=...)
    is the bicgstab-exact-residue option (default false: the residue is
    ||r|| of the BiCGStab recurrence)."""

    def __init__(self):
        super().__init__()
        self.exact_residue = False

    def setup(self, exact_residue=None):
        if exact_residue is not None:
            self.exact_residue = bool(exact_residue)
        return True

    def solve(self, b, x):
        m, plan, pre = self._handles()
        st = self._state()
        _lib.check(_lib.lib().tnlmi_bicgstab_solve(m, plan, pre, int(self.exact_residue), _f64(b), _f64(x),
                                                   ctypes.byref(st), _stream()))
        return self._finish(st)


class TFQMR(IterativeSolver):
    """Solvers::Linear::TFQMR (TFQMR.hpp:12-121): the residue is the quasi-residual
    tau sqrt(it + 1) / ||M^-1 b||."""

    def solve(self, b, x):
        m, plan, pre = self._handles()
        st = self._state()
        _lib.check(_lib.lib().tnlmi_tfqmr_solve(m, plan, pre, _f64(b), _f64(x), ctypes.byref(st), _stream()))
        return self._finish(st)


class BICGStabL(IterativeSolver):
    """Solvers::Linear::BICGStabL (BICGStabL.hpp:37-257): BiCGStab(ell);
    setup(ell=..., exact_residue=...) are "bicgstab-ell" (default 1) and
    "bicgstab-exact-residue" (default false).  One iteration per Bi-CG step."""

    def __init__(self):
        super().__init__()
        self.ell = 1
        self.exact_residue = False

    def setup(self, ell=None, exact_residue=None):
        if ell is not None:
            self.ell = int(ell)
        if exact_residue is not None:
            self.exact_residue = bool(exact_residue)
        return True

    def solve(self, b, x):
        m, plan, pre = self._handles()
        st = self._state()
        _lib.check(_lib.lib().tnlmi_bicgstabl_solve(m, plan, pre, self.ell, int(self.exact_residue), _f64(b),
                                                    _f64(x), ctypes.byref(st), _stream()))
        return self._finish(st)


class IDRs(IterativeSolver):
    """Solvers::Linear::IDRs (IDRs.hpp:47-386): IDR(s) with the reference's
    shadow space; setShadowSpaceDimension (idr-s, default 4),
    setResidualSmoothing (idr-residual-smoothing, default false).  A breakdown
    returns False."""

    def __init__(self):
        super().__init__()
        self.s = 4
        self.smoothing = False

    def setShadowSpaceDimension(self, s):
        self.s = int(s)

    def setResidualSmoothing(self, smoothing):
        self.smoothing = bool(smoothing)

    def solve(self, b, x):
        m, plan, pre = self._handles()
        st = self._state()
        _lib.check(_lib.lib().tnlmi_idrs_solve(m, plan, pre, self.s, int(self.smoothing), _f64(b), _f64(x),
                                               ctypes.byref(st), _stream()))
        return self._finish(st)


class Jacobi(IterativeSolver):
    def __init__(self):
        super().__init__()
        self.omega = 1.0
        self.residue_period = 4

    def setOmega(self, w):
        self.omega = float(w)

    def setResiduePeriod(self, p):
        self.residue_period = int(p)

    def solve(self, b, x):
        m, plan, _ = self._handles()
        st = self._state()
        _lib.check(_lib.lib().tnlmi_jacobi_solve(m, plan, self.omega, self.residue_period, _f64(b), _f64(x),
                                                 ctypes.byref(st), _stream()))
        return self._finish(st)


class MixedCG(IterativeSolver):
    """FP64 iterative refinement around an FP32-arithmetic Jacobi-CG on a copy
    of the matrix with FP32 or BF16 values (tnlmi_ir_cg_solve).  The stopping
    rule is the FP64 true residual; getIterations() counts refinement steps and
    getInnerIterations() the inner CG iterations."""

    def __init__(self):
        super().__init__()
        self.matrix_lo = None
        self.plan_lo = None
        self.inner_tolerance = 1e-4
        self.inner_max_iterations = 10_000
        self.inner_iterations = 0

    def setInnerMatrix(self, matrix_lo, plan=None):
        self.matrix_lo = matrix_lo
        self.plan_lo = plan

    def setInnerTolerance(self, tol):
        self.inner_tolerance = float(tol)

    def getInnerIterations(self):
        return self.inner_iterations

    def solve(self, b, x):
        m, plan, _ = self._handles()
        if self.matrix_lo is None:
            raise _lib.LogicError(_lib.LOGIC_ERROR, "No inner matrix was set. Call setInnerMatrix() first.")
        st = self._state()
        inner = ctypes.c_int64()
        plan_lo = self.plan_lo.handle if self.plan_lo is not None else None
        _lib.check(_lib.lib().tnlmi_ir_cg_solve(m, ctypes.byref(self.matrix_lo._m), plan, plan_lo,
                                                self.inner_tolerance, self.inner_max_iterations, _f64(b), _f64(x),
                                                ctypes.byref(st), ctypes.byref(inner), _stream()))
        self.inner_iterations = inner.value
        return self._finish(st)


class ReliableMixedCG(IterativeSolver):
    """Jacobi-CG in FP32 on the FP32- or BF16-valued copy of the matrix with
    reliable updates: x accumulated in FP64, the residual replaced by the FP64
    b - A x whenever the FP32 residual has fallen by `delta` since the last
    replacement (tnlmi_mixed_cg_solve).  Stops on the FP64 true residual;
    getIterations() counts FP32 iterations, getUpdates() the FP64 residuals."""

    def __init__(self):
        super().__init__()
        self.matrix_lo = None
        self.delta = 0.1
        self.updates = 0

    def setInnerMatrix(self, matrix_lo):
        self.matrix_lo = matrix_lo

    def setDelta(self, delta):
        self.delta = float(delta)

    def getUpdates(self):
        return self.updates

    def solve(self, b, x):
        m, plan, _ = self._handles()
        if self.matrix_lo is None:
            raise _lib.LogicError(_lib.LOGIC_ERROR, "No inner matrix was set. Call setInnerMatrix() first.")
        st = self._state()
        upd = ctypes.c_int64()
        _lib.check(_lib.lib().tnlmi_mixed_cg_solve(m, ctypes.byref(self.matrix_lo._m), plan, self.delta, _f64(b),
                                                   _f64(x), ctypes.byref(st), ctypes.byref(upd), _stream()))
        self.updates = upd.value
        return self._finish(st)


def true_residual(matrix, x, b):
    """LinearResidueGetter::getResidue: ||A x - b|| / ||b||."""
    r = ctypes.c_double()
    _lib.check(_lib.lib().tnlmi_true_residual(ctypes.byref(matrix._m), _f64(x), _f64(b), ctypes.byref(r), _stream()))
    return r.value


def dot(a, b):
    """(a, b) of FP64 device vectors (host scalar, like the reference's operator)."""
    import torch
    out = torch.empty(1, dtype=torch.float64, device="cuda")
    _lib.check(_lib.lib().tnlmi_dot(a.numel(), _f64(a), _f64(b), ctypes.c_void_p(out.data_ptr()), _stream()))
    return float(out.item())


def l2Norm(a):
    import torch
    out = torch.empty(1, dtype=torch.float64, device="cuda")
    _lib.check(_lib.lib().tnlmi_nrm2(a.numel(), _f64(a), ctypes.c_void_p(out.data_ptr()), _stream()))
    return float(out.item())
