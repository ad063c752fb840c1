"""ctypes binding of libtnlmi.so (the C ABI declared in include/tnlmi.h).

The shared library is built in-tree (``tnl_amd/libtnlmi.so``) by
``__graft_entry__.build()``; loading fails loudly when it is missing: there is
no CPU fallback for any operation of this package.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtnlmi.so")

c_i32p = ctypes.POINTER(ctypes.c_int32)
c_i64 = ctypes.c_int64
c_vp = ctypes.c_void_p

# status codes (tnlmi_status)
OK, INVALID_ARGUMENT, RUNTIME_ERROR, OVERFLOW_ERROR, BACKEND_ERROR, NOT_IMPLEMENTED, BAD_ALLOC, LOGIC_ERROR = range(8)


class TnlmiError(RuntimeError):
    """Base of the errors raised for a non-OK status."""

    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


class InvalidArgument(TnlmiError, ValueError):
    """std::invalid_argument in the reference."""


class OverflowError_(TnlmiError, OverflowError):
    """std::overflow_error in the reference."""


class NotImplementedError_(TnlmiError, NotImplementedError):
    """Exceptions::NotImplementedError in the reference."""


class LogicError(TnlmiError):
    """std::logic_error in the reference."""


class BackendError(TnlmiError):
    """Exceptions::BackendRuntimeError / BackendBadAlloc in the reference."""


_ERRORS = {
    INVALID_ARGUMENT: InvalidArgument,
    RUNTIME_ERROR: TnlmiError,
    OVERFLOW_ERROR: OverflowError_,
    BACKEND_ERROR: BackendError,
    NOT_IMPLEMENTED: NotImplementedError_,
    BAD_ALLOC: BackendError,
    LOGIC_ERROR: LogicError,
}

_lib = None


def lib():
    """The loaded library (loads it on first use, raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "(there is no CPU fallback)")
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def check(status):
    if status != OK:
        msg = lib().tnlmi_last_error().decode()
        raise _ERRORS.get(status, TnlmiError)(status, msg)


class Launch(ctypes.Structure):
    _fields_ = [("mapping", ctypes.c_int32), ("threads_per_segment", ctypes.c_int32)]


class Matrix(ctypes.Structure):
    _fields_ = [
        ("rows", ctypes.c_int64),
        ("columns", ctypes.c_int64),
        ("segments", ctypes.c_void_p),
        ("value_type", ctypes.c_int),
        ("values", ctypes.c_void_p),
        ("column_indexes", ctypes.c_void_p),
        ("matrix_type", ctypes.c_int32),
    ]


class SolverState(ctypes.Structure):
    """tnlmi_solver_state: IterativeSolver settings in, results out."""
    _fields_ = [
        ("max_iterations", ctypes.c_int64),
        ("min_iterations", ctypes.c_int64),
        ("convergence_residue", ctypes.c_double),
        ("divergence_residue", ctypes.c_double),
        ("iterations", ctypes.c_int64),
        ("residue", ctypes.c_double),
        ("converged", ctypes.c_int32),
    ]


# name -> (restype, argtypes)
_PROTOS = {
    "tnlmi_last_error": (ctypes.c_char_p, []),
    "tnlmi_version": (ctypes.c_char_p, []),
    "tnlmi_segments_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int32, c_vp, c_i64, c_vp,
                                             ctypes.POINTER(c_vp)]),
    "tnlmi_segments_destroy": (None, [c_vp]),
    "tnlmi_segments_count": (c_i64, [c_vp]),
    "tnlmi_segments_storage_size": (c_i64, [c_vp]),
    "tnlmi_segments_element_count": (c_i64, [c_vp]),
    "tnlmi_segments_format": (ctypes.c_int, [c_vp]),
    "tnlmi_segments_get_array": (ctypes.c_int, [c_vp, ctypes.c_int, ctypes.POINTER(c_vp),
                                                ctypes.POINTER(c_i64)]),
    "tnlmi_segments_ellpack_dims": (ctypes.c_int, [c_vp, c_i32p, c_i32p]),
    "tnlmi_segments_adaptive_blocks": (ctypes.c_int, [c_vp, ctypes.c_int32, ctypes.POINTER(c_vp),
                                                      ctypes.POINTER(c_i64)]),
    "tnlmi_segments_global_indexes": (ctypes.c_int, [c_vp, c_vp, c_vp]),
    "tnlmi_matrix_fill": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, ctypes.c_int, c_vp, ctypes.c_int, c_vp, c_vp]),
    "tnlmi_spmv": (ctypes.c_int, [ctypes.POINTER(Matrix), ctypes.c_int, c_vp, c_vp, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_double, c_i64, c_i64, ctypes.POINTER(Launch), c_vp]),
    "tnlmi_spmv_plan_create": (ctypes.c_int, [ctypes.POINTER(Matrix), ctypes.c_int32, c_vp, ctypes.POINTER(c_vp)]),
    "tnlmi_spmv_plan_destroy": (None, [c_vp]),
    "tnlmi_spmv_plan_execute": (ctypes.c_int, [c_vp, ctypes.c_int, c_vp, c_vp, ctypes.c_int, ctypes.c_double,
                                               ctypes.c_double, c_i64, c_i64, c_vp]),
    "tnlmi_dot": (ctypes.c_int, [c_i64, c_vp, c_vp, c_vp, c_vp]),
    "tnlmi_nrm2": (ctypes.c_int, [c_i64, c_vp, c_vp, c_vp]),
    "tnlmi_dot_noalloc": (ctypes.c_int, [c_i64, c_vp, c_vp, c_vp, c_vp]),
    "tnlmi_axpby": (ctypes.c_int, [c_i64, ctypes.c_double, c_vp, ctypes.c_double, c_vp, c_vp]),
    "tnlmi_preconditioner_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(Matrix), c_vp, ctypes.POINTER(c_vp)]),
    "tnlmi_preconditioner_destroy": (None, [c_vp]),
    "tnlmi_preconditioner_solve": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp]),
    "tnlmi_cg_solve": (ctypes.c_int, [ctypes.POINTER(Matrix), c_vp, c_vp, c_vp, c_vp, ctypes.POINTER(SolverState),
                                      c_vp]),
    "tnlmi_cg_solve_true_residual": (ctypes.c_int, [ctypes.POINTER(Matrix), c_vp, c_vp, ctypes.c_int32, c_vp, c_vp,
                                                    ctypes.POINTER(SolverState), c_vp]),
    "tnlmi_gmres_solve": (ctypes.c_int, [ctypes.POINTER(Matrix), c_vp, c_vp, ctypes.c_int, ctypes.c_int32,
                                         ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, c_vp, c_vp,
                                         ctypes.POINTER(SolverState), c_vp]),
    "tnlmi_gmres_mixed_solve": (ctypes.c_int, [ctypes.POINTER(Matrix), ctypes.POINTER(Matrix), c_vp, c_vp,
                                               ctypes.c_int, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                               ctypes.c_int32, c_vp, c_vp, ctypes.POINTER(SolverState), c_vp]),
    "tnlmi_bicgstab_solve": (ctypes.c_int, [ctypes.POINTER(Matrix), c_vp, c_vp, ctypes.c_int32, c_vp, c_vp,
                                            ctypes.POINTER(SolverState), c_vp]),
    "tnlmi_bicgstabl_solve": (ctypes.c_int, [ctypes.POINTER(Matrix), c_vp, c_vp, ctypes.c_int32, ctypes.c_int32, c_vp,
                                             c_vp, ctypes.POINTER(SolverState), c_vp]),
    "tnlmi_idrs_solve": (ctypes.c_int, [ctypes.POINTER(Matrix), c_vp, c_vp, ctypes.c_int32, ctypes.c_int32, c_vp, c_vp,
                                        ctypes.POINTER(SolverState), c_vp]),
    "tnlmi_tfqmr_solve": (ctypes.c_int, [ctypes.POINTER(Matrix), c_vp, c_vp, c_vp, c_vp, ctypes.POINTER(SolverState),
                                         c_vp]),
    "tnlmi_jacobi_solve": (ctypes.c_int, [ctypes.POINTER(Matrix), c_vp, ctypes.c_double, c_i64, c_vp, c_vp,
                                          ctypes.POINTER(SolverState), c_vp]),
    "tnlmi_true_residual": (ctypes.c_int, [ctypes.POINTER(Matrix), c_vp, c_vp, ctypes.POINTER(ctypes.c_double), c_vp]),
    "tnlmi_ir_cg_solve": (ctypes.c_int, [ctypes.POINTER(Matrix), ctypes.POINTER(Matrix), c_vp, c_vp, ctypes.c_double,
                                         c_i64, c_vp, c_vp, ctypes.POINTER(SolverState), ctypes.POINTER(c_i64),
                                         c_vp]),
    "tnlmi_mixed_cg_solve": (ctypes.c_int, [ctypes.POINTER(Matrix), ctypes.POINTER(Matrix), c_vp, ctypes.c_double,
                                            c_vp, c_vp, ctypes.POINTER(SolverState), ctypes.POINTER(c_i64), c_vp]),
    "tnlmi_spai_parameter_count": (c_i64, []),
    "tnlmi_spai_create": (ctypes.c_int, [c_vp, ctypes.POINTER(c_vp)]),
    "tnlmi_spai_destroy": (None, [c_vp]),
    "tnlmi_spai_get_parameters": (ctypes.c_int, [c_vp, c_vp, c_vp]),
    "tnlmi_spai_apply": (ctypes.c_int, [c_vp, ctypes.POINTER(Matrix), c_vp, c_vp]),
    "tnlmi_spai_loss_grad": (ctypes.c_int, [c_vp, ctypes.POINTER(Matrix), c_vp, ctypes.c_int32,
                                            ctypes.POINTER(ctypes.c_double), c_vp, c_vp]),
    "tnlmi_spai_train_step": (ctypes.c_int, [c_vp, ctypes.POINTER(Matrix), ctypes.c_int32, ctypes.c_uint64,
                                             ctypes.c_float, ctypes.POINTER(ctypes.c_double), c_vp]),
    "tnlmi_bf16_gemm64": (ctypes.c_int, [c_i64, c_vp, c_vp, c_vp, c_vp]),
    "tnlmi_batched_cg_create": (ctypes.c_int, [ctypes.POINTER(Matrix), ctypes.POINTER(Matrix), c_vp, c_i64,
                                               ctypes.c_int32, ctypes.c_int32, c_vp, ctypes.POINTER(c_vp)]),
    "tnlmi_batched_cg_destroy": (None, [c_vp]),
    "tnlmi_batched_cg_solve": (ctypes.c_int, [c_vp, ctypes.c_int32, c_vp, c_vp, ctypes.c_double, c_i64,
                                              ctypes.c_double, c_vp, c_vp, ctypes.c_int32, c_vp]),
    "tnlmi_matrix_row_lengths": (ctypes.c_int, [ctypes.POINTER(Matrix), c_vp, c_vp]),
    "tnlmi_matrix_modified": (ctypes.c_int, [ctypes.POINTER(Matrix)]),
    "tnlmi_matrix_copy": (ctypes.c_int, [ctypes.POINTER(Matrix), ctypes.POINTER(Matrix), c_vp, c_vp]),
    "tnlmi_matrix_add_element": (ctypes.c_int, [ctypes.POINTER(Matrix), ctypes.c_int64, ctypes.c_int64,
                                                ctypes.c_double, ctypes.c_double, c_vp]),
    "tnlmi_matrix_add_elements": (ctypes.c_int, [ctypes.POINTER(Matrix), c_i64, c_vp, c_vp, c_vp, c_vp,
                                                 ctypes.POINTER(c_i64), c_vp]),
    "tnlmi_matrix_row_counts": (ctypes.c_int, [ctypes.POINTER(Matrix), ctypes.c_int32, c_vp, c_vp]),
    "tnlmi_matrix_get_element": (ctypes.c_int, [ctypes.POINTER(Matrix), ctypes.c_int64, ctypes.c_int64,
                                                ctypes.POINTER(ctypes.c_double), c_vp]),
    "tnlmi_mtx_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(c_vp)]),
    "tnlmi_mtx_parse": (ctypes.c_int, [ctypes.c_char_p, c_i64, ctypes.POINTER(c_vp)]),
    "tnlmi_host_csr_get": (ctypes.c_int, [c_vp, ctypes.POINTER(c_i64), ctypes.POINTER(c_i64), ctypes.POINTER(c_i64),
                                          ctypes.POINTER(c_vp), ctypes.POINTER(c_vp), ctypes.POINTER(c_vp)]),
    "tnlmi_host_csr_destroy": (None, [c_vp]),
    "tnlmi_matrix_save": (ctypes.c_int, [ctypes.POINTER(Matrix), ctypes.c_char_p, c_vp]),
    "tnlmi_spmv_transposed": (ctypes.c_int, [ctypes.POINTER(Matrix), ctypes.c_int, c_vp, c_vp, ctypes.c_double,
                                             ctypes.c_double, c_i64, c_i64, c_vp]),
    "tnlmi_matrix_serialization_type": (ctypes.c_int, [ctypes.POINTER(Matrix), ctypes.c_char_p, c_i64,
                                                       ctypes.POINTER(c_i64)]),
    "tnlmi_matrix_file_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(c_vp)]),
    "tnlmi_matrix_file_info": (ctypes.c_int, [c_vp, ctypes.POINTER(c_i64), ctypes.POINTER(c_i64),
                                              ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                              ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int),
                                              ctypes.POINTER(c_i64), ctypes.POINTER(ctypes.c_int)]),
    "tnlmi_matrix_file_type": (ctypes.c_char_p, [c_vp]),
    "tnlmi_matrix_file_segments": (ctypes.c_int, [c_vp, c_vp, ctypes.POINTER(c_vp)]),
    "tnlmi_matrix_file_upload": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp]),
    "tnlmi_matrix_file_destroy": (None, [c_vp]),
    "tnlmi_gen_batched_diffusion": (None, [c_i64, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64, c_vp, c_vp, c_vp,
                                           c_vp, c_vp, c_vp]),
    "tnlmi_gen_convdiff": (ctypes.c_double, [c_i64, ctypes.c_uint64, c_vp, c_vp, c_vp]),
    "tnlmi_gen_laplace2d": (c_i64, [c_i64, c_i64, c_vp, c_vp, c_vp]),
    "tnlmi_gen_poisson3d": (c_i64, [c_i64, c_i64, c_i64, c_vp, c_vp, c_vp]),
    "tnlmi_gen_zipf_sizes": (None, [c_i64, ctypes.c_double, ctypes.c_int32, ctypes.c_double, ctypes.c_int32,
                                    ctypes.c_int32, ctypes.c_uint64, c_vp]),
    "tnlmi_gen_random_columns": (None, [c_i64, c_i64, c_vp, ctypes.c_uint64, c_vp]),
    "tnlmi_gen_uniform": (None, [c_i64, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, ctypes.c_uint64, c_vp]),
    "tnlmi_gen_normal": (None, [c_i64, ctypes.c_uint64, ctypes.c_uint64, c_vp]),
}

# every symbol include/tnlmi.h declares (checked by the CPU tests)
DECLARED = list(_PROTOS)


def _declare(L):
    for name, (res, args) in _PROTOS.items():
        f = getattr(L, name, None)
        if f is None:
            continue
        f.restype = res
        f.argtypes = args
