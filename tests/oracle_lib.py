"""ctypes access to the CPU oracle (oracle/liboracle.so) and to the reference
build (oracle/_ref/libtnlref.so), plus the KAT fixtures.  Test infrastructure."""
import ctypes
import json
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libtnlref.so")
KATS = os.path.join(ROOT, "tests", "golden", "spmv_kats.json")

I64 = ctypes.c_int64


def P(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class SolverState(ctypes.Structure):
    _fields_ = [("max_iterations", ctypes.c_int64), ("min_iterations", ctypes.c_int64),
                ("convergence_residue", ctypes.c_double), ("divergence_residue", ctypes.c_double),
                ("iterations", ctypes.c_int64), ("residue", ctypes.c_double), ("converged", ctypes.c_int)]


def state(tol=1e-6, maxit=10**9, minit=0, div=np.finfo(np.float64).max):
    return SolverState(maxit, minit, tol, div, 0, 0.0, 0)


class Oracle:
    def __init__(self):
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built: run __graft_entry__.build()")
        self.L = ctypes.CDLL(ORACLE_SO)
        for name in ["orc_chunked_ellpack", "orc_adaptive_csr_blocks", "orc_add_elements"]:
            getattr(self.L, name).restype = ctypes.c_int64
        self.L.orc_true_residual.restype = ctypes.c_double

    def csr_offsets(self, sizes):
        sizes = i32(sizes)
        off = np.zeros(len(sizes) + 1, np.int32)
        self.L.orc_csr_offsets(P(sizes), I64(len(sizes)), P(off))
        return off

    def ellpack_dims(self, sizes, col_major=True, alignment=32):
        sizes = i32(sizes)
        w, a = ctypes.c_int32(), ctypes.c_int32()
        rc = self.L.orc_ellpack_dims(P(sizes), I64(len(sizes)), int(col_major), alignment, ctypes.byref(w),
                                     ctypes.byref(a))
        return rc, w.value, a.value

    def sliced(self, sizes, S):
        sizes = i32(sizes)
        ns = (len(sizes) + S - 1) // S
        off = np.zeros(ns + 1, np.int32)
        ss = np.zeros(ns, np.int32)
        self.L.orc_sliced_ellpack(P(sizes), I64(len(sizes)), S, P(off), P(ss))
        return off, ss

    def biellpack(self, sizes, warp=32, brute=False):
        sizes = i32(sizes)
        n = len(sizes)
        strips = (n + warp - 1) // warp
        logw = int(np.log2(warp))
        perm = np.zeros(n, np.int32)
        gp = np.zeros(strips * (logw + 1) + 1, np.int32)
        f = self.L.orc_biellpack_bruteforce if brute else self.L.orc_biellpack
        f(P(sizes), I64(n), warp, P(perm), P(gp))
        return perm, gp

    def chunked(self, sizes, C=256, desired=16):
        sizes = i32(sizes)
        n = len(sizes)
        cap = max(n, 1) * C
        sl = np.zeros(4 * max(n, 1), np.int32)
        s2c, s2s = np.zeros(n, np.int32), np.zeros(n, np.int32)
        sp = np.zeros(n + 1, np.int32)
        c2s = np.full(cap, -1, np.int32)
        st = np.zeros(1, np.int32)
        ns = self.L.orc_chunked_ellpack(P(sizes), I64(n), C, desired, P(sl), P(s2c), P(s2s), P(sp), P(c2s),
                                        I64(cap), P(st))
        return dict(nslices=ns, slices=sl[:4 * ns], seg_to_chunk=s2c, seg_to_slice=s2s, seg_ptrs=sp,
                    chunks_to_seg=c2s[:ns * C], storage=int(st[0]))

    def adaptive_blocks(self, offsets, value_size, host_warp=32, device_warp=64):
        offsets = i32(offsets)
        n = len(offsets) - 1
        cap = 8 * (n + 16)
        b = np.zeros(6 * cap, np.int32)
        c = self.L.orc_adaptive_csr_blocks(P(offsets), I64(n), value_size, host_warp, device_warp, P(b), I64(cap))
        return b[:6 * c].reshape(-1, 6)

    def sorted_segments(self, sizes, sigma=-1):
        sizes = i32(sizes)
        n = len(sizes)
        ss, inv, perm = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n, np.int32)
        self.L.orc_sorted_segments(P(sizes), I64(n), I64(sigma), P(ss), P(inv), P(perm))
        return ss, inv, perm

    def gidx(self, fmt, sizes, col_major=True, **kw):
        """CSR-ordered global indexes of every (row, local) slot for format `fmt` (names as in tnl_amd)."""
        sizes = i32(sizes)
        n = len(sizes)
        out = np.zeros(int(sizes.sum()), np.int32)
        if fmt == "Ellpack":
            _, w, a = self.ellpack_dims(sizes, col_major)
            self.L.orc_gidx_ellpack(P(sizes), I64(n), w, a, int(col_major), P(out))
        elif fmt == "SlicedEllpack":
            S = kw["slice_size"]
            off, ss = self.sliced(sizes, S)
            self.L.orc_gidx_sliced(P(sizes), I64(n), S, P(off), P(ss), int(col_major), P(out))
        elif fmt == "BiEllpack":
            perm, gp = self.biellpack(sizes)
            self.L.orc_gidx_biellpack(P(sizes), I64(n), 32, P(perm), P(gp), int(col_major), P(out))
        elif fmt == "ChunkedEllpack":
            c = self.chunked(sizes)
            sl = np.zeros(max(4 * c["nslices"], 4), np.int32)
            sl[:len(c["slices"])] = c["slices"]
            self.L.orc_gidx_chunked(P(sizes), I64(n), 256, P(sl), P(c["seg_to_chunk"]), P(c["seg_to_slice"]),
                                    int(col_major), P(out))
        elif fmt in ("CSR", "AdaptiveCSR"):
            out = np.concatenate([np.arange(o, o + s, dtype=np.int32) for o, s in
                                  zip(self.csr_offsets(sizes)[:-1], sizes)]) if n else out
        elif fmt == "SortedCSR":
            ss, inv, perm = self.sorted_segments(sizes)
            so = self.csr_offsets(ss)
            out = np.concatenate([np.arange(so[perm[r]], so[perm[r]] + sizes[r], dtype=np.int32)
                                  for r in range(n)]) if n else out
        else:
            raise ValueError(fmt)
        return out

    def spmv(self, off, cols, vals, x, y=None, alpha=1.0, beta=0.0, begin=0, end=0, dtype=np.float64):
        off, cols = i32(off), i32(cols)
        n = len(off) - 1
        vals = np.ascontiguousarray(vals, dtype=dtype)
        x = np.ascontiguousarray(x, dtype=dtype)
        y = np.zeros(n, dtype) if y is None else np.array(y, dtype=dtype)
        if dtype == np.float64:
            self.L.orc_spmv_f64(I64(n), P(off), P(cols), P(vals), P(x), P(y), ctypes.c_double(alpha),
                                ctypes.c_double(beta), I64(begin), I64(end))
        else:
            self.L.orc_spmv_f32(I64(n), P(off), P(cols), P(vals), P(x), P(y), ctypes.c_float(alpha),
                                ctypes.c_float(beta), I64(begin), I64(end))
        return y

    def spmv_symmetric(self, off, cols, vals, x, y=None, alpha=1.0, beta=0.0, begin=0, end=0):
        off, cols, vals, x = i32(off), i32(cols), f64(vals), f64(x)
        n = len(off) - 1
        y = np.zeros(n) if y is None else f64(y).copy()
        self.L.orc_spmv_symmetric(I64(n), P(off), P(cols), P(vals), P(x), P(y), ctypes.c_double(alpha),
                                  ctypes.c_double(beta), I64(begin), I64(end))
        return y

    def spmv_transposed(self, off, cols, vals, ncols, x, y=None, alpha=1.0, beta=0.0, begin=0, end=0):
        off, cols, vals, x = i32(off), i32(cols), f64(vals), f64(x)
        n = len(off) - 1
        y = np.zeros(ncols) if y is None else f64(y).copy()
        self.L.orc_spmv_transposed(I64(n), I64(ncols), P(off), P(cols), P(vals), P(x), P(y), ctypes.c_double(alpha),
                                   ctypes.c_double(beta), I64(begin), I64(end))
        return y

    def add_elements(self, off, ops, symmetric=False):
        """Empty rows of capacities diff(off), then addElement per op (row, column, value,
        multiplicator) until a row is full.  Returns (CSR-ordered columns, values, failing op or -1)."""
        off = i32(off)
        cols, vals = np.full(int(off[-1]), -1, np.int32), np.zeros(int(off[-1]))
        pos = np.ascontiguousarray([(o[0], o[1]) for o in ops], np.int64).reshape(-1)
        val = np.ascontiguousarray([(o[2], o[3]) for o in ops], np.float64).reshape(-1)
        failed = self.L.orc_add_elements(I64(len(off) - 1), P(off), P(cols), P(vals), int(symmetric), P(pos), P(val),
                                         I64(len(ops)))
        return cols, vals, int(failed)

    def spmv_omp(self, off, cols, vals, x, threads):
        off, cols, vals, x = i32(off), i32(cols), f64(vals), f64(x)
        n = len(off) - 1
        y = np.zeros(n)
        self.L.orc_spmv_f64_omp(I64(n), P(off), P(cols), P(vals), P(x), P(y), threads)
        return y

    def cg(self, off, cols, vals, b, x0=None, precond=0, **kw):
        off, cols, vals, b = i32(off), i32(cols), f64(vals), f64(b)
        n = len(b)
        x = np.zeros(n) if x0 is None else f64(x0).copy()
        st = state(**kw)
        self.L.orc_cg(I64(n), P(off), P(cols), P(vals), P(b), P(x), precond, ctypes.byref(st))
        return x, st

    def bicgstab(self, off, cols, vals, b, x0=None, precond=0, exact=False, **kw):
        off, cols, vals, b = i32(off), i32(cols), f64(vals), f64(b)
        n = len(b)
        x = np.zeros(n) if x0 is None else f64(x0).copy()
        st = state(**kw)
        self.L.orc_bicgstab(I64(n), P(off), P(cols), P(vals), P(b), P(x), precond, int(exact), ctypes.byref(st))
        return x, st

    def tfqmr(self, off, cols, vals, b, x0=None, precond=0, **kw):
        off, cols, vals, b = i32(off), i32(cols), f64(vals), f64(b)
        n = len(b)
        x = np.zeros(n) if x0 is None else f64(x0).copy()
        st = state(**kw)
        self.L.orc_tfqmr(I64(n), P(off), P(cols), P(vals), P(b), P(x), precond, ctypes.byref(st))
        return x, st

    def bicgstabl(self, off, cols, vals, b, x0=None, precond=0, ell=1, exact=False, **kw):
        off, cols, vals, b = i32(off), i32(cols), f64(vals), f64(b)
        n = len(b)
        x = np.zeros(n) if x0 is None else f64(x0).copy()
        st = state(**kw)
        self.L.orc_bicgstabl(I64(n), P(off), P(cols), P(vals), P(b), P(x), precond, int(ell), int(exact),
                             ctypes.byref(st))
        return x, st

    def idrs(self, off, cols, vals, b, x0=None, precond=0, s=4, smoothing=False, **kw):
        off, cols, vals, b = i32(off), i32(cols), f64(vals), f64(b)
        n = len(b)
        x = np.zeros(n) if x0 is None else f64(x0).copy()
        st = state(**kw)
        self.L.orc_idrs(I64(n), P(off), P(cols), P(vals), P(b), P(x), precond, int(s), int(smoothing),
                        ctypes.byref(st))
        return x, st

    def gmres(self, off, cols, vals, b, x0=None, precond=0, variant=3, rmin=10, rmax=10, dmin=3, dmax=3, **kw):
        off, cols, vals, b = i32(off), i32(cols), f64(vals), f64(b)
        n = len(b)
        x = np.zeros(n) if x0 is None else f64(x0).copy()
        st = state(**kw)
        self.L.orc_gmres(I64(n), P(off), P(cols), P(vals), P(b), P(x), precond, variant, rmin, rmax, dmin, dmax,
                         ctypes.byref(st))
        return x, st

    def jacobi(self, off, cols, vals, b, x0=None, omega=1.0, period=4, **kw):
        off, cols, vals, b = i32(off), i32(cols), f64(vals), f64(b)
        n = len(b)
        x = np.zeros(n) if x0 is None else f64(x0).copy()
        st = state(**kw)
        self.L.orc_jacobi(I64(n), P(off), P(cols), P(vals), P(b), P(x), ctypes.c_double(omega), I64(period),
                          ctypes.byref(st))
        return x, st

    def ilu0(self, off, cols, vals):
        off, cols, vals = i32(off), i32(cols), f64(vals)
        n = len(off) - 1
        nnz = int(off[-1])
        lp, up = np.zeros(n + 1, np.int32), np.zeros(n + 1, np.int32)
        lc, uc = np.zeros(nnz, np.int32), np.zeros(nnz, np.int32)
        lv, uv = np.zeros(nnz), np.zeros(nnz)
        self.L.orc_ilu0_factor(I64(n), P(off), P(cols), P(vals), P(lp), P(lc), P(lv), P(up), P(uc), P(uv))
        return (lp, lc[:lp[-1]], lv[:lp[-1]]), (up, uc[:up[-1]], uv[:up[-1]])

    def ilu0_solve(self, L, U, b):
        b = f64(b)
        x = np.zeros(len(b))
        self.L.orc_ilu0_solve(I64(len(b)), P(L[0]), P(L[1]), P(L[2]), P(U[0]), P(U[1]), P(U[2]), P(b), P(x))
        return x

    def true_residual(self, off, cols, vals, x, b):
        off, cols, vals, x, b = i32(off), i32(cols), f64(vals), f64(x), f64(b)
        return self.L.orc_true_residual(I64(len(b)), P(off), P(cols), P(vals), P(x), P(b))


class Reference:
    """TNL's own Host code path (oracle/_ref/libtnlref.so)."""

    FORMATS = dict(CSR=0, AdaptiveCSR=1, Ellpack=2, SlicedEllpack32=3, SlicedEllpack64=4, BiEllpack=5,
                   ChunkedEllpack=6, SortedCSR=7)

    @classmethod
    def load(cls):
        if not os.path.exists(REF_SO):
            return None
        r = cls()
        r.L = ctypes.CDLL(REF_SO)
        for name in ["ref_chunked_ellpack", "ref_adaptive_csr_blocks", "ref_fill", "ref_spmv", "ref_save",
                     "ref_load", "ref_spmv_symmetric", "ref_spmv_transposed", "ref_save_symmetric",
                     "ref_add_elements", "ref_row_counts", "ref_copy"]:
            getattr(r.L, name).restype = ctypes.c_int64
        r.L.ref_last_error.restype = ctypes.c_char_p
        r.L.ref_spmv_csr_time.restype = ctypes.c_double
        r.L.ref_true_residual.restype = ctypes.c_double
        return r

    def csr_offsets(self, sizes):
        sizes = i32(sizes)
        off = np.zeros(len(sizes) + 1, np.int32)
        self.L.ref_csr_offsets(P(sizes), I64(len(sizes)), P(off))
        return off

    def ellpack_dims(self, sizes, col_major=True):
        sizes = i32(sizes)
        w, a = ctypes.c_int32(), ctypes.c_int32()
        rc = self.L.ref_ellpack_dims(P(sizes), I64(len(sizes)), int(col_major), ctypes.byref(w), ctypes.byref(a))
        return rc, w.value, a.value

    def sliced(self, sizes, S):
        sizes = i32(sizes)
        ns = (len(sizes) + S - 1) // S
        off = np.zeros(ns + 1, np.int32)
        ss = np.zeros(ns, np.int32)
        self.L.ref_sliced_ellpack(P(sizes), I64(len(sizes)), S, P(off), P(ss))
        return off, ss

    def biellpack(self, sizes):
        sizes = i32(sizes)
        n = len(sizes)
        strips = (n + 31) // 32
        perm = np.zeros(n, np.int32)
        gp = np.zeros(strips * 6 + 1, np.int32)
        self.L.ref_biellpack(P(sizes), I64(n), P(perm), P(gp))
        return perm, gp

    def chunked(self, sizes):
        sizes = i32(sizes)
        n = len(sizes)
        cap = max(n, 1) * 256
        sl = np.zeros(4 * max(n, 1), np.int32)
        s2c, s2s = np.zeros(n, np.int32), np.zeros(n, np.int32)
        sp = np.zeros(n + 1, np.int32)
        c2s = np.full(cap, -1, np.int32)
        st = np.zeros(1, np.int32)
        ns = self.L.ref_chunked_ellpack(P(sizes), I64(n), P(sl), P(s2c), P(s2s), P(sp), P(c2s), I64(cap), P(st))
        return dict(nslices=ns, slices=sl[:4 * ns], seg_to_chunk=s2c, seg_to_slice=s2s, seg_ptrs=sp,
                    chunks_to_seg=c2s[:ns * 256], storage=int(st[0]))

    def adaptive_blocks(self, sizes, value_size):
        sizes = i32(sizes)
        n = len(sizes)
        cap = 8 * (n + 16)
        b = np.zeros(6 * cap, np.int32)
        c = self.L.ref_adaptive_csr_blocks(P(sizes), I64(n), value_size, P(b), I64(cap))
        if c == -2:
            return None  # LONG rows: not constructible by a host-only build of the reference
        return b[:6 * c].reshape(-1, 6)

    def sorted_segments(self, sizes):
        sizes = i32(sizes)
        n = len(sizes)
        eo, inv, perm = np.zeros(n + 1, np.int32), np.zeros(n, np.int32), np.zeros(n, np.int32)
        self.L.ref_sorted_segments(P(sizes), I64(n), P(eo), P(inv), P(perm))
        return eo, inv, perm

    def fill(self, fmt, off, cols, vals, ncols, col_major=True):
        off, cols, vals = i32(off), i32(cols), f64(vals)
        n = len(off) - 1
        cap = int(self.L.ref_fill(self.FORMATS[fmt], int(col_major), I64(n), I64(ncols), P(off), P(cols), P(vals),
                                  None, None, I64(0)))
        oc, ov = np.zeros(cap, np.int32), np.zeros(cap)
        self.L.ref_fill(self.FORMATS[fmt], int(col_major), I64(n), I64(ncols), P(off), P(cols), P(vals), P(oc),
                        P(ov), I64(cap))
        return oc, ov

    def copy(self, src_fmt, dst_fmt, off, cols, vals, ncols, src_col_major=True, dst_col_major=True):
        """TNL's copySparseMatrix (SparseOperations.hpp:600-690) from a matrix of
        format src_fmt built from the CSR arrays into one of format dst_fmt: the
        target's column indexes and values."""
        off, cols, vals = i32(off), i32(cols), f64(vals)
        n = len(off) - 1
        args = (self.FORMATS[src_fmt], int(src_col_major), self.FORMATS[dst_fmt], int(dst_col_major), I64(n),
                I64(ncols), P(off), P(cols), P(vals))
        cap = int(self.L.ref_copy(*args, None, None, I64(0)))
        oc, ov = np.zeros(cap, np.int32), np.zeros(cap)
        self.L.ref_copy(*args, P(oc), P(ov), I64(cap))
        return oc, ov

    def add_elements(self, fmt, off, ncols, ops, symmetric=False, col_major=True):
        """TNL's SparseMatrix (general or symmetric) with row capacities diff(off), then
        addElement(row, column, value, multiplicator) for each op until one throws.
        Returns (storage columns, storage values, index of the failing op or -1, its message)."""
        off = i32(off)
        n = len(off) - 1
        cols, vals = np.full(int(off[-1]), -1, np.int32), np.zeros(int(off[-1]))
        pos = np.ascontiguousarray([(o[0], o[1]) for o in ops], np.int64).reshape(-1)
        val = np.ascontiguousarray([(o[2], o[3]) for o in ops], np.float64).reshape(-1)
        failed = np.zeros(1, np.int64)
        args = (self.FORMATS[fmt], int(col_major), int(symmetric), I64(n), I64(ncols), P(off), P(cols), P(vals),
                P(pos), P(val), I64(len(ops)), P(failed))
        cap = int(self.L.ref_add_elements(*args, None, None, I64(0)))
        oc, ov = np.zeros(cap, np.int32), np.zeros(cap)
        self.L.ref_add_elements(*args, P(oc), P(ov), I64(cap))
        msg = self.L.ref_last_error().decode() if failed[0] >= 0 else ""
        return oc, ov, int(failed[0]), msg

    def row_counts(self, fmt, off, cols, vals, ncols, symmetric=False, col_major=True):
        """getRowCapacity per row, getRowCapacities, getCompressedRowLengths and
        getNonzeroElementsCount of the matrix TNL builds from these CSR arrays (as fill)."""
        off, cols, vals = i32(off), i32(cols), f64(vals)
        n = len(off) - 1
        cap, stored, nz = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n, np.int32)
        total = self.L.ref_row_counts(self.FORMATS[fmt], int(col_major), int(symmetric), I64(n), I64(ncols), P(off),
                                      P(cols), P(vals), P(cap), P(stored), P(nz))
        return cap, stored, nz, int(total)

    def save_file(self, fmt, off, cols, vals, ncols, path, col_major=True):
        """SparseMatrix::save of the matrix TNL builds from these CSR arrays (as fill)."""
        off, cols, vals = i32(off), i32(cols), f64(vals)
        rc = self.L.ref_save(self.FORMATS[fmt], int(col_major), I64(len(off) - 1), I64(ncols), P(off), P(cols),
                             P(vals), os.fsencode(path))
        if rc < 0:
            raise RuntimeError(self.L.ref_last_error().decode())

    def load_file(self, fmt, path, col_major=True):
        """SparseMatrix::load into TNL's matrix of this format: (rows, columns), column indexes, values."""
        dims = np.zeros(2, np.int64)
        st = self.L.ref_load(self.FORMATS[fmt], int(col_major), os.fsencode(path), P(dims), None, None, I64(0))
        if st < 0:
            raise RuntimeError(self.L.ref_last_error().decode())
        oc, ov = np.zeros(st, np.int32), np.zeros(st)
        self.L.ref_load(self.FORMATS[fmt], int(col_major), os.fsencode(path), P(dims), P(oc), P(ov), I64(st))
        return (int(dims[0]), int(dims[1])), oc, ov

    def spmv(self, fmt, off, cols, vals, x, y=None, alpha=1.0, beta=0.0, begin=0, end=0, col_major=True):
        off, cols, vals, x = i32(off), i32(cols), f64(vals), f64(x)
        n = len(off) - 1
        y = np.zeros(n) if y is None else f64(y).copy()
        self.L.ref_spmv(self.FORMATS[fmt], int(col_major), I64(n), I64(len(x)), P(off), P(cols), P(vals), P(x),
                        P(y), ctypes.c_double(alpha), ctypes.c_double(beta), I64(begin), I64(end))
        return y

    def spmv_symmetric(self, fmt, off, cols, vals, x, y=None, alpha=1.0, beta=0.0, begin=0, end=0, col_major=True):
        """vectorProduct of TNL's SymmetricMatrix built from lower-triangle CSR arrays."""
        off, cols, vals, x = i32(off), i32(cols), f64(vals), f64(x)
        n = len(off) - 1
        y = np.zeros(n) if y is None else f64(y).copy()
        self.L.ref_spmv_symmetric(self.FORMATS[fmt], int(col_major), I64(n), I64(len(x)), P(off), P(cols), P(vals),
                                  P(x), P(y), ctypes.c_double(alpha), ctypes.c_double(beta), I64(begin), I64(end))
        return y

    def spmv_transposed(self, fmt, off, cols, vals, ncols, x, y=None, alpha=1.0, beta=0.0, begin=0, end=0,
                        col_major=True):
        off, cols, vals, x = i32(off), i32(cols), f64(vals), f64(x)
        n = len(off) - 1
        y = np.zeros(ncols) if y is None else f64(y).copy()
        self.L.ref_spmv_transposed(self.FORMATS[fmt], int(col_major), I64(n), I64(ncols), P(off), P(cols), P(vals),
                                   P(x), P(y), ctypes.c_double(alpha), ctypes.c_double(beta), I64(begin), I64(end))
        return y

    def spmv_symmetric_time(self, off, cols, vals, x, reps, threads):
        """Median seconds of TNL's Host SymmetricMatrix<CSR> vectorProduct (lower-triangle
        CSR arrays, the matrix built once) and its last result."""
        off, cols, vals, x = i32(off), i32(cols), f64(vals), f64(x)
        n = len(off) - 1
        y = np.zeros(n)
        self.L.ref_spmv_symmetric_time.restype = ctypes.c_double
        t = self.L.ref_spmv_symmetric_time(I64(n), P(off), P(cols), P(vals), P(x), P(y), int(reps), int(threads))
        return t, y

    def save_symmetric_file(self, fmt, off, cols, vals, ncols, path, col_major=True):
        off, cols, vals = i32(off), i32(cols), f64(vals)
        rc = self.L.ref_save_symmetric(self.FORMATS[fmt], int(col_major), I64(len(off) - 1), I64(ncols), P(off),
                                       P(cols), P(vals), os.fsencode(path))
        if rc < 0:
            raise RuntimeError(self.L.ref_last_error().decode())

    def spmv_f32(self, off, cols, vals, x):
        off, cols = i32(off), i32(cols)
        vals = np.ascontiguousarray(vals, np.float32)
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros(len(off) - 1, np.float32)
        self.L.ref_spmv_csr_f32(I64(len(off) - 1), I64(len(x)), P(off), P(cols), P(vals), P(x), P(y))
        return y

    def spmv_time(self, off, cols, vals, x, reps, threads):
        off, cols, vals, x = i32(off), i32(cols), f64(vals), f64(x)
        y = np.zeros(len(off) - 1)
        t = self.L.ref_spmv_csr_time(I64(len(off) - 1), I64(len(x)), P(off), P(cols), P(vals), P(x), P(y), reps,
                                     threads)
        return t, y

    def _solve(self, fn, off, cols, vals, b, x0, *args, **kw):
        off, cols, vals, b = i32(off), i32(cols), f64(vals), f64(b)
        n = len(b)
        x = np.zeros(n) if x0 is None else f64(x0).copy()
        st = state(**kw)
        fn(I64(n), P(off), P(cols), P(vals), P(b), P(x), *args, ctypes.byref(st))
        return x, st

    def cg(self, off, cols, vals, b, x0=None, precond=0, **kw):
        return self._solve(self.L.ref_cg, off, cols, vals, b, x0, precond, **kw)

    def gmres(self, off, cols, vals, b, x0=None, precond=0, variant=3, rmin=10, rmax=10, dmin=3, dmax=3, **kw):
        return self._solve(self.L.ref_gmres, off, cols, vals, b, x0, precond, variant, rmin, rmax, dmin, dmax, **kw)

    def bicgstab(self, off, cols, vals, b, x0=None, precond=0, exact=False, **kw):
        return self._solve(self.L.ref_bicgstab, off, cols, vals, b, x0, precond, int(exact), **kw)

    def tfqmr(self, off, cols, vals, b, x0=None, precond=0, **kw):
        return self._solve(self.L.ref_tfqmr, off, cols, vals, b, x0, precond, **kw)

    def bicgstabl(self, off, cols, vals, b, x0=None, precond=0, ell=1, exact=False, **kw):
        return self._solve(self.L.ref_bicgstabl, off, cols, vals, b, x0, precond, int(ell), int(exact), **kw)

    def idrs(self, off, cols, vals, b, x0=None, precond=0, s=4, smoothing=False, **kw):
        return self._solve(self.L.ref_idrs, off, cols, vals, b, x0, precond, int(s), int(smoothing), **kw)

    def jacobi(self, off, cols, vals, b, x0=None, omega=1.0, period=4, **kw):
        return self._solve(self.L.ref_jacobi, off, cols, vals, b, x0, ctypes.c_double(omega), I64(period), **kw)

    def read_mtx(self, path):
        """MatrixReader< CSR Host >::readMtx: (rows, columns, offsets, columns, values);
        RuntimeError with the reference's message on a bad file."""
        r, c, z = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        if self.L.ref_read_mtx(os.fsencode(path), ctypes.byref(r), ctypes.byref(c), ctypes.byref(z)) != 0:
            self.L.ref_last_error.restype = ctypes.c_char_p
            raise RuntimeError(self.L.ref_last_error().decode())
        off = np.zeros(r.value + 1, np.int32)
        col = np.zeros(z.value, np.int32)
        val = np.zeros(z.value)
        self.L.ref_read_mtx_arrays(P(off), P(col), P(val))
        return r.value, c.value, off, col, val

    def set_threads(self, threads):
        """OpenMP threads of TNL's Host algorithms in later calls."""
        self.L.ref_set_threads(ctypes.c_int(int(threads)))

    def ilu0_apply(self, off, cols, vals, b):
        off, cols, vals, b = i32(off), i32(cols), f64(vals), f64(b)
        x = np.zeros(len(b))
        self.L.ref_ilu0_apply(I64(len(b)), P(off), P(cols), P(vals), P(b), P(x))
        return x

    def true_residual(self, off, cols, vals, x, b):
        off, cols, vals, x, b = i32(off), i32(cols), f64(vals), f64(x), f64(b)
        return self.L.ref_true_residual(I64(len(b)), P(off), P(cols), P(vals), P(x), P(b))


# ------------------------------------------------------------------------ KATs
def load_kats():
    with open(KATS) as f:
        return json.load(f)["cases"]


def kat_csr(case):
    """(offsets, cols, vals, x, y_expected, columns) of a KAT case; rows are CSR
    by capacity with unused slots holding column -1."""
    rows, ncols = case["rows"], case["columns"]
    ent = case["entries"]
    if ent == "lower_triangular_jplus1":
        ent = [[i, j, j + 1] for i in range(rows) for j in range(i + 1)]
    elif ent == "long_rows_jplusr":
        ent = [[r, j, j + r] for r in range(rows) for j in range(ncols)]
    per = [[] for _ in range(rows)]
    for r, c, v in ent:
        per[r].append((c, v))
    caps = case["capacities"] or [len(p) for p in per]
    off = np.zeros(rows + 1, np.int32)
    off[1:] = np.cumsum(caps)
    cols = np.full(off[-1], -1, np.int32)
    vals = np.zeros(off[-1])
    for r in range(rows):
        for k, (c, v) in enumerate(sorted(per[r])):
            cols[off[r] + k] = c
            vals[off[r] + k] = v
    return off, cols, vals, np.array(case["x"], np.float64), np.array(case["y"], np.float64), ncols


def dense(off, cols, vals, ncols):
    n = len(off) - 1
    A = np.zeros((n, ncols))
    for r in range(n):
        for g in range(off[r], off[r + 1]):
            if cols[g] >= 0:
                A[r, cols[g]] += vals[g]
    return A
