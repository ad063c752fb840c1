"""Batched CG over many independent SPD systems (tnlmi_batched_cg_*), the
seeded variable-coefficient diffusion batch, and the autotuner that picks the
segment format, slice size, workgroup size and rows per wavefront of the
persistent kernel."""
import ctypes
import json
import os
import time

import numpy as np

from . import _lib
from . import SparseMatrix, _stream, _ptr, CSR, ELLPACK, SLICED_ELLPACK, FORMAT_NAMES


def diffusion_batch(nsys, smin=16, smax=128, seed=0):
    """The batch as one block-diagonal CSR matrix: (sides, offsets, columns,
    values, sys_rows) with sys_rows[k] .. sys_rows[k+1] the rows of system k."""
    L = _lib.lib()
    sides = np.zeros(nsys, np.int32)
    rows, nnz = ctypes.c_int64(), ctypes.c_int64()
    L.tnlmi_gen_batched_diffusion(nsys, smin, smax, seed, _ptr(sides), ctypes.byref(rows), ctypes.byref(nnz), None,
                                  None, None)
    off = np.zeros(rows.value + 1, np.int32)
    col = np.zeros(nnz.value, np.int32)
    val = np.zeros(nnz.value, np.float64)
    L.tnlmi_gen_batched_diffusion(nsys, smin, smax, seed, _ptr(sides), ctypes.byref(rows), ctypes.byref(nnz),
                                  _ptr(off), _ptr(col), _ptr(val))
    sys_rows = np.zeros(nsys + 1, np.int64)
    sys_rows[1:] = np.cumsum(sides.astype(np.int64) ** 2)
    return sides, off, col, val, sys_rows


class BatchedCG:
    """One persistent launch solving every diagonal block of A (FP64 values).
    lanes_per_row (1, 2, 4, 8): the lanes that sum one row, so a wavefront
    covers 64 / lanes_per_row rows."""

    def __init__(self, A, sys_rows, A_lo=None, block_size=512, lanes_per_row=1, stream=None):
        self.sys_rows = np.ascontiguousarray(sys_rows, np.int64)
        self.nsys = len(self.sys_rows) - 1
        self.A, self.A_lo = A, A_lo
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().tnlmi_batched_cg_create(ctypes.byref(A._m), ctypes.byref(A_lo._m) if A_lo else None,
                                                      _ptr(self.sys_rows), self.nsys, block_size, lanes_per_row,
                                                      _stream(stream), ctypes.byref(h)))
        self.handle = h
        import torch
        self.iters = torch.zeros(self.nsys, dtype=torch.int64, device="cuda")
        self.res = torch.zeros(self.nsys, dtype=torch.float64, device="cuda")

    def __del__(self):
        h = getattr(self, "handle", None)
        try:
            if h is not None and h.value and _lib._lib is not None:
                _lib.lib().tnlmi_batched_cg_destroy(h)
        except (AttributeError, TypeError):
            pass
        self.handle = None

    def solve(self, b, x, tolerance=1e-10, max_iterations=100_000, mixed=False, inner_tolerance=1e-4,
              use_graph=True, stream=None):
        """Asynchronous; returns the per-system (iterations, residues) device tensors."""
        _lib.check(_lib.lib().tnlmi_batched_cg_solve(self.handle, int(mixed), ctypes.c_void_p(b.data_ptr()),
                                                     ctypes.c_void_p(x.data_ptr()), tolerance, max_iterations,
                                                     inner_tolerance, ctypes.c_void_p(self.iters.data_ptr()),
                                                     ctypes.c_void_p(self.res.data_ptr()), int(use_graph),
                                                     _stream(stream)))
        return self.iters, self.res


CANDIDATES = [(CSR, 0), (SLICED_ELLPACK, 32), (SLICED_ELLPACK, 64), (ELLPACK, 0)]
BLOCKS = [256, 512, 1024]
LANES = [1, 2, 4]  # lanes per row: 64, 32 or 16 rows per wavefront


def _cache_path():
    return os.environ.get("TNL_AMD_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "tnl_amd", "autotune.json"))


def _device_name():
    import torch
    p = torch.cuda.get_device_properties(0)
    return f"{p.name}:{getattr(p, 'gcnArchName', '')}"


def autotune(off, cols, vals, sys_rows, sample=256, candidates=CANDIDATES, blocks=BLOCKS, lanes=LANES,
             tolerance=1e-10, cache=True, log=None):
    """Times every (format, slice size, workgroup size, rows per wavefront)
    candidate on the first `sample` systems and returns the fastest as a dict;
    results are cached in a JSON file keyed by the batch's shape, the device,
    the candidate lists and the tolerance."""
    import torch
    nsys = len(sys_rows) - 1
    key = (f"batched_cg:nsys={nsys}:rows={int(sys_rows[-1])}:nnz={int(off[-1])}:sample={sample}"
           f":device={_device_name()}:candidates={sorted((int(f), int(p)) for f, p in candidates)}"
           f":blocks={sorted(int(b) for b in blocks)}:lanes={sorted(int(t) for t in lanes)}"
           f":tolerance={float(tolerance)!r}")
    path = _cache_path()
    if cache and os.path.exists(path):
        with open(path) as f:
            db = json.load(f)
        if key in db:
            return db[key]
    k = min(sample, nsys)
    rows = int(sys_rows[k])
    o = off[: rows + 1]
    c, v = cols[: int(o[-1])], vals[: int(o[-1])]
    b = torch.ones(rows, dtype=torch.float64, device="cuda")
    x = torch.zeros(rows, dtype=torch.float64, device="cuda")
    trials = []
    for fmt, param in candidates:
        A = SparseMatrix.from_csr(rows, rows, o, c, v, fmt=fmt, param=param)
        for bs in blocks:
            for lpr in lanes:
                bc = BatchedCG(A, sys_rows[: k + 1], block_size=bs, lanes_per_row=lpr)
                x.zero_()
                bc.solve(b, x, tolerance, use_graph=False)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(2):
                    x.zero_()
                    bc.solve(b, x, tolerance, use_graph=False)
                torch.cuda.synchronize()
                t = (time.perf_counter() - t0) / 2
                trials.append(dict(format=FORMAT_NAMES[fmt], format_id=fmt, param=param, block=bs, lanes_per_row=lpr,
                                   rows_per_wavefront=64 // lpr, seconds=t))
                if log:
                    log(f"autotune {FORMAT_NAMES[fmt]}{param or ''} block {bs} lanes/row {lpr}: {t * 1e3:.2f} ms")
        del A
    best = min(trials, key=lambda d: d["seconds"])
    result = dict(best, trials=trials)
    if cache:
        db = {}
        if os.path.exists(path):
            with open(path) as f:
                db = json.load(f)
        db[key] = result
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            json.dump(db, f, indent=1)
    return result
