"""Benchmarks of the hot path; each workload prints one JSON line (rank 0).

Default workload (`--workload spmv-powerlaw`): FP64 CSR SpMV y = A x on the
power-law matrix through the reference's own call, A.vectorProduct(x, y) with
the default launch configuration (tnlmi_spmv).  For this matrix the library
runs the CSR matrix's automatic windowed plan, built by the first call
(w_phase1 forms the products of each 8192-column window with x in LDS,
w_phase2 sums them per row block in LDS, w_fixup adds the pieces of rows
longer than a block); the XCD column-split plan and the plain tiled CSR
kernel are timed beside it and reported in `config`.  4,194,304 x 4,194,304,
row lengths ~ Zipf(2) on
[1, 4096] scaled to mean 16 (67,108,864 nonzeros), uniform distinct sorted
columns, values ~ N(0,1), x ~ U[-1,1], int32 indices, FP64 values/vectors, all
resident in HBM before the timed region; value = algorithmic bytes / time with
bytes per SpMV = 8*nnz + 4*nnz + 4*(N+1) + 8*N (x, once) + 8*N (y).
The same line carries `time_to_solution`: Jacobi-CG to a 1e-10 relative FP64
true residual on the 3D 7-point Poisson 256^3 matrix, FP64 CG (int32 columns,
and the 8-bit column codes the mixed solver reads) against the reliable-update
FP32 CG on BF16 and FP32 values, every variant stopped on the same rule
(`--no-solve` skips it).

Other workloads (same JSON contract, their own `config.workload`):
  spmv-laplace   FP64 CSR SpMV on the 2D 5-point Laplacian 512x512, 100 timed
                 repetitions, cold (a 512 MiB stream between repetitions) and warm
  spmv-formats   the power-law matrix in every segment format, FP64 and FP32
  cg-poisson     time to a 1e-10 relative FP64 residual on the 3D 7-point Poisson
                 256^3 matrix: FP64 Jacobi-CG vs FP64 iterative refinement around an
                 FP32- or BF16-valued inner Jacobi-CG
  batched-cg     4,096 variable-coefficient diffusion systems (sides log-uniform on
                 [16, 128]) in one persistent graph-launched batched CG, vs a loop of
                 per-system CG solves
  gmres-convdiff GMRES(30) to a 1e-10 true residual on 1024^2 convection-diffusion
                 with the learned sparse approximate inverse (trained first on 2,000
                 seeded 256^2 instances) vs Jacobi and ILU(0)

All inputs are synthetic, from the package's seeded host generators.
`--gpus N` with N > 1 exits: every GPU box of this build has one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SpMV GB/s (% of HBM peak); time-to-1e-10 residual: mixed/learned vs plain FP64"
HBM_PEAK_GBS = 8000.0
N_ROWS = 4_194_304
MEAN_NNZ = 16.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def spmv_bytes(n, nnz, sv=8, sx=8):
    return sv * nnz + 4 * nnz + 4 * (n + 1) + sx * n + sx * n


def _oracle_lib():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    return oracle_lib


def cpu_spmv_baseline(off, cols, vals, x, y_gpu, nbytes, budget_s):
    """TNL's Host CSR vectorProduct (OpenMP) timed on this host's cores on the
    same matrix; bounded to about budget_s seconds of CPU work."""
    import numpy as np
    oracle_lib = _oracle_lib()
    ref = oracle_lib.Reference.load()
    cores = cpu_cores()
    kind = "reference" if ref is not None else "port"
    if ref is not None:
        t1, y = ref.spmv_time(off, cols, vals, x, 1, cores)
        reps = int(max(3, min(200, budget_s / max(t1, 1e-6))))
        t, y = ref.spmv_time(off, cols, vals, x, reps, cores)
    else:
        orc = oracle_lib.Oracle()
        t0 = time.perf_counter()
        y = orc.spmv_omp(off, cols, vals, x, cores)
        t1 = time.perf_counter() - t0
        reps = int(max(3, min(200, budget_s / max(t1, 1e-6))))
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            y = orc.spmv_omp(off, cols, vals, x, cores)
            ts.append(time.perf_counter() - t0)
        t = sorted(ts)[len(ts) // 2]
    err = float(np.abs(y_gpu - y).max() / max(np.abs(y).max(), 1e-300))
    # the same product on one thread (a few repetitions within a quarter of the budget)
    reps1 = int(max(1, min(5, budget_s / 4 / max(t * cores, 1e-6))))
    if ref is not None:
        t_one, _ = ref.spmv_time(off, cols, vals, x, reps1, 1)
    else:
        ts = []
        for _ in range(reps1):
            t0 = time.perf_counter()
            orc.spmv_omp(off, cols, vals, x, 1)
            ts.append(time.perf_counter() - t0)
        t_one = sorted(ts)[len(ts) // 2]
    return dict(value=round(nbytes / t / 1e9, 3), unit="GB/s", cores=cores, kind=kind,
                sample=f"median of {reps} full products of the same {len(y)}-row matrix (TNL Host CSR "
                       f"vectorProduct, OpenMP schedule(dynamic,100), {cores} threads)",
                one_thread_value=round(nbytes / t_one / 1e9, 3), cpu_model=cpu_model(),
                seconds_per_spmv=t), err


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def timed(op, steps, warmup, flush=None):
    """Average ms per op over `steps` ops (HIP events on torch's current
    stream, the stream every op here is launched on).  With `flush`, each op
    is timed alone after the flush and the median is returned."""
    import torch
    for _ in range(warmup):
        op()
    torch.cuda.synchronize()
    if flush is None:
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(steps):
            op()
        stop.record()
        torch.cuda.synchronize()
        return start.elapsed_time(stop) / steps
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for a, b in ev:
        flush()
        a.record()
        op()
        b.record()
    torch.cuda.synchronize()
    ts = sorted(a.elapsed_time(b) for a, b in ev)
    return ts[len(ts) // 2]


def base_record(args, value, unit, ms, dtype, data, config, higher=True):
    return {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(ms, 6), "higher_is_better": higher, "scaling": "weak", "vs_baseline": None,
            "dtype": dtype, "data": data, "config": config}


def roofline(gbs, traffic=None):
    return {"bound": "hbm", "achieved": round(gbs, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic}


# --------------------------------------------------------------- workloads
def wl_spmv_powerlaw(args):
    import torch
    import tnl_amd

    t0 = time.perf_counter()
    off, cols, vals = tnl_amd.zipf_csr(args.rows, s=2.0, kmax=4096, mean=MEAN_NNZ, seed=0)
    x = tnl_amd.uniform(args.rows, -1.0, 1.0, seed=0)
    n, nnz = args.rows, int(off[-1])
    log(f"generated {n} rows, {nnz} nnz in {time.perf_counter() - t0:.1f}s")

    A = tnl_amd.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=tnl_amd.CSR)
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros(n, dtype=torch.float64, device="cuda")
    nbytes = spmv_bytes(n, nnz)
    flops = 2 * nnz

    # plain tiled CSR kernel over the CSR arrays themselves (the BlockMerged
    # launch configuration, which never takes the automatic plan; reported beside)
    plain = (tnl_amd.BLOCK_MERGED, 1)
    ms_plain = timed(lambda: A.vectorProduct(xd, yd, launch=plain), max(10, args.steps // 4),
                     max(2, args.warmup // 4))
    log(f"CSR f64 SpMV, plain tiled kernel: {ms_plain * 1e3:.1f} us, {nbytes / (ms_plain * 1e-3) / 1e9:.1f} GB/s")
    # the XCD column-split plan (8 column ranges; reported beside the headline)
    P8 = tnl_amd.SpmvPlan(A, 8)
    ms_split = timed(lambda: P8.vectorProduct(xd, yd), max(10, args.steps // 4), max(2, args.warmup // 4))
    log(f"CSR f64 SpMV, column-split plan: {ms_split * 1e3:.1f} us, {nbytes / (ms_split * 1e-3) / 1e9:.1f} GB/s")
    del P8
    torch.cuda.empty_cache()
    # the headline: the reference's default call, A.vectorProduct(x, y) with the
    # default launch configuration.  For this matrix (x wider than an L2, columns
    # far from the diagonal) the library builds the windowed plan on the first
    # call (here: the first warm-up step) and keeps it with the matrix.
    t0 = time.perf_counter()
    A.vectorProduct(xd, yd)
    torch.cuda.synchronize()
    log(f"first vectorProduct (builds the automatic windowed plan) {time.perf_counter() - t0:.2f}s")
    ms = timed(lambda: A.vectorProduct(xd, yd), args.steps, args.warmup)
    gbs = nbytes / (ms * 1e-3) / 1e9
    log(f"CSR f64 SpMV, A.vectorProduct (windowed plan): {ms * 1e3:.1f} us/step, {gbs:.1f} GB/s, "
        f"{flops / (ms * 1e-3) / 1e9:.1f} GFLOP/s")

    if args.copy_calibration:
        a = torch.empty(2**26, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_windowed_zipf4m_f64.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    out = base_record(
        args, round(gbs, 3), "GB/s", ms, "f64",
        "synthetic (seeded Zipf(2) power-law matrix generated on the host, N(0,1) values, U[-1,1] x)",
        {"workload": "FP64 CSR SpMV, power-law matrix", "rows": n, "nnz": nnz,
         "row_lengths": "Zipf(s=2.0) on [1,4096] scaled to mean 16", "index": "int32",
         "call": "A.vectorProduct(x, y) (tnlmi_spmv, default launch configuration)",
         "format": "CSR; vectorProduct runs the automatic windowed plan (8192-column windows: products formed "
                   "with x in LDS, then summed per row block in LDS; built by the first call, outside the "
                   "timing; one SpMV = w_phase1 + w_phase2 + w_fixup)", "bytes_per_spmv": nbytes,
         "gflops": round(flops / (ms * 1e-3) / 1e9, 3),
         "column_split_plan_GBs": round(nbytes / (ms_split * 1e-3) / 1e9, 3),
         "column_split_plan_ms": round(ms_split, 6),
         "plain_csr_kernel_GBs": round(nbytes / (ms_plain * 1e-3) / 1e9, 3),
         "plain_csr_kernel_ms": round(ms_plain, 6)})
    out["roofline"] = roofline(gbs, traffic)
    out["roofline"]["kernel"] = "one SpMV of the windowed plan (w_phase1 + w_phase2 + w_fixup), HIP events"
    if not args.no_cpu:
        y_gpu = yd.cpu().numpy()
        cb, err = cpu_spmv_baseline(off, cols, vals, x, y_gpu, nbytes, args.cpu_budget)
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample", "one_thread_value",
                                                  "cpu_model")}
        out["parity_max_rel_err_vs_cpu"] = err
        log(f"CPU baseline: {cb['value']} GB/s ({cb['seconds_per_spmv'] * 1e3:.1f} ms/SpMV, {cb['cores']} threads, "
            f"{cb['kind']}), GPU vs CPU max rel err {err:.2e}")
    else:
        out["cpu_baseline"] = None
    del A, xd, yd
    torch.cuda.empty_cache()
    if not args.no_solve:
        # the metric's second half: time to a 1e-10 relative FP64 residual on the
        # 3D 7-point Poisson 256^3 matrix, mixed precision against plain FP64
        t0 = time.perf_counter()
        out["time_to_solution"] = time_to_solution_poisson(args.side, args.delta, log_prefix="[time-to-1e-10] ")
        log(f"[time-to-1e-10] done in {time.perf_counter() - t0:.1f}s")
    return out


def wl_spmv_laplace(args):
    import torch
    import tnl_amd

    off, cols, vals = tnl_amd.laplace2d(512, 512)
    n, nnz = len(off) - 1, int(off[-1])
    x = tnl_amd.uniform(n, -1.0, 1.0, seed=0)
    A = tnl_amd.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=tnl_amd.CSR)
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros(n, dtype=torch.float64, device="cuda")
    nbytes = spmv_bytes(n, nnz)
    reps = 100
    # cold: stream 512 MiB (twice the 256 MB last-level cache) before each product
    big_a = torch.empty(2**26, dtype=torch.float64, device="cuda")
    big_b = torch.empty_like(big_a)
    ms_cold = timed(lambda: A.vectorProduct(xd, yd), reps, 5, flush=lambda: big_b.copy_(big_a))
    ms_warm = timed(lambda: A.vectorProduct(xd, yd), reps, 5)
    formats = {}
    for name, fmt, prm in (("Ellpack", tnl_amd.ELLPACK, 0), ("SlicedEllpack64", tnl_amd.SLICED_ELLPACK, 64)):
        B = tnl_amd.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=fmt, param=prm)
        t = timed(lambda: B.vectorProduct(xd, yd), reps, 5, flush=lambda: big_b.copy_(big_a))
        formats[name] = {"cold_ms": round(t, 6), "cold_GBs": round(nbytes / (t * 1e-3) / 1e9, 3)}
    A.vectorProduct(xd, yd)
    torch.cuda.synchronize()
    gbs = nbytes / (ms_cold * 1e-3) / 1e9
    log(f"Laplacian 512^2 CSR f64: cold {ms_cold * 1e3:.1f} us ({gbs:.1f} GB/s), warm {ms_warm * 1e3:.1f} us "
        f"({nbytes / (ms_warm * 1e-3) / 1e9:.1f} GB/s)")
    out = base_record(args, round(gbs, 3), "GB/s", ms_cold, "f64",
                      "synthetic (2D 5-point Laplacian, U[-1,1] x from seed 0)",
                      {"workload": "FP64 CSR SpMV, 2D Laplacian 512x512 (cold cache: 512 MiB streamed before "
                                   "each timed product; median of 100)", "rows": n, "nnz": nnz,
                       "bytes_per_spmv": nbytes, "warm_ms": round(ms_warm, 6),
                       "warm_GBs": round(nbytes / (ms_warm * 1e-3) / 1e9, 3), "other_formats": formats})
    out["roofline"] = roofline(gbs)
    if not args.no_cpu:
        cb, err = cpu_spmv_baseline(off, cols, vals, x, yd.cpu().numpy(), nbytes, args.cpu_budget)
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample", "one_thread_value",
                                                  "cpu_model")}
        out["parity_max_rel_err_vs_cpu"] = err
    else:
        out["cpu_baseline"] = None
    return out


def wl_spmv_formats(args):
    import numpy as np
    import torch
    import tnl_amd

    off, cols, vals = tnl_amd.zipf_csr(args.rows, s=2.0, kmax=4096, mean=MEAN_NNZ, seed=0)
    x = tnl_amd.uniform(args.rows, -1.0, 1.0, seed=0)
    n, nnz = args.rows, int(off[-1])
    xd = torch.from_numpy(x).cuda()
    xf = xd.float()
    yd = torch.zeros(n, dtype=torch.float64, device="cuda")
    yf = torch.zeros(n, dtype=torch.float32, device="cuda")
    y_ref = None
    steps = max(5, args.steps // 10)
    results = {}
    best = None
    variants = [("CSR (tiled, default)", tnl_amd.CSR, 0, None),
                ("CSR scalar (Fixed TPS 1)", tnl_amd.CSR, 0, (tnl_amd.FIXED, 1)),
                ("CSR vector (Warp)", tnl_amd.CSR, 0, (tnl_amd.WARP, 64)),
                ("CSR light (Fixed TPS 16)", tnl_amd.CSR, 0, (tnl_amd.FIXED, 16)),
                ("CSR dynamic grouping", tnl_amd.CSR, 0, (tnl_amd.DYNAMIC_GROUPING, 1)),
                ("AdaptiveCSR", tnl_amd.ADAPTIVE_CSR, 0, None),
                ("Ellpack", tnl_amd.ELLPACK, 0, None),
                ("SlicedEllpack64", tnl_amd.SLICED_ELLPACK, 64, None),
                ("BiEllpack", tnl_amd.BIELLPACK, 32, None),
                ("ChunkedEllpack", tnl_amd.CHUNKED_ELLPACK, 0, None),
                ("SortedCSR", tnl_amd.SORTED_CSR, 0, None)]
    A0 = tnl_amd.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=tnl_amd.CSR)
    P = tnl_amd.SpmvPlan(A0, 8)
    P.vectorProduct(xd, yd)
    torch.cuda.synchronize()
    y_ref = yd.cpu().numpy()
    scale = np.abs(y_ref).max()
    del A0
    for name, fmt, prm, launch in variants:
        rec = {}
        for prec in ("f64", "f32"):
            try:
                A = tnl_amd.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=fmt, param=prm,
                                                  value_dtype=torch.float64 if prec == "f64" else torch.float32)
            except tnl_amd._lib.OverflowError_ as e:
                rec = {"result": f"N/A: {e} (TNL throws std::overflow_error as well)"}
                break
            xin, yout = (xd, yd) if prec == "f64" else (xf, yf)
            ms = timed(lambda: A.vectorProduct(xin, yout, launch=launch), steps, 2)
            err = float(np.abs(yout.double().cpu().numpy() - y_ref).max() / scale)
            nb = spmv_bytes(n, nnz, 8 if prec == "f64" else 4, 8 if prec == "f64" else 4)
            rec[prec] = {"ms": round(ms, 4), "GBs": round(nb / (ms * 1e-3) / 1e9, 1), "max_rel_err": err}
            log(f"{name} {prec}: {ms * 1e3:.0f} us {nb / (ms * 1e-3) / 1e9:.0f} GB/s err {err:.1e}")
            del A
            torch.cuda.empty_cache()
        results[name] = rec
    nb64, nb32 = spmv_bytes(n, nnz), spmv_bytes(n, nnz, 4, 4)
    msp = timed(lambda: P.vectorProduct(xd, yd), steps, 2)
    results["CSR + column-split plan"] = {"f64": {"ms": round(msp, 4), "GBs": round(nb64 / (msp * 1e-3) / 1e9, 1)}}
    best = nb64 / (msp * 1e-3) / 1e9
    out = base_record(args, round(best, 3), "GB/s", msp, "f64",
                      "synthetic (seeded Zipf(2) power-law matrix, N(0,1) values, U[-1,1] x)",
                      {"workload": "SpMV on the power-law matrix in every segment format, FP64 and FP32 "
                                   "(value: the CSR column-split plan, FP64)", "rows": n, "nnz": nnz,
                       "bytes_per_spmv_f64": nb64, "bytes_per_spmv_f32": nb32, "formats": results})
    out["roofline"] = roofline(best)
    out["cpu_baseline"] = None
    return out


def _cpu_cg_per_iteration(off, cols, vals, b):
    """Seconds per iteration of TNL's Host Jacobi-CG (oracle/_ref, OpenMP, all
    host cores) on the same matrix: (time of 21 iterations - time of 1) / 20."""
    oracle_lib = _oracle_lib()
    ref = oracle_lib.Reference.load()
    if ref is None:
        return None
    ts = []
    for it in (1, 21):
        t0 = time.perf_counter()
        ref.cg(off, cols, vals, b, precond=1, tol=1e-300, maxit=it)
        ts.append(time.perf_counter() - t0)
    return max(ts[1] - ts[0], 1e-9) / 20


def time_to_solution_poisson(side, delta, log_prefix=""):
    """Jacobi-CG to ||b - A x|| / ||b|| <= 1e-10 on the 3D 7-point Poisson
    matrix of side^3 rows (b ~ N(0,1), seed 1), every variant stopped on that
    same true-residual rule: FP64 CG (int32 columns, and the 8-bit column
    codes the mixed solver reads) against the reliable-update FP32 CG on BF16
    and FP32 values; TNL's own rule (the preconditioned residue) reported
    beside them.  Returns a dict of results and the speed-ups."""
    import torch
    import tnl_amd
    from tnl_amd import solvers

    t0 = time.perf_counter()
    off, cols, vals = tnl_amd.poisson3d(side, side, side)
    n, nnz = len(off) - 1, int(off[-1])
    b = tnl_amd.normal(n, seed=1)
    log(f"{log_prefix}Poisson {side}^3: {n} rows, {nnz} nnz, generated in {time.perf_counter() - t0:.1f}s")
    bd = torch.from_numpy(b).cuda()
    tol = 1e-10
    res = {}
    A = tnl_amd.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=tnl_amd.ELLPACK)
    D = solvers.Diagonal(A)

    def run(label, solver):
        solver.setMatrix(A)
        solver.setConvergenceResidue(tol)
        solver.setMaxIterations(20_000)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        solver.solve(bd, x)  # warm-up (kernels, column codes, plan)
        x.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ok = solver.solve(bd, x)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        tr = solvers.true_residual(A, x, bd)
        rec = {"seconds": round(t, 4), "iterations": solver.getIterations(), "true_residual": tr,
               "converged_below_1e-10": bool(ok and tr <= tol)}
        if hasattr(solver, "getUpdates"):
            rec["fp64_residual_updates"] = solver.getUpdates()
        log(f"{log_prefix}{label}: {rec}")
        res[label] = rec

    for codes in (False, True):
        cg = solvers.TrueResidualCG(column_codes=codes)
        cg.setPreconditioner(D)
        run(f"fp64_cg[Ellpack, {'8-bit column codes' if codes else 'int32 columns'}]", cg)
    tnl_rule = solvers.CG()
    tnl_rule.setPreconditioner(D)
    run("fp64_cg_tnl_rule[Ellpack, int32 columns]", tnl_rule)
    for lo_name, lo_dtype in (("bf16", torch.bfloat16), ("fp32", torch.float32)):
        A_lo = tnl_amd.SparseMatrix(A.rows, A.columns, A.segments, A.values.to(lo_dtype), A.column_indexes)
        m = solvers.ReliableMixedCG()
        m.setInnerMatrix(A_lo)
        m.setDelta(delta)
        run(f"reliable_{lo_name}[Ellpack, 8-bit column codes]", m)
        del A_lo
    del A, D
    torch.cuda.empty_cache()
    ok = {k: v for k, v in res.items() if v["converged_below_1e-10"]}
    t32 = ok.get("fp64_cg[Ellpack, int32 columns]", {}).get("seconds")
    t8 = ok.get("fp64_cg[Ellpack, 8-bit column codes]", {}).get("seconds")
    mixed = {k: v for k, v in ok.items() if k.startswith("reliable")}
    best = min(mixed, key=lambda k: mixed[k]["seconds"]) if mixed else None
    tmx = mixed[best]["seconds"] if best else None
    return {"workload": f"Poisson {side}^3 ({n} rows, {nnz} nonzeros), Jacobi-CG, b ~ N(0,1) seed 1",
            "stopping_rule": "||b - A x|| / ||b|| <= 1e-10 for every variant (FP64 b - A x confirmed at the stop)",
            "best_mixed": best,
            "speedup_vs_fp64_int32_columns": round(t32 / tmx, 3) if t32 and tmx else None,
            "speedup_vs_fp64_same_column_codes": round(t8 / tmx, 3) if t8 and tmx else None,
            "results": res}


def wl_cg_poisson(args):
    import torch
    import tnl_amd
    from tnl_amd import solvers

    side = args.side
    t0 = time.perf_counter()
    off, cols, vals = tnl_amd.poisson3d(side, side, side)
    n, nnz = len(off) - 1, int(off[-1])
    b = tnl_amd.normal(n, seed=1)
    log(f"Poisson {side}^3: {n} rows, {nnz} nnz, generated in {time.perf_counter() - t0:.1f}s")
    bd = torch.from_numpy(b).cuda()
    tol = 1e-10
    res = {}

    def run(label, solver, A_check, per_it_bytes):
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ok = solver.solve(bd, x)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        tr = solvers.true_residual(A_check, x, bd)
        its = solver.getIterations()
        rec = {"seconds": round(t, 4), "converged": ok, "iterations": its, "true_residual": tr,
               "ms_per_iteration": round(t / max(its, 1) * 1e3, 4)}
        if per_it_bytes:
            rec["GBs_per_iteration_bytes"] = round(per_it_bytes * its / t / 1e9, 1)
        for extra in ("getInnerIterations", "getUpdates"):
            if hasattr(solver, extra):
                rec[extra[3:].lower()] = getattr(solver, extra)()
        log(f"{label}: {t:.3f}s, {rec}")
        res[label] = rec
        return rec

    nvec64, nvec32 = 11 * 8 * n, 11 * 4 * n  # vector passes of one fused iteration
    for fname, fmt in (("CSR", tnl_amd.CSR), ("Ellpack", tnl_amd.ELLPACK)):
        A = tnl_amd.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=fmt)
        ob = 4 * (n + 1) if fmt == tnl_amd.CSR else 0
        # warm-up of every kernel on a few iterations
        w = solvers.CG()
        w.setMatrix(A)
        w.setPreconditioner(solvers.Diagonal(A))
        w.setMaxIterations(3)
        w.solve(bd, torch.zeros(n, dtype=torch.float64, device="cuda"))
        # plain FP64 Jacobi-CG (TNL's CG with the Diagonal preconditioner)
        cg = solvers.CG()
        cg.setMatrix(A)
        cg.setPreconditioner(solvers.Diagonal(A))
        cg.setConvergenceResidue(tol)
        cg.setMaxIterations(20_000)
        run(f"fp64_cg[{fname}]", cg, A, 12 * nnz + ob + 16 * n + nvec64)
        for lo_name, lo_dtype, sv in (("fp32", torch.float32, 4), ("bf16", torch.bfloat16, 2)):
            A_lo = tnl_amd.SparseMatrix(A.rows, A.columns, A.segments, A.values.to(lo_dtype), A.column_indexes)
            m = solvers.ReliableMixedCG()
            m.setMatrix(A)
            m.setInnerMatrix(A_lo)
            m.setDelta(args.delta)
            m.setConvergenceResidue(tol)
            m.setMaxIterations(20_000)
            m.solve(bd, torch.zeros(n, dtype=torch.float64, device="cuda"))  # warm-up
            run(f"reliable_{lo_name}[{fname}]", m, A, (sv + 4) * nnz + ob + 8 * n + nvec32)
            if fname == "CSR" and lo_name == "fp32":
                ir = solvers.MixedCG()
                ir.setMatrix(A)
                ir.setInnerMatrix(A_lo)
                ir.setInnerTolerance(1e-4)
                ir.setConvergenceResidue(tol)
                ir.setMaxIterations(100)
                run("restarted_ir_fp32_inner1e-4[CSR]", ir, A, None)
            del A_lo
        del A
        torch.cuda.empty_cache()
    ok = {k: v for k, v in res.items() if v["converged"] and v["true_residual"] <= tol}
    fp64 = {k: v for k, v in ok.items() if k.startswith("fp64")}
    mixed = {k: v for k, v in ok.items() if k.startswith("reliable")}
    # TNL's CG stops on its preconditioned residue (here sqrt(6) x the true one): count it only if the true
    # residual also met the tolerance, else take the slowest FP64 time as the baseline
    base_k = min(fp64, key=lambda k: fp64[k]["seconds"]) if fp64 else max(
        (k for k in res if k.startswith("fp64")), key=lambda k: res[k]["seconds"])
    best_k = min(mixed, key=lambda k: mixed[k]["seconds"])
    t64, tmx = res[base_k]["seconds"], mixed[best_k]["seconds"]
    speedup = t64 / tmx
    lo = "bf16" if "bf16" in best_k else "fp32"
    out = base_record(args, round(speedup, 4), "x (best FP64 CG time / best mixed-precision time, to 1e-10)",
                      tmx * 1e3, f"{lo} values, f32 vectors and accumulation, f64 residual and x",
                      "synthetic (3D 7-point Poisson, b ~ N(0,1) from seed 1)",
                      {"workload": f"time to ||b-Ax||/||b|| <= 1e-10 on Poisson {side}^3, Jacobi-CG", "rows": n,
                       "nnz": nnz, "fp64_baseline": base_k, "best_mixed": best_k, "fp64_seconds": t64,
                       "mixed_seconds": tmx, "delta": args.delta, "results": res,
                       "note": "FP64 CG = TNL's CG (stops on the preconditioned residue sqrt(r.z)/||M^-1 b||); mixed "
                               "= FP32 CG with reliable updates stopping on the FP64 true residual; GBs = bytes of "
                               "one fused iteration (SpMV with x once + 11 vector passes) x iterations / time"})
    out["roofline"] = roofline(mixed[best_k]["GBs_per_iteration_bytes"])
    out["roofline"]["kernel"] = f"whole {best_k} solve (fused SpMV + vector kernels)"
    if not args.no_cpu:
        A_off, A_cols, A_vals = off, cols, vals
        per_it = _cpu_cg_per_iteration(A_off, A_cols, A_vals, b)
        if per_it:
            its = res[base_k]["iterations"]
            out["cpu_baseline"] = {"value": round(per_it * its, 3), "unit": "s (extrapolated FP64 Jacobi-CG solve)",
                                   "cores": cpu_cores(), "kind": "reference",
                                   "sample": f"TNL Host Jacobi-CG, (21 - 1 iterations) timed on the same matrix: "
                                             f"{per_it * 1e3:.1f} ms/iteration x {its} GPU iterations"}
    out.setdefault("cpu_baseline", None)
    return out


def wl_batched_cg(args):
    import numpy as np
    import torch
    import tnl_amd
    from tnl_amd import batched as B
    from tnl_amd import solvers

    nsys = args.systems
    t0 = time.perf_counter()
    sides, off, cols, vals, sys_rows = B.diffusion_batch(nsys, 16, 128, seed=11)
    n, nnz = int(sys_rows[-1]), int(off[-1])
    b = tnl_amd.normal(n, seed=12)
    log(f"batch: {nsys} systems, {n} rows, {nnz} nnz in {time.perf_counter() - t0:.1f}s")
    tune = B.autotune(off, cols, vals, sys_rows, sample=256, cache=False, log=log)
    fmt = tune["format_id"]
    A = tnl_amd.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=fmt, param=tune["param"])
    A32 = tnl_amd.SparseMatrix(A.rows, A.columns, A.segments, A.values.float(), A.column_indexes)
    bc = B.BatchedCG(A, sys_rows, A_lo=A32, block_size=tune["block"], lanes_per_row=tune["lanes_per_row"])
    bd = torch.from_numpy(b).cuda()
    tol = 1e-10
    out_res = {}
    for label, mixed in (("fp64", False), ("mixed", True)):
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        bc.solve(bd, x, tol, mixed=mixed, use_graph=True)  # capture + warm-up
        torch.cuda.synchronize()
        ts = []
        for _ in range(max(1, args.steps // 100)):
            x.zero_()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            bc.solve(bd, x, tol, mixed=mixed, use_graph=True)
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1) / 1e3)
        its = bc.iters.cpu().numpy()
        res = bc.res.cpu().numpy()
        xs = x.cpu().numpy()
        # true residual of a sample of systems (FP64 on the host)
        orc = _oracle_lib().Oracle()
        worst = 0.0
        for k in range(0, nsys, max(1, nsys // 64)):
            r0, r1 = int(sys_rows[k]), int(sys_rows[k + 1])
            o = off[r0:r1 + 1] - off[r0]
            worst = max(worst, orc.true_residual(o, cols[off[r0]:off[r1]] - r0, vals[off[r0]:off[r1]], xs[r0:r1],
                                                 b[r0:r1]))
        # algorithmic bytes: per CG iteration per system one FP64 SpMV + 10 vector passes
        rows_k = np.diff(sys_rows)
        nnz_k = off[sys_rows[1:]] - off[sys_rows[:-1]]
        per_it = 12 * nnz_k + 4 * (rows_k + 1) + 16 * rows_k + 80 * rows_k
        if mixed:
            per_it = 8 * nnz_k + 4 * (rows_k + 1) + 8 * rows_k + 40 * rows_k
        tsec = sorted(ts)[len(ts) // 2]
        out_res[label] = {"seconds": round(tsec, 5), "iterations_total": int(its.sum()),
                          "iterations_max": int(its.max()), "all_below_tol": bool((res < tol).all()),
                          "worst_sampled_true_residual": worst,
                          "GBs": round(float((per_it * its).sum()) / tsec / 1e9, 1)}
        log(f"batched {label}: {out_res[label]}")
    # per-system loop of plain CG solves (one solver call per system) on a sample
    sample = min(nsys, args.loop_sample)
    mats = []
    for k in range(sample):
        r0, r1 = int(sys_rows[k]), int(sys_rows[k + 1])
        o = (off[r0:r1 + 1] - off[r0]).astype(np.int32)
        Ak = tnl_amd.SparseMatrix.from_csr(r1 - r0, r1 - r0, o, cols[off[r0]:off[r1]] - r0, vals[off[r0]:off[r1]])
        mats.append((Ak, solvers.Diagonal(Ak), bd[r0:r1].clone(), torch.zeros(r1 - r0, dtype=torch.float64,
                                                                               device="cuda")))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for Ak, Mk, bk, xk in mats:
        s = solvers.CG()
        s.setMatrix(Ak)
        s.setPreconditioner(Mk)
        s.setConvergenceResidue(tol)
        s.solve(bk, xk)
    torch.cuda.synchronize()
    t_loop = (time.perf_counter() - t0) * nsys / sample
    best = min(out_res["fp64"]["seconds"], out_res["mixed"]["seconds"])
    out = base_record(args, round(t_loop / best, 3), "x (per-system CG loop time / batched time)", best * 1e3,
                      "f64 (and f32 inner / f64 refined)",
                      "synthetic (seeded variable-coefficient diffusion batch, b ~ N(0,1))",
                      {"workload": f"{nsys} independent SPD systems, Jacobi-CG to 1e-10 each", "rows": n, "nnz": nnz,
                       "autotune": {k: tune[k] for k in ("format", "param", "block", "rows_per_wavefront")},
                       "batched": out_res, "per_system_loop_seconds": round(t_loop, 3),
                       "loop_sample": f"{sample} systems timed, scaled by {nsys}/{sample}"})
    out["roofline"] = roofline(out_res["fp64"]["GBs"])
    out["roofline"]["kernel"] = "batched_cg64 (FP64)"
    out["roofline"]["note"] = ("achieved counts the bytes one CG iteration of each system is designed to move; a "
                               "system (<= 16,384 rows) stays in its CU and its XCD's L2 during its solve, so this "
                               "is not an HBM figure")
    out["cpu_baseline"] = None
    return out


def wl_gmres_convdiff(args):
    import numpy as np
    import torch
    import tnl_amd
    from tnl_amd import solvers
    from tnl_amd import spai

    tol = 1e-10
    todo = set(args.solvers.split(","))
    # training: seeded side x side instances, one Adam step each, on the GPU
    net = spai.Network(seed=args.spai_seed)
    t0 = time.perf_counter()
    hist = spai.train(net, sizes=(args.spai_side,), instances=args.spai_instances, steps_per_instance=1,
                      probes=args.spai_probes, lr=args.spai_lr, seed=100_000, log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    # the solve: an unseen m x m instance
    m = args.side
    off, cols, vals, pe = tnl_amd.convdiff(m, seed=args.solve_seed)
    n, nnz = len(off) - 1, int(off[-1])
    A = tnl_amd.SparseMatrix.from_csr(n, n, off, cols, vals)
    b = torch.from_numpy(tnl_amd.normal(n, seed=3)).cuda()
    log(f"convection-diffusion {m}^2: {n} rows, {nnz} nnz, Peclet {pe:.1f}")
    res = {}

    def run(label, make_pre, max_its, progress=None):
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pre, extra = make_pre()
        torch.cuda.synchronize()
        t_setup = time.perf_counter() - t0
        ok, its, r = spai.gmres_solve(A, b, x, pre, tolerance=tol, restart=30, max_iterations=max_its,
                                      variant=args.gmres_variant, log=progress)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        rec = {"seconds": round(t, 4), "setup_seconds": round(t_setup, 4), "converged": ok, "iterations": its,
               "true_residual": r, "ms_per_iteration": round((t - t_setup) / max(its, 1) * 1e3, 4)}
        rec.update(extra)
        log(f"{label}: {rec}")
        res[label] = rec
        return rec

    # warm-up of the GMRES kernels
    w = solvers.GMRES()
    w.setMatrix(A)
    w.setMaxIterations(5)
    w.solve(b, torch.zeros(n, dtype=torch.float64, device="cuda"))
    if "jacobi" in todo:
        run("jacobi", lambda: (solvers.Diagonal(A), {}), args.max_iterations)

    def learned():
        return solvers.Explicit(net.inverse(A)), {}
    if "learned_spai" in todo:
        run("learned_spai", learned, args.max_iterations)
    if "ilu0" in todo:
        run("ilu0", lambda: (solvers.ILU0(A), {}), args.ilu_max_iterations, progress=lambda m: log(f"[ilu0] {m}"))

    # the nonsymmetric alternative on the same system: BiCGStab (BICGStab.hpp),
    # restarted from its own x with a tighter target until the true residual
    # meets the tolerance (its stopping residue is the preconditioned one)
    def true_residual(x):
        y = torch.zeros(n, dtype=torch.float64, device="cuda")
        A.vectorProduct(x, y)
        return float(torch.linalg.norm(b - y) / torch.linalg.norm(b))

    def run_bicgstab(label, make_pre):
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pre, _ = make_pre()
        torch.cuda.synchronize()
        t_setup = time.perf_counter() - t0
        s = solvers.BICGStab()
        s.setMatrix(A)
        s.setPreconditioner(pre)
        s.setMaxIterations(args.max_iterations)
        target, its, ok, r = tol, 0, False, float("nan")
        for _ in range(6):
            s.setConvergenceResidue(target)
            ok = s.solve(b, x)
            its += s.getIterations()
            r = true_residual(x)
            if r <= tol or not ok:
                break
            target *= 0.5 * tol / r
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        rec = {"seconds": round(t, 4), "setup_seconds": round(t_setup, 4), "converged": bool(ok and r <= tol),
               "iterations": its, "true_residual": r}
        log(f"{label}: {rec}")
        res[label] = rec

    if "bicgstab" in todo:
        run_bicgstab("bicgstab_jacobi", lambda: (solvers.Diagonal(A), {}))
        run_bicgstab("bicgstab_learned_spai", learned)
    # probe loss of the two inverses on the solve matrix
    Z = torch.from_numpy(np.random.default_rng(5).integers(0, 2, (4, n)) * 2.0 - 1.0).cuda()
    L_learned, _ = net.loss_grad(A, Z)
    L_jacobi, _ = spai.Network(seed=args.spai_seed).loss_grad(A, Z)
    both = "jacobi" in res and "learned_spai" in res
    t_j = res["jacobi"]["seconds"] if both else 0.0
    t_l = res["learned_spai"]["seconds"] if both else 0.0
    ok = both and res["jacobi"]["converged"] and res["learned_spai"]["converged"]
    speedup = t_j / t_l if ok else 0.0
    nb = spmv_bytes(n, nnz)
    out = base_record(args, round(speedup, 4), "x (Jacobi-GMRES(30) time / learned-SPAI-GMRES(30) time, to 1e-10)",
                      t_l * 1e3, "f64 solve, bf16 network (f32 accumulation)",
                      "synthetic (seeded convection-diffusion instances; training on the GPU from random init)",
                      {"workload": f"GMRES(30) to ||b-Ax||/||b|| <= 1e-10 on {m}^2 upwind convection-diffusion",
                       "rows": n, "nnz": nnz, "peclet": pe, "results": res,
                       "training": {"instances": args.spai_instances, "side": args.spai_side,
                                    "probes": args.spai_probes, "lr": args.spai_lr, "seconds": round(t_train, 2),
                                    "loss_first_50": float(np.mean(hist[:50])) if len(hist) else None,
                                    "loss_last_50": float(np.mean(hist[-50:])) if len(hist) else None},
                       "probe_loss_on_solve_matrix": {"jacobi": L_jacobi, "learned": L_learned},
                       "gmres_variant": args.gmres_variant,
                       "note": "left-preconditioned GMRES(30); the solve time of the learned preconditioner includes "
                               "the network's inference on the solve matrix; ILU(0) capped at --ilu-max-iterations; "
                               "bicgstab_* are BiCGStab on the same system to the same true residual (beside the "
                               "GMRES comparison the value is quoted on)"})
    its_l = res["learned_spai"]["iterations"] if both else 0
    out["roofline"] = roofline(2 * nb * its_l / (t_l - res["learned_spai"]["setup_seconds"]) / 1e9 if both else 0.0)
    out["roofline"]["kernel"] = "SpMV with A and with M per GMRES iteration (orthogonalisation not counted)"
    out["cpu_baseline"] = None
    return out


def wl_spmv_symmetric(args):
    """FP64 SpMV of the 3D Poisson matrix 256^3 stored as a SymmetricMatrix
    (lower triangle and diagonal: 66.9 M of the 117 M entries) in CSR and
    Ellpack, beside the full matrix in the same formats."""
    import numpy as np
    import torch
    import tnl_amd

    off, cols, vals = tnl_amd.poisson3d(256, 256, 256)
    n = len(off) - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(off))
    keep = cols <= rows
    loff = np.zeros(n + 1, np.int32)
    loff[1:] = np.cumsum(np.bincount(rows[keep], minlength=n))
    lcols, lvals = cols[keep], vals[keep]
    del rows
    xd = torch.from_numpy(tnl_amd.uniform(n, -1.0, 1.0, seed=0)).cuda()
    yd = torch.zeros(n, dtype=torch.float64, device="cuda")
    steps = max(10, args.steps // 4)
    res = {}
    for name, fmt, sym, (o, c, v) in (("CSR full", tnl_amd.CSR, False, (off, cols, vals)),
                                      ("Ellpack full", tnl_amd.ELLPACK, False, (off, cols, vals)),
                                      ("CSR symmetric", tnl_amd.CSR, True, (loff, lcols, lvals)),
                                      ("Ellpack symmetric", tnl_amd.ELLPACK, True, (loff, lcols, lvals))):
        A = tnl_amd.SparseMatrix.from_csr(n, n, o, c, v, fmt=fmt, symmetric=sym)
        ms = timed(lambda: A.vectorProduct(xd, yd), steps, 5)
        res[name] = {"ms": round(ms, 6), "stored_entries": int(len(c)),
                     "GBs_of_stored_bytes": round(spmv_bytes(n, len(c)) / (ms * 1e-3) / 1e9, 3)}
        y_now = yd.cpu().numpy()
        if name == "CSR full":
            y_full = y_now
        else:
            res[name]["max_rel_diff_vs_full"] = float(np.abs(y_now - y_full).max() / np.abs(y_full).max())
        log(f"Poisson 256^3 {name}: {ms * 1e3:.1f} us")
        del A
    best = res["Ellpack symmetric"]
    nbytes = spmv_bytes(n, len(lcols))
    gbs = nbytes / (best["ms"] * 1e-3) / 1e9
    out = base_record(args, round(gbs, 3), "GB/s", best["ms"], "f64", "synthetic (3D 7-point Poisson 256^3, U[-1,1] x)",
                      {"workload": "FP64 SpMV, 3D Poisson 256^3 as a SymmetricMatrix (lower triangle + diagonal), "
                                   "Ellpack; atomic adds of the mirrored terms", "rows": n,
                       "stored_entries": int(len(lcols)), "bytes_per_spmv": nbytes, "variants": res})
    out["roofline"] = roofline(gbs)
    out["cpu_baseline"] = None
    ref = _oracle_lib().Reference.load() if not args.no_cpu else None
    if ref is not None:
        # TNL's Host SymmetricMatrix<CSR> product of the same lower triangle, all host cores
        cores = cpu_cores()
        x = xd.cpu().numpy()
        t1, _ = ref.spmv_symmetric_time(loff, lcols, lvals, x, 1, cores)
        reps = int(max(3, min(50, args.cpu_budget / max(t1, 1e-6))))
        t, y_cpu = ref.spmv_symmetric_time(loff, lcols, lvals, x, reps, cores)
        out["cpu_baseline"] = {"value": round(nbytes / t / 1e9, 3), "unit": "GB/s", "cores": cores, "kind": "reference",
                               "sample": f"median of {reps} products of the same lower triangle (TNL Host "
                                         f"SymmetricMatrix<CSR> vectorProduct, OpenMP atomics, {cores} threads)"}
        out["parity_max_rel_err_vs_cpu"] = float(np.abs(y_now - y_cpu).max() / np.abs(y_cpu).max())
        log(f"CPU baseline: {out['cpu_baseline']['value']} GB/s ({t * 1e3:.1f} ms/SpMV, {cores} threads), "
            f"GPU vs CPU max rel err {out['parity_max_rel_err_vs_cpu']:.2e}")
    return out


WORKLOADS = {"spmv-powerlaw": wl_spmv_powerlaw, "spmv-laplace": wl_spmv_laplace, "spmv-formats": wl_spmv_formats,
             "spmv-symmetric": wl_spmv_symmetric, "cg-poisson": wl_cg_poisson, "batched-cg": wl_batched_cg,
             "gmres-convdiff": wl_gmres_convdiff}


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--workload", choices=sorted(WORKLOADS), default="spmv-powerlaw")
    p.add_argument("--rows", type=int, default=N_ROWS)
    p.add_argument("--side", type=int, default=None, help="grid side (cg-poisson 256, gmres-convdiff 1024)")
    p.add_argument("--delta", type=float, default=0.1, help="cg-poisson: reliable-update threshold")
    p.add_argument("--systems", type=int, default=4096, help="batched-cg system count")
    p.add_argument("--loop-sample", type=int, default=4096, help="batched-cg: systems timed in the per-system loop")
    p.add_argument("--max-iterations", type=int, default=100_000, help="gmres-convdiff: GMRES iteration cap")
    p.add_argument("--ilu-max-iterations", type=int, default=20_000, help="gmres-convdiff: cap for ILU(0)")
    p.add_argument("--solve-seed", type=int, default=7)
    p.add_argument("--gmres-variant", default="CGSR", choices=["CGS", "CGSR", "MGS", "MGSR"])
    p.add_argument("--solvers", default="jacobi,learned_spai,ilu0,bicgstab",
                   help="gmres-convdiff: the solves to run (value needs jacobi and learned_spai)")
    p.add_argument("--spai-instances", type=int, default=2000)
    p.add_argument("--spai-side", type=int, default=256)
    p.add_argument("--spai-probes", type=int, default=4)
    p.add_argument("--spai-lr", type=float, default=1e-3)
    p.add_argument("--spai-seed", type=int, default=0)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--no-solve", action="store_true",
                   help="spmv-powerlaw: skip the time-to-1e-10 sub-record (Poisson 256^3 CG, mixed vs FP64)")
    p.add_argument("--cpu-budget", type=float, default=15.0)
    p.add_argument("--copy-calibration", action="store_true",
                   help="after the timed region, copy 512 MiB three times (the known-size kernel that scales the "
                        "FETCH_SIZE / WRITE_SIZE counters of a profiling run)")
    args = p.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this build runs on one GPU (the value is the one "
              "GPU's throughput)", file=sys.stderr)
        sys.exit(2)
    if args.side is None:
        args.side = 1024 if args.workload == "gmres-convdiff" else 256
    import torch
    torch.cuda.set_device(0)
    out = WORKLOADS[args.workload](args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
