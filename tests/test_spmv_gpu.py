"""SpMV parity on the GPU, through the C ABI, against the pinned oracle.

* every known-answer case of the reference's SparseMatrixVectorProductTest.hpp,
  for every format, organisation and launch configuration: exact;
* the 2D Laplacian 512x512 (x ~ U[-1,1], seed 0): FP64 relative error <= 1e-12;
* random matrices with padding slots, alpha/beta and row ranges: <= 1e-12;
* FP32 and BF16-stored values: against the oracle in FP32/FP64 with the
  tolerance stated per case;
* the full power-law matrix (4,194,304 rows, 67M nonzeros): FP64 <= 1e-12
  against the oracle's OpenMP product in every format and through the CSR
  launch configurations of the reference (scalar, vector, light, block,
  dynamic grouping); FP32 values and vectors in every format within the
  per-row FP32 rounding bound;
* error behaviour of the reference (size mismatch, invalid threads per segment).
"""
import numpy as np
import pytest

import oracle_lib as OL

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
tnl = pytest.importorskip("tnl_amd")

FORMATS = [(tnl.CSR, 0), (tnl.ADAPTIVE_CSR, 0), (tnl.ELLPACK, 0), (tnl.SLICED_ELLPACK, 32),
           (tnl.SLICED_ELLPACK, 64), (tnl.BIELLPACK, 32), (tnl.CHUNKED_ELLPACK, 0), (tnl.SORTED_CSR, 0)]
FMT_IDS = [f"{tnl.FORMAT_NAMES[f]}{p or ''}" for f, p in FORMATS]
CSR_LAUNCHES = ([(tnl.AUTO, 1), (tnl.WARP, 1), (tnl.BLOCK, 1), (tnl.BLOCK_MERGED, 1)]
                + [(tnl.FIXED, t) for t in (1, 2, 4, 8, 16, 32, 64, 128, 256)]
                + [(tnl.DYNAMIC_GROUPING, t) for t in (1, 4, 32, 256)])


def gpu(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def run(A, x, y0=None, dtype=torch.float64, **kw):
    xd = gpu(x, dtype)
    yd = gpu(y0, dtype) if y0 is not None else torch.zeros(A.rows, dtype=dtype, device="cuda")
    A.vectorProduct(xd, yd, **kw)
    torch.cuda.synchronize()
    return yd.double().cpu().numpy()


def relerr(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)) if len(b) else 0.0


@pytest.mark.parametrize("fmt,param", FORMATS, ids=FMT_IDS)
@pytest.mark.parametrize("org", [tnl.COLUMN_MAJOR, tnl.ROW_MAJOR], ids=["cm", "rm"])
def test_kats_all_formats(fmt, param, org):
    for case in OL.load_kats():
        off, cols, vals, x, y, ncols = OL.kat_csr(case)
        A = tnl.SparseMatrix.from_csr(len(off) - 1, ncols, off, cols, vals, fmt=fmt, param=param, organization=org)
        launches = CSR_LAUNCHES if fmt == tnl.CSR else [None]
        for launch in launches:
            got = run(A, x, np.full(len(y), -1.0), launch=launch)
            assert np.array_equal(got, y), (case["name"], launch)
        if case["rows"] <= 8:  # small integers: exact in FP32 and in BF16 storage
            A32 = tnl.SparseMatrix.from_csr(len(off) - 1, ncols, off, cols, vals, fmt=fmt, param=param,
                                            organization=org, value_dtype=torch.float32)
            assert np.array_equal(run(A32, x, dtype=torch.float32), y), case["name"]
            Ab = tnl.SparseMatrix.from_csr(len(off) - 1, ncols, off, cols, vals, fmt=fmt, param=param,
                                           organization=org, value_dtype=torch.bfloat16)
            assert np.array_equal(run(Ab, x, dtype=torch.float32), y), case["name"]


@pytest.mark.parametrize("fmt,param", FORMATS, ids=FMT_IDS)
def test_laplacian_512(orc, fmt, param):
    off, cols, vals = tnl.laplace2d(512, 512)
    n = len(off) - 1
    x = tnl.uniform(n, -1.0, 1.0, seed=0)
    y_ref = orc.spmv(off, cols, vals, x)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=fmt, param=param)
    assert relerr(run(A, x), y_ref) <= 1e-12
    if fmt == tnl.CSR:
        for launch in CSR_LAUNCHES:
            assert relerr(run(A, x, launch=launch), y_ref) <= 1e-12, launch


def random_csr(n, m, maxlen, rng, pad=0.2, long_rows=()):
    sizes = rng.integers(0, maxlen + 1, n).astype(np.int32)
    for r, L in long_rows:
        sizes[r] = L
    off = np.zeros(n + 1, np.int32)
    off[1:] = np.cumsum(sizes)
    cols = np.full(off[-1], -1, np.int32)
    vals = np.zeros(off[-1])
    for r in range(n):
        used = int(sizes[r]) - int(rng.random() < pad and sizes[r] > 0)
        c = np.sort(rng.choice(m, size=min(used, m), replace=False))
        cols[off[r]:off[r] + len(c)] = c
        vals[off[r]:off[r] + len(c)] = rng.standard_normal(len(c))
    return off, cols, vals


@pytest.mark.parametrize("fmt,param", FORMATS, ids=FMT_IDS)
@pytest.mark.parametrize("org", [tnl.COLUMN_MAJOR, tnl.ROW_MAJOR], ids=["cm", "rm"])
def test_random_alpha_beta_range(orc, fmt, param, org):
    rng = np.random.default_rng(fmt * 10 + org)
    for n, m, maxlen, longs in [(1, 1, 1, ()), (97, 130, 20, ()), (3001, 2500, 40, [(5, 1500), (2000, 2400)])]:
        off, cols, vals = random_csr(n, m, maxlen, rng, long_rows=longs)
        A = tnl.SparseMatrix.from_csr(n, m, off, cols, vals, fmt=fmt, param=param, organization=org)
        x = rng.standard_normal(m)
        y0 = rng.standard_normal(n)
        for alpha, beta, b, e in [(1, 0, 0, 0), (-2.0, 0, 0, 0), (1, 0.5, 0, 0), (0.75, -1.25, n // 3, n - n // 5)]:
            want = orc.spmv(off, cols, vals, x, y=y0, alpha=alpha, beta=beta, begin=b, end=e)
            got = run(A, x, y0, alpha=alpha, beta=beta, begin=b, end=e)
            assert relerr(got, want) <= 1e-12, (n, alpha, beta, b, e)
            if b or e:  # rows outside [begin, end) are untouched
                assert np.array_equal(got[:b], y0[:b]) and np.array_equal(got[e:], y0[e:])


def bf16_round(v):
    """float64 -> float32 (nearest even) -> bfloat16 (nearest even), as float64."""
    u = np.ascontiguousarray(v, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("fmt,param", FORMATS, ids=FMT_IDS)
def test_fp32_and_bf16(orc, fmt, param):
    rng = np.random.default_rng(99)
    n = 5000
    off, cols, vals = random_csr(n, n, 30, rng, long_rows=[(17, 3000)])
    x = rng.standard_normal(n)
    lens = np.diff(off).astype(np.float64)
    absrow = orc.spmv(off, cols, np.abs(vals), np.abs(x))
    # FP32 storage, vectors and accumulation, any summation order:
    # |error| <= (len + 3) * 2^-24 * sum_j |a_ij x_j| (products, sums, rounding of x)
    tol32 = (lens + 3) * 2.0**-24 * absrow * 1.01 + 1e-30
    y = orc.spmv(off, cols, vals, x)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=fmt, param=param, value_dtype=torch.float32)
    got = run(A, x.astype(np.float32), dtype=torch.float32)
    assert np.all(np.abs(got - y) <= tol32 + np.abs(y) * 2.0**-24)
    # BF16 storage (round to nearest even), FP32 vectors and accumulation
    vb = bf16_round(vals)
    x32 = x.astype(np.float32).astype(np.float64)
    yb = orc.spmv(off, cols, vb, x32)
    Ab = tnl.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=fmt, param=param, value_dtype=torch.bfloat16)
    got = run(Ab, x.astype(np.float32), dtype=torch.float32)
    assert np.all(np.abs(got - yb) <= tol32 + np.abs(yb) * 2.0**-24)
    # BF16 storage with FP64 vectors and accumulation: FP64 bar
    got = run(Ab, x, dtype=torch.float64)
    assert relerr(got, orc.spmv(off, cols, vb, x)) <= 1e-12


@pytest.fixture(scope="module")
def powerlaw():
    off, cols, vals = tnl.zipf_csr(4_194_304, seed=0)
    n = len(off) - 1
    x = tnl.uniform(n, -1.0, 1.0, seed=0)
    return n, off, cols, vals, x


def test_powerlaw_full_size(orc, powerlaw):
    # the benchmark matrix at full size: 4,194,304 rows, mean 16 nonzeros per row
    n, off, cols, vals, x = powerlaw
    assert int(off[-1]) == 16 * n
    sizes = np.diff(off)
    assert sizes.min() >= 1 and sizes.max() <= 4096
    want = orc.spmv_omp(off, cols, vals, x, 16)
    for fmt, param in [(tnl.CSR, 0), (tnl.SLICED_ELLPACK, 64), (tnl.BIELLPACK, 32), (tnl.CHUNKED_ELLPACK, 0),
                       (tnl.ADAPTIVE_CSR, 0), (tnl.SORTED_CSR, 0)]:
        A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=fmt, param=param)
        assert relerr(run(A, x), want) <= 1e-12, tnl.FORMAT_NAMES[fmt]
        del A
        torch.cuda.empty_cache()
    # Ellpack is infeasible here: width 4096 x 4,194,304 rows overflows int32 (Ellpack.hpp:91-95)
    with pytest.raises(OverflowError):
        tnl.Segments(tnl.ELLPACK, sizes)


# the CSR launch configurations of SparseMatrixVectorProductTest.hpp at the
# benchmark matrix's size: scalar (Fixed 1), VariableVector (Fixed 2..32),
# LightMultivector (Fixed 64..256), Vector (Warp), Block, BlockMerged and
# DynamicGrouping (ReducingOperations_CSR.h:87-332)
POWERLAW_LAUNCHES = [(tnl.FIXED, 1), (tnl.FIXED, 4), (tnl.FIXED, 16), (tnl.FIXED, 64), (tnl.FIXED, 256),
                     (tnl.WARP, 1), (tnl.BLOCK, 1), (tnl.BLOCK_MERGED, 1), (tnl.DYNAMIC_GROUPING, 1),
                     (tnl.DYNAMIC_GROUPING, 32), (tnl.AUTO, 1)]


def test_powerlaw_full_size_csr_launches(orc, powerlaw):
    n, off, cols, vals, x = powerlaw
    want = orc.spmv_omp(off, cols, vals, x, 16)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals)
    for launch in POWERLAW_LAUNCHES:
        print("power-law CSR FP64 launch", launch, flush=True)
        assert relerr(run(A, x, launch=launch), want) <= 1e-12, launch


def test_powerlaw_full_size_fp32(orc, powerlaw):
    # FP32 values, vectors and accumulation in any summation order:
    # |y - y_exact| <= (len + 3) 2^-24 sum_j |a_ij x_j| per row (products,
    # sums and the rounding of x), y_exact from the FP64 oracle on the
    # FP32-rounded inputs
    n, off, cols, vals, x = powerlaw
    v32 = vals.astype(np.float32)
    x32 = x.astype(np.float32)
    v64, xr = v32.astype(np.float64), x32.astype(np.float64)
    want = orc.spmv_omp(off, cols, v64, xr, 16)
    absrow = orc.spmv_omp(off, cols, np.abs(v64), np.abs(xr), 16)
    lens = np.diff(off).astype(np.float64)
    tol = (lens + 3) * 2.0**-24 * absrow * 1.01 + np.abs(want) * 2.0**-24 + 1e-30
    for fmt, param in [(tnl.CSR, 0), (tnl.ADAPTIVE_CSR, 0), (tnl.SLICED_ELLPACK, 64), (tnl.BIELLPACK, 32),
                       (tnl.CHUNKED_ELLPACK, 0), (tnl.SORTED_CSR, 0)]:
        print("power-law FP32", tnl.FORMAT_NAMES[fmt], flush=True)
        A = tnl.SparseMatrix.from_csr(n, n, off, cols, v32, fmt=fmt, param=param, value_dtype=torch.float32)
        launches = [None, (tnl.FIXED, 1), (tnl.WARP, 1), (tnl.FIXED, 16), (tnl.DYNAMIC_GROUPING, 1)] \
            if fmt == tnl.CSR else [None]
        for launch in launches:
            got = run(A, x32, dtype=torch.float32, launch=launch)
            bad = np.nonzero(np.abs(got - want) > tol)[0]
            assert len(bad) == 0, (tnl.FORMAT_NAMES[fmt], launch, bad[:5])
        del A
        torch.cuda.empty_cache()


def test_auto_plan_follows_the_matrix(orc):
    # 2^20 rows and columns with 4 columns per row spread over the whole
    # width: the automatic launch runs the windowed plan the library builds
    # and keeps (the explicit plan's result bit for bit, the oracle's within
    # 1e-12); a changed element reaches the next product (the library drops
    # its plan), and so does a direct write announced by modified()
    n = 1 << 20
    rng = np.random.default_rng(5)
    q = n // 4
    cols = (np.arange(4, dtype=np.int64) * q + rng.integers(0, q, (n, 4))).astype(np.int32).reshape(-1)
    off = np.arange(0, 4 * n + 1, 4, dtype=np.int32)
    vals = rng.standard_normal(4 * n)
    x = tnl.uniform(n, -1.0, 1.0, seed=3)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals)
    y_auto = run(A, x)
    P = tnl.SpmvPlan(A, 0)
    y_plan = torch.zeros(n, dtype=torch.float64, device="cuda")
    P.vectorProduct(gpu(x), y_plan)
    torch.cuda.synchronize()
    assert np.array_equal(y_auto, y_plan.cpu().numpy())
    assert relerr(y_auto, orc.spmv(off, cols, vals, x)) <= 1e-12
    assert np.array_equal(run(A, x), y_auto)  # the kept plan again
    vals2 = vals.copy()
    vals2[28] = 100.0
    A.setElement(7, int(cols[28]), 100.0)
    y2 = run(A, x)
    assert relerr(y2, orc.spmv(off, cols, vals2, x)) <= 1e-12
    A.values.mul_(2.0)
    A.modified()
    assert relerr(run(A, x), 2.0 * orc.spmv(off, cols, vals2, x)) <= 1e-12
    # FP32 values and vectors take the plan too; a launch configuration of the
    # reference runs its own kernel (scalar: rows summed left to right)
    A32 = tnl.SparseMatrix.from_csr(n, n, off, cols, vals2, value_dtype=torch.float32)
    y32 = run(A32, x.astype(np.float32), dtype=torch.float32)
    want32 = orc.spmv(off, cols, vals2.astype(np.float32).astype(np.float64), x.astype(np.float32))
    assert relerr(y32, want32) <= 1e-6
    assert relerr(run(A, x, launch=(tnl.FIXED, 1)), 2.0 * orc.spmv(off, cols, vals2, x)) <= 1e-12


def test_errors():
    off, cols, vals = tnl.laplace2d(8, 8)
    A = tnl.SparseMatrix.from_csr(64, 64, off, cols, vals)
    x = torch.zeros(64, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        A.vectorProduct(torch.zeros(63, dtype=torch.float64, device="cuda"), x)
    with pytest.raises(ValueError):
        A.vectorProduct(x, torch.zeros(65, dtype=torch.float64, device="cuda"))
    with pytest.raises(ValueError):
        A.vectorProduct(x, x.clone(), launch=(tnl.FIXED, 3))
    with pytest.raises(tnl.TnlmiError):
        A.vectorProduct(x, x.clone(), launch=(tnl.FIXED, 512))
    with pytest.raises(ValueError):
        A.vectorProduct(x, x.clone(), begin=10, end=5)


def test_csr_short_rows_one_launch(orc):
    # no row longer than 32 elements: the tiled kernel's owner finishes its last
    # row past the tile end (no carries, no fixup launch) and one lane sums each
    # row left to right from zero: bit-exact with TNL's host loop.  Matrices of
    # at most 2^22 elements get 256-element tiles, larger ones 512.
    for side in (1100, 512):  # 6,046,000 elements (512-element tiles), then 1,308,672 (256)
        off, cols, vals = tnl.laplace2d(side, side)
        n = len(off) - 1
        x = tnl.uniform(n, -1.0, 1.0, seed=0)
        want = orc.spmv(off, cols, vals, x)
        for fmt in (tnl.CSR, tnl.SORTED_CSR):
            A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=fmt)
            assert np.array_equal(run(A, x), want), (side, fmt)
    ref = OL.Reference.load()
    if ref is not None:
        A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals)
        assert np.array_equal(run(A, x), ref.spmv("CSR", off, cols, vals, x))
    rng = np.random.default_rng(31)
    # longest row 32 (one launch) and 33 (carries): empty rows, padding slots,
    # rows of exactly 32 (8 rows per tile), row ranges
    for n, m, maxlen, full in [(5000, 7000, 32, False), (4000, 9000, 32, True), (5000, 7000, 33, False),
                               (700, 40, 12, False)]:
        off, cols, vals = random_csr(n, m, maxlen, rng)
        if full:
            off, cols, vals = random_csr(n, m, maxlen, rng, pad=0.0, long_rows=[(r, 32) for r in range(n)])
        for fmt in (tnl.CSR, tnl.SORTED_CSR):
            A = tnl.SparseMatrix.from_csr(n, m, off, cols, vals, fmt=fmt)
            xs = rng.standard_normal(m)
            y0 = rng.standard_normal(n)
            for alpha, beta, b, e in [(1, 0, 0, 0), (-2.0, 0.5, 0, 0), (0.75, -1.25, n // 3, n - n // 5)]:
                want = orc.spmv(off, cols, vals, xs, y=y0, alpha=alpha, beta=beta, begin=b, end=e)
                got = run(A, xs, y0, alpha=alpha, beta=beta, begin=b, end=e)
                assert relerr(got, want) <= 1e-13, (n, maxlen, fmt, alpha, beta, b, e)
                if maxlen <= 32 and (alpha, beta) == (1, 0):
                    assert np.array_equal(got, want), (n, maxlen, fmt)
                if b or e:
                    assert np.array_equal(got[:b], y0[:b]) and np.array_equal(got[e:], y0[e:])
        A = tnl.SparseMatrix.from_csr(n, m, off, cols, vals, value_dtype=torch.float32)
        xs = rng.standard_normal(m).astype(np.float32)
        want = orc.spmv(off, cols, vals.astype(np.float32).astype(np.float64), xs.astype(np.float64))
        got = run(A, xs, dtype=torch.float32)
        assert relerr(got, want) <= 1e-5, (n, maxlen)


@pytest.mark.parametrize("parts", [8, 16, 0], ids=["split8", "split16", "windowed"])
def test_column_split_plan(orc, parts):
    rng = np.random.default_rng(parts)
    for n, m, maxlen, longs in [(1, 1, 1, ()), (300, 700, 30, ()), (20000, 30000, 40, [(3, 5000), (19999, 2500)])]:
        off, cols, vals = random_csr(n, m, maxlen, rng, long_rows=longs)
        A = tnl.SparseMatrix.from_csr(n, m, off, cols, vals)
        P = tnl.SpmvPlan(A, parts)
        x = rng.standard_normal(m)
        y0 = rng.standard_normal(n)
        for alpha, beta, b, e in [(1, 0, 0, 0), (-2.0, 0.5, 0, 0), (0.75, -1.25, n // 3, n - n // 5)]:
            want = orc.spmv(off, cols, vals, x, y=y0, alpha=alpha, beta=beta, begin=b, end=e)
            xd, yd = gpu(x), gpu(y0)
            P.vectorProduct(xd, yd, alpha=alpha, beta=beta, begin=b, end=e)
            torch.cuda.synchronize()
            got = yd.cpu().numpy()
            assert relerr(got, want) <= 1e-12, (n, alpha, beta, b, e)
            if b or e:
                assert np.array_equal(got[:b], y0[:b]) and np.array_equal(got[e:], y0[e:])
        del A  # the plan owns a copy of the matrix data
        xd, yd = gpu(x), torch.zeros(n, dtype=torch.float64, device="cuda")
        P.vectorProduct(xd, yd)
        torch.cuda.synchronize()
        assert relerr(yd.cpu().numpy(), orc.spmv(off, cols, vals, x)) <= 1e-12
    # full-size power-law matrix
    off, cols, vals = tnl.zipf_csr(4_194_304, seed=0)
    n = len(off) - 1
    x = tnl.uniform(n, -1.0, 1.0, seed=0)
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals)
    P = tnl.SpmvPlan(A, parts)
    del A
    xd, yd = gpu(x), torch.zeros(n, dtype=torch.float64, device="cuda")
    P.vectorProduct(xd, yd)
    torch.cuda.synchronize()
    assert relerr(yd.cpu().numpy(), orc.spmv_omp(off, cols, vals, x, 16)) <= 1e-12


def test_windowed_plan_matches_tnl_bit_for_bit(orc):
    # rows summed by one thread: products rounded to FP64, then summed from
    # zero left to right in column order -- TNL's host loop exactly
    # (rows of at most 256 nonzeros); longer rows are summed by a wave in a
    # fixed order and agree to rounding
    rng = np.random.default_rng(5)
    n, m = 60000, 300001  # 37 column windows, the last one partial
    off, cols, vals = random_csr(n, m, 24, rng, long_rows=[(11, 256), (12, 257), (4000, 3000), (59999, 700)])
    A = tnl.SparseMatrix.from_csr(n, m, off, cols, vals)
    P = tnl.SpmvPlan(A, 0)
    x = rng.standard_normal(m)
    xd, yd = gpu(x), torch.zeros(n, dtype=torch.float64, device="cuda")
    P.vectorProduct(xd, yd)
    torch.cuda.synchronize()
    got, want = yd.cpu().numpy(), orc.spmv(off, cols, vals, x)
    short = np.diff(off) <= 256
    assert short[11] and not short[12]
    assert np.array_equal(got[short], want[short])
    assert relerr(got[~short], want[~short]) <= 1e-13
    ref = OL.Reference.load()  # TNL's own Host CSR vectorProduct, where built
    if ref is not None:
        assert np.array_equal(got[short], ref.spmv("CSR", off, cols, vals, x)[short])


@pytest.mark.parametrize("slots,rounds", [("block", "across"), ("window", "grouped"), ("block", "grouped")])
def test_windowed_plan_more_windows_than_a_sweep(orc, monkeypatch, slots, rounds):
    # 611 column windows: phase 2 reads windows past its 512-window sweep on
    # demand; most chunks are empty or hold one or two products.  Every slot
    # layout and read order gives TNL's host loop exactly (rows <= 256)
    monkeypatch.setenv("TNLMI_WINDOW_SLOTS", slots)
    monkeypatch.setenv("TNLMI_WINDOW_ROUNDS", rounds)
    rng = np.random.default_rng(8)
    n, m = 20000, 5_000_000
    off, cols, vals = random_csr(n, m, 40, rng, long_rows=[(3, 2000), (19999, 300)])
    A = tnl.SparseMatrix.from_csr(n, m, off, cols, vals)
    P = tnl.SpmvPlan(A, 0)
    x = rng.standard_normal(m)
    yd = torch.zeros(n, dtype=torch.float64, device="cuda")
    P.vectorProduct(gpu(x), yd)
    torch.cuda.synchronize()
    got, want = yd.cpu().numpy(), orc.spmv(off, cols, vals, x)
    short = np.diff(off) <= 256
    assert np.array_equal(got[short], want[short])
    assert relerr(got[~short], want[~short]) <= 1e-13


def test_windowed_plan_fp32_and_errors(orc):
    rng = np.random.default_rng(6)
    n, m = 5000, 20000
    off, cols, vals = random_csr(n, m, 30, rng, long_rows=[(7, 9000)])
    A = tnl.SparseMatrix.from_csr(n, m, off, cols, vals, value_dtype=torch.float32)
    P = tnl.SpmvPlan(A, 0)
    x = rng.standard_normal(m).astype(np.float32)
    xd, yd = gpu(x).float(), torch.zeros(n, dtype=torch.float32, device="cuda")
    P.vectorProduct(xd, yd)
    torch.cuda.synchronize()
    v32 = vals.astype(np.float32).astype(np.float64)
    want = orc.spmv(off, cols, v32, x.astype(np.float64))
    bound = 1e-6 * np.bincount(np.repeat(np.arange(n), np.diff(off)), weights=np.abs(v32 * np.where(
        cols >= 0, x.astype(np.float64)[np.maximum(cols, 0)], 0.0)), minlength=n) * 64
    assert np.all(np.abs(yd.cpu().numpy() - want) <= bound + 1e-30)
    with pytest.raises(NotImplementedError):  # FP32 plan, FP64 vectors
        P.vectorProduct(gpu(x.astype(np.float64)), torch.zeros(n, dtype=torch.float64, device="cuda"))
    with pytest.raises(tnl.TnlmiError):  # parts must be 0, 8 or 16
        tnl.SpmvPlan(A, 4)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
def test_windowed_plan_rows_longer_than_a_block(orc, dtype):
    # a row block holds 10000 FP64 / 20000 FP32 nonzeros: longer rows are cut
    # into pieces whose partial sums are added in order; empty rows around them
    rng = np.random.default_rng(7)
    m = 70000
    longs = [(0, 12000), (5, 10000), (6, 10001), (9, 45000), (40, 20001), (77, 25000)]
    off, cols, vals = random_csr(80, m, 12, rng, long_rows=longs)
    sizes = np.diff(off).copy()
    sizes[[1, 2, 78, 79]] = 0  # empty rows after the first long row and after the last one
    off, cols, vals = _resize_rows(off, cols, vals, sizes)
    npd = np.float64 if dtype == torch.float64 else np.float32
    v = vals.astype(npd).astype(np.float64)
    A = tnl.SparseMatrix.from_csr(80, m, off, cols, vals, value_dtype=dtype)
    P = tnl.SpmvPlan(A, 0)
    x = rng.standard_normal(m).astype(npd)
    y0 = rng.standard_normal(80).astype(npd)
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    for alpha, beta, b, e in [(1, 0, 0, 0), (-2.0, 0.5, 0, 0), (0.75, -1.25, 6, 78)]:
        want = orc.spmv(off, cols, v, x.astype(np.float64), y=y0.astype(np.float64), alpha=alpha, beta=beta,
                        begin=b, end=e)
        xd, yd = gpu(x).to(dtype), gpu(y0).to(dtype)
        P.vectorProduct(xd, yd, alpha=alpha, beta=beta, begin=b, end=e)
        torch.cuda.synchronize()
        got = yd.cpu().numpy().astype(np.float64)
        assert relerr(got, want) <= tol, (alpha, beta, b, e)
        if b or e:
            assert np.array_equal(got[:b], y0[:b]) and np.array_equal(got[e:], y0[e:])
    # all rows empty
    Z = tnl.SparseMatrix.from_csr(5, 9, np.zeros(6, np.int32), np.zeros(0, np.int32), np.zeros(0))
    yd = gpu(np.ones(5))
    tnl.SpmvPlan(Z, 0).vectorProduct(gpu(np.ones(9)), yd, alpha=1.0, beta=0.5)
    torch.cuda.synchronize()
    assert np.array_equal(yd.cpu().numpy(), np.full(5, 0.5))


def _resize_rows(off, cols, vals, sizes):
    """The CSR arrays with row r cut to its first sizes[r] entries."""
    keep = np.concatenate([np.arange(off[r], off[r] + sizes[r]) for r in range(len(sizes))]).astype(np.int64)
    new = np.zeros(len(sizes) + 1, np.int32)
    new[1:] = np.cumsum(sizes)
    return new, cols[keep], vals[keep]
