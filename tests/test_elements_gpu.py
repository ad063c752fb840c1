"""setElement / addElement / getElement on the GPU through the C ABI, every
format and organisation, general and symmetric: the storage after each
operation list equals the oracle's (bit-exact, mapped through the format's
global indexes) and TNL's own Host build's storage arrays; the capacity error
comes at the same operation with the reference's message."""
import numpy as np
import pytest

import oracle_lib as OL

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
tnl = pytest.importorskip("tnl_amd")
from tnl_amd import solvers  # noqa: E402

FORMATS = [(tnl.CSR, 0, "CSR"), (tnl.ADAPTIVE_CSR, 0, "AdaptiveCSR"), (tnl.ELLPACK, 0, "Ellpack"),
           (tnl.SLICED_ELLPACK, 32, "SlicedEllpack32"), (tnl.SLICED_ELLPACK, 64, "SlicedEllpack64"),
           (tnl.BIELLPACK, 32, "BiEllpack"), (tnl.CHUNKED_ELLPACK, 0, "ChunkedEllpack"), (tnl.SORTED_CSR, 0, "SortedCSR")]


@pytest.fixture(scope="module")
def orc():
    return OL.Oracle()


def apply(A, ops, failed):
    """Applies ops up to the failing one, which must raise the reference's error."""
    for i, (r, c, v, mult) in enumerate(ops):
        if i == failed:
            with pytest.raises(tnl.LogicError, match="The capacity of the sparse matrix row number"):
                A.addElement(r, c, v, mult)
            return
        if mult == 0.0:
            A.setElement(r, c, v)
        else:
            A.addElement(r, c, v, mult)


CSR_FAMILY = (tnl.CSR, tnl.ADAPTIVE_CSR, tnl.SORTED_CSR)


@pytest.mark.parametrize("fmt,param,name", FORMATS, ids=[f[2] for f in FORMATS])
@pytest.mark.parametrize("org", [tnl.COLUMN_MAJOR, tnl.ROW_MAJOR], ids=["cm", "rm"])
@pytest.mark.parametrize("symmetric", [False, True], ids=["general", "symmetric"])
def test_elements_match_oracle_and_tnl(orc, fmt, param, name, org, symmetric):
    # a row has getSegmentSize slots: its capacity in the CSR formats (checked
    # against the oracle in CSR order), the padded width in the Ellpack
    # formats (checked against TNL's own Host build, storage arrays bit-exact)
    if org == tnl.ROW_MAJOR and fmt in CSR_FAMILY:
        pytest.skip("CSR formats have one organisation")
    ref = OL.Reference.load()
    if ref is None and fmt not in CSR_FAMILY:
        pytest.skip("the padded formats are checked against TNL's Host build (oracle/_ref), not built")
    rng = np.random.default_rng(fmt * 4 + org * 2 + symmetric)
    n = m = 70
    mults = [0.0, 1.0, 0.5, -2.0]
    for trial in range(2):
        caps = rng.integers(0, 6, n)
        off = np.zeros(n + 1, np.int32)
        off[1:] = np.cumsum(caps)
        ops = [(int(rng.integers(n)), int(rng.integers(m)), float(rng.standard_normal()),
                mults[rng.integers(len(mults))]) for _ in range(150)]
        want_c, want_v, failed = orc.add_elements(off, ops, symmetric=symmetric)
        if ref is not None:
            rc, rv, rfailed, _ = ref.add_elements("CSR" if fmt == tnl.ADAPTIVE_CSR else name, off, m, ops,
                                                  symmetric=symmetric, col_major=org == tnl.COLUMN_MAJOR)
            if fmt in CSR_FAMILY:
                assert rfailed == failed
            failed = rfailed
        A = tnl.SparseMatrix.from_csr(n, m, off, np.full(off[-1], -1, np.int32), np.zeros(off[-1]), fmt=fmt,
                                      param=param, organization=org, symmetric=symmetric)
        apply(A, ops, failed)
        if ref is not None:
            assert np.array_equal(A.column_indexes.cpu().numpy()[: len(rc)], rc), (name, trial)
            assert np.array_equal(A.values.cpu().numpy()[: len(rv)], rv), (name, trial)
        if fmt in CSR_FAMILY:
            g = A.segments.global_indexes().long()
            got_c = A.column_indexes[g].cpu().numpy()
            got_v = A.values[g].cpu().numpy()
            assert np.array_equal(got_c, want_c) and np.array_equal(got_v, want_v), (name, trial, failed)
            # getElement: stored elements, and absent positions read 0
            for r in range(0, n, 7):
                for c in range(0, m, 5):
                    rr, cc = (max(r, c), min(r, c)) if symmetric else (r, c)
                    seg = slice(off[rr], off[rr + 1])
                    hit = np.nonzero(want_c[seg] == cc)[0]
                    expect = float(want_v[seg][hit[0]]) if len(hit) else 0.0
                    assert A.getElement(r, c) == expect, (name, r, c)


def test_elements_ellpack_rows_have_the_full_width():
    # Ellpack: every row has width = the largest capacity (EllpackBase.hpp), so a
    # row of capacity 1 takes three elements; SlicedEllpack: its slice's width
    for fmt, param in ((tnl.ELLPACK, 0), (tnl.SLICED_ELLPACK, 32)):
        off = np.array([0, 1, 4], np.int32)
        A = tnl.SparseMatrix.from_csr(2, 6, off, np.full(4, -1, np.int32), np.zeros(4), fmt=fmt, param=param)
        A.setElement(0, 4, 1.0)  # [4, -, -]
        A.setElement(0, 0, 2.0)  # before 4: [0, 4, -]
        A.setElement(0, 2, 3.0)  # before 4: [0, 2, 4]
        assert [A.getElement(0, c) for c in range(6)] == [2.0, 0.0, 3.0, 0.0, 1.0, 0.0]
        with pytest.raises(tnl.LogicError, match="The capacity of the sparse matrix row number 0 was exceeded."):
            A.setElement(0, 5, 9.0)  # all three slots hold smaller columns


def test_elements_fp32_and_errors(orc):
    off = np.array([0, 2, 4], np.int32)
    A = tnl.SparseMatrix.from_csr(2, 6, off, np.full(4, -1, np.int32), np.zeros(4), value_dtype=torch.float32)
    A.setElement(0, 3, 1.5)
    A.addElement(0, 3, 0.25, 0.5)  # 0.5 * 1.5 + 0.25 = 1.0
    A.setElement(1, 5, -2.0)
    A.addElement(1, 0, 3.0)
    assert A.getElement(0, 3) == 1.0 and A.getElement(1, 0) == 3.0 and A.getElement(1, 5) == -2.0
    assert A.getElement(1, 2) == 0.0
    with pytest.raises(tnl.TnlmiError):
        A.setElement(2, 0, 1.0)
    with pytest.raises(tnl.TnlmiError):
        A.getElement(0, 6)
    A.setElement(0, 1, 7.0)  # row 0 full: [1, 3]
    with pytest.raises(tnl.LogicError, match="row number 0 was exceeded"):
        A.setElement(0, 4, 1.0)
    y = torch.zeros(2, dtype=torch.float32, device="cuda")
    A.vectorProduct(torch.ones(6, dtype=torch.float32, device="cuda"), y)
    assert y.cpu().tolist() == [8.0, 1.0]


@pytest.mark.parametrize("fmt,param,name", FORMATS, ids=[f[2] for f in FORMATS])
@pytest.mark.parametrize("symmetric", [False, True], ids=["general", "symmetric"])
def test_row_counts_match_tnl(fmt, param, name, symmetric):
    # getRowCapacity (the segment size), getRowCapacities (stored slots),
    # getCompressedRowLengths (stored nonzeros), getNonzeroElementsCount; stored
    # zeros and padding slots at row ends
    rng = np.random.default_rng(40 + fmt * 2 + symmetric)
    n = 300
    sizes = rng.integers(0, 9, n).astype(np.int32)
    sizes[[3, 200]] = [40, 70]
    off = np.zeros(n + 1, np.int32)
    off[1:] = np.cumsum(sizes)
    cols = np.full(off[-1], -1, np.int32)
    vals = np.zeros(off[-1])
    for r in range(n):
        hi = r + 1 if symmetric else n
        used = min(int(sizes[r]) - int(rng.random() < 0.3 and sizes[r] > 0), hi)
        c = np.sort(rng.choice(hi, size=used, replace=False))
        cols[off[r]:off[r] + used] = c
        vals[off[r]:off[r] + used] = np.where(rng.random(used) < 0.2, 0.0, rng.standard_normal(used))
    A = tnl.SparseMatrix.from_csr(n, n, off, cols, vals, fmt=fmt, param=param, symmetric=symmetric)
    row_of = np.repeat(np.arange(n), np.diff(off))
    stored = cols >= 0
    want_stored = np.bincount(row_of[stored], minlength=n)
    want_nz = np.bincount(row_of[stored & (vals != 0)], minlength=n)
    assert np.array_equal(A.getRowCapacities().cpu().numpy(), want_stored)
    assert np.array_equal(A.getCompressedRowLengths().cpu().numpy(), want_nz)
    want_total = int(stored.sum() + (stored & (cols != row_of)).sum()) if symmetric else int(want_nz.sum())
    assert A.getNonzeroElementsCount() == want_total
    ref = OL.Reference.load()
    if ref is not None:
        cap, rstored, rnz, total = ref.row_counts("CSR" if fmt == tnl.ADAPTIVE_CSR else name, off, cols, vals, n,
                                                  symmetric=symmetric)
        assert np.array_equal(A._row_counts(0).cpu().numpy(), cap)
        assert np.array_equal(rstored, want_stored) and np.array_equal(rnz, want_nz) and total == want_total


@pytest.mark.parametrize("fmt,param", [(tnl.ELLPACK, 0), (tnl.SLICED_ELLPACK, 32)], ids=["Ellpack", "SlicedEllpack32"])
@pytest.mark.parametrize("org", [tnl.COLUMN_MAJOR, tnl.ROW_MAJOR], ids=["cm", "rm"])
def test_slots_past_the_capacity_are_walked(fmt, param, org):
    # rows 0 and 2 ask for one slot but have the padded width 3
    # (getSegmentSize); the elements set past the capacity take part in every
    # per-row walk of the reference -- the symmetric and transposed products,
    # copySparseMatrix and the Diagonal preconditioner's diagonal (SparseMatrixBase.hpp:
    # 298-329, 541-588; SparseOperations.hpp:544-597 via row.getSize();
    # Diagonal.hpp:27-30) -- checked against the dense closed forms and TNL's
    # own storage after the same operations
    off = np.array([0, 1, 4, 5], np.int32)
    n = 3
    ops_general = [(0, 1, 3.0, 0.0), (0, 2, 4.0, 0.0), (0, 0, 2.0, 0.0), (1, 1, 5.0, 0.0), (2, 0, 1.0, 0.0),
                   (2, 2, 6.0, 0.0)]  # row 2's diagonal lands in slot 1, past its capacity
    ops_lower = [(0, 0, 2.0, 0.0), (1, 0, 3.0, 0.0), (1, 1, 5.0, 0.0), (2, 0, 1.0, 0.0), (2, 1, -2.0, 0.0),
                 (2, 2, 6.0, 0.0)]
    x_h = np.array([0.5, -1.25, 2.0])
    x = torch.tensor(x_h, device="cuda")
    ref = OL.Reference.load()
    name = "Ellpack" if fmt == tnl.ELLPACK else "SlicedEllpack32"
    for symmetric, ops in ((False, ops_general), (True, ops_lower)):
        D = np.zeros((n, n))
        for r, c, v, _ in ops:
            D[r, c] = v
        A = tnl.SparseMatrix.from_csr(n, n, off, np.full(5, -1, np.int32), np.zeros(5), fmt=fmt, param=param,
                                      organization=org, symmetric=symmetric)
        apply(A, ops, -1)
        if ref is not None:
            rc, rv, rfailed, _ = ref.add_elements(name, off, n, ops, symmetric=symmetric,
                                                  col_major=org == tnl.COLUMN_MAJOR)
            assert rfailed == -1
            assert np.array_equal(A.column_indexes.cpu().numpy()[: len(rc)], rc)
            assert np.array_equal(A.values.cpu().numpy()[: len(rv)], rv)
        full = D + D.T - np.diag(np.diag(D)) if symmetric else D
        y = torch.zeros(n, dtype=torch.float64, device="cuda")
        A.vectorProduct(x, y)
        assert np.allclose(y.cpu().numpy(), full @ x_h, rtol=0, atol=1e-15), symmetric
        if not symmetric:
            yt = torch.zeros(n, dtype=torch.float64, device="cuda")
            A.transposedVectorProduct(x, yt)
            assert np.allclose(yt.cpu().numpy(), D.T @ x_h, rtol=0, atol=1e-15)
        B = A.convert(tnl.CSR)
        C = B.convert(fmt, param=param, organization=org)
        for M in (B, C):
            # getElement of a symmetric matrix reads (c, r) above the diagonal (SparseMatrixBase.hpp:238-262)
            got = np.array([[M.getElement(r, c) for c in range(n)] for r in range(n)])
            assert np.array_equal(got, full), symmetric
        P = solvers.Diagonal(A)
        b = torch.tensor([4.0, 10.0, 12.0], dtype=torch.float64, device="cuda")
        z = torch.zeros(n, dtype=torch.float64, device="cuda")
        P.solve(b, z)
        assert np.array_equal(z.cpu().numpy(), b.cpu().numpy() / np.diag(D)), symmetric


@pytest.mark.parametrize("fmt,param,name", FORMATS, ids=[f[2] for f in FORMATS])
@pytest.mark.parametrize("symmetric", [False, True], ids=["general", "symmetric"])
def test_add_elements_batch_matches_one_by_one(fmt, param, name, symmetric):
    """addElements (one launch for the whole list, rows in parallel) leaves the
    storage TNL's addElement calls in list order leave, and equals the one-by-one
    GPU path; mixed setElement (multiplicator 0) and addElement operations."""
    ref = OL.Reference.load()
    rng = np.random.default_rng(fmt * 2 + symmetric + 100)
    n = m = 300
    caps = rng.integers(1, 9, n)
    off = np.zeros(n + 1, np.int32)
    off[1:] = np.cumsum(caps)
    # operations that never overflow a row: each stored row receives at most its capacity of distinct
    # columns (a symmetric matrix stores row >= column; an operation names either (r, c) or (c, r))
    ops = []
    for r in range(n):
        pool = np.arange(r + 1) if symmetric else np.arange(m)
        cols = rng.choice(pool, size=min(int(caps[r]), len(pool)), replace=False)
        for _ in range(int(rng.integers(0, 12))):
            c = int(cols[rng.integers(len(cols))])
            rr, cc = (c, r) if symmetric and rng.random() < 0.5 else (r, c)
            ops.append((rr, cc, float(rng.standard_normal()), [0.0, 1.0, 0.5][rng.integers(3)]))
    order = rng.permutation(len(ops))
    ops = [ops[i] for i in order]
    empty = (np.full(off[-1], -1, np.int32), np.zeros(off[-1]))
    A = tnl.SparseMatrix.from_csr(n, m, off, *empty, fmt=fmt, param=param, symmetric=symmetric)
    B = tnl.SparseMatrix.from_csr(n, m, off, *empty, fmt=fmt, param=param, symmetric=symmetric)
    rows = [o[0] for o in ops]
    cols = [o[1] for o in ops]
    vals = [o[2] for o in ops]
    mls = [o[3] for o in ops]
    A.addElements(rows, cols, vals, mls)
    apply(B, ops, -1)
    assert np.array_equal(A.column_indexes.cpu().numpy(), B.column_indexes.cpu().numpy())
    assert np.array_equal(A.values.cpu().numpy(), B.values.cpu().numpy())
    if ref is not None:
        rc, rv, rfailed, _ = ref.add_elements("CSR" if fmt == tnl.ADAPTIVE_CSR else name, off, m, ops,
                                              symmetric=symmetric)
        assert rfailed == -1
        assert np.array_equal(A.column_indexes.cpu().numpy()[: len(rc)], rc)
        assert np.array_equal(A.values.cpu().numpy()[: len(rv)], rv)
    # setElement batch (multiplicator 0) over the same positions
    A.setElementsBatch(rows, cols, [2.0 * v for v in vals])
    B2 = tnl.SparseMatrix.from_csr(n, m, off, *empty, fmt=fmt, param=param, symmetric=symmetric)
    apply(B2, ops, -1)
    apply(B2, [(r, c, 2.0 * v, 0.0) for (r, c, v, _) in ops], -1)
    assert np.array_equal(A.values.cpu().numpy(), B2.values.cpu().numpy())


def test_add_elements_batch_errors():
    n = 6
    off = np.array([0, 1, 2, 3, 4, 5, 6], np.int32)
    empty = (np.full(6, -1, np.int32), np.zeros(6))
    A = tnl.SparseMatrix.from_csr(n, n, off, *empty)
    # row 2 takes one column, then a larger one: its capacity is exceeded at list index 3;
    # the other rows' operations are applied, row 2 keeps its first element
    rows, cols, vals = [0, 2, 4, 2, 5], [0, 1, 4, 3, 5], [1.0, 2.0, 3.0, 4.0, 5.0]
    with pytest.raises(tnl.LogicError, match="The capacity of the sparse matrix row number 2 was exceeded"):
        A.addElements(rows, cols, vals)
    got = {(r, c): A.getElement(r, c) for r, c in zip(rows, cols)}
    assert got == {(0, 0): 1.0, (2, 1): 2.0, (4, 4): 3.0, (2, 3): 0.0, (5, 5): 5.0}
    # a position outside the matrix: nothing is written
    B = tnl.SparseMatrix.from_csr(n, n, off, *empty)
    with pytest.raises(tnl.InvalidArgument):
        B.addElements([0, 6], [0, 0], [1.0, 1.0])
    assert B.getElement(0, 0) == 0.0
    with pytest.raises(tnl.InvalidArgument):
        B.addElements([0, 1], [0], [1.0, 1.0])
