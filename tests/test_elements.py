"""SparseMatrix::setElement / addElement / getElement (SparseMatrixBase.hpp:152-265).

CPU: the oracle's restatement against a hand-derived sequence (in-place update,
insertion into padding, sorted insertion with the shift, the reference's
full-row shift that drops the last slot, and the capacity error) and against
TNL's own Host SparseMatrix (general and symmetric) on random operation lists.
"""
import numpy as np
import pytest

import oracle_lib as OL


@pytest.fixture(scope="module")
def orc():
    return OL.Oracle()


def random_ops(rng, n, m, count):
    mults = [0.0, 1.0, 0.5, -2.0, 3.0]
    return [(int(rng.integers(n)), int(rng.integers(m)), float(rng.standard_normal()),
             float(mults[rng.integers(len(mults))])) for _ in range(count)]


def test_hand_derived_sequence(orc):
    off = [0, 3, 3]  # row 0: three slots, row 1: none
    ops = [(0, 5, 1.0, 0.0),    # padding slot 0
           (0, 2, 2.0, 0.0),    # before 5: [2, 5, -1]
           (0, 9, 3.0, 0.0),    # padding slot 2: [2, 5, 9]
           (0, 2, 10.0, 0.5),   # update: 0.5 * 2 + 10 = 11
           (0, 1, 4.0, 0.0),    # full row, 1 < 2: shift right, 9 falls off: [1, 2, 5]
           (0, 20, 7.0, 0.0)]   # full row, no larger column: capacity error
    cols, vals, failed = orc.add_elements(off, ops)
    assert failed == 5
    assert cols.tolist() == [1, 2, 5] and vals.tolist() == [4.0, 11.0, 1.0]
    cols, vals, failed = orc.add_elements(off, [(1, 0, 1.0, 0.0)])
    assert failed == 0 and cols.tolist() == [-1, -1, -1]
    # symmetric: (1, 3) is stored as (3, 1)
    cols, vals, failed = orc.add_elements([0, 1, 2, 3, 5], [(1, 3, 2.5, 0.0), (3, 3, 1.0, 0.0)], symmetric=True)
    assert failed == -1 and cols.tolist() == [-1, -1, -1, 1, 3] and vals.tolist() == [0, 0, 0, 2.5, 1.0]


@pytest.mark.parametrize("symmetric", [False, True], ids=["general", "symmetric"])
def test_oracle_vs_tnl(orc, symmetric):
    ref = OL.Reference.load()
    if ref is None:
        pytest.skip("TNL's Host build (oracle/_ref) is not built")
    rng = np.random.default_rng(3 + symmetric)
    n = m = 40
    for trial in range(6):
        caps = rng.integers(0, 7, n)
        off = np.zeros(n + 1, np.int32)
        off[1:] = np.cumsum(caps)
        ops = random_ops(rng, n, m, 120)
        cols, vals, failed = orc.add_elements(off, ops, symmetric=symmetric)
        rc, rv, rfailed, msg = ref.add_elements("CSR", off, m, ops, symmetric=symmetric)
        assert failed == rfailed, trial
        assert np.array_equal(cols, rc) and np.array_equal(vals, rv), trial
        if failed >= 0:
            r, c = ops[failed][:2]
            row = max(r, c) if symmetric else r
            assert msg == f"The capacity of the sparse matrix row number {row} was exceeded."
