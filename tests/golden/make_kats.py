"""Writes spmv_kats.json: the known-answer SpMV cases of the reference's
src/UnitTests/Matrices/SparseMatrixVectorProductTest.hpp as data.

Each case: matrix (rows, columns, row capacities, entries (row, col, value)),
the input vector x and the expected output y exactly as the reference test
states them (hand-written constants or the closed forms written in the test).
Run from the repository root: python tests/golden/make_kats.py
"""
import json
import os


def small(name, rows, cols, entries, x, y, caps=None, src=""):
    return dict(name=name, rows=rows, columns=cols, capacities=caps, entries=entries, x=[x] * cols, y=y, source=src)


cases = []
# test_VectorProduct_zeroMatrix (:14-74): capacities [1,2,1,1], no entries, x = 1
cases.append(small("zeroMatrix", 4, 4, [], 1, [0, 0, 0, 0], caps=[1, 2, 1, 1], src=":14-74"))
# test_VectorProduct_smallMatrix1 (:76-137)
cases.append(small("smallMatrix1", 4, 4, [[0, 0, 1], [1, 1, 2], [1, 3, 3], [2, 1, 4], [3, 2, 5]], 2,
                   [2, 10, 8, 10], src=":76-137"))
# test_VectorProduct_smallMatrix2 (:139-199)
cases.append(small("smallMatrix2", 4, 4, [[0, 0, 1], [0, 1, 2], [0, 2, 3], [1, 3, 4], [2, 0, 5], [2, 1, 6],
                                          [2, 2, 7], [3, 1, 8]], 2, [12, 8, 36, 16], src=":139-199"))
# test_VectorProduct_smallMatrix3 (:201-265)
cases.append(small("smallMatrix3", 4, 4, [[0, 0, 1], [0, 1, 2], [0, 2, 3], [1, 1, 4], [1, 2, 5], [1, 3, 6],
                                          [2, 0, 7], [2, 1, 8], [2, 2, 9], [3, 1, 10], [3, 2, 11], [3, 3, 12]], 2,
                   [12, 30, 48, 66], src=":201-265"))
# test_VectorProduct_mediumSizeMatrix1 (:267-343)
m4 = [[0, 0, 1], [0, 1, 2], [0, 2, 3], [0, 5, 4],
      [1, 1, 5], [1, 2, 6], [1, 3, 7], [1, 4, 8],
      [2, 0, 9], [2, 1, 10], [2, 2, 11], [2, 3, 12], [2, 4, 13],
      [3, 1, 14], [3, 2, 15], [3, 3, 16], [3, 4, 17],
      [4, 2, 18], [4, 3, 19], [4, 4, 20], [4, 5, 21],
      [5, 3, 22], [5, 4, 23], [5, 5, 24], [5, 6, 25],
      [6, 0, 26], [6, 1, 27], [6, 2, 28], [6, 3, 29], [6, 4, 30],
      [7, 0, 31], [7, 1, 32], [7, 2, 33], [7, 3, 34], [7, 4, 35]]
cases.append(small("mediumSizeMatrix1", 8, 8, m4, 2, [20, 52, 110, 124, 156, 188, 280, 330], src=":267-343"))
# test_VectorProduct_mediumSizeMatrix2 (:345-418)
m5 = [[0, 0, 1], [0, 1, 2], [0, 2, 3], [0, 4, 4], [0, 5, 5], [0, 7, 1],
      [1, 1, 6], [1, 3, 7], [1, 7, 1],
      [2, 1, 8], [2, 2, 9], [2, 4, 10], [2, 7, 1],
      [3, 1, 11], [3, 2, 12], [3, 3, 13], [3, 4, 14], [3, 7, 1],
      [4, 1, 15], [4, 7, 1],
      [5, 1, 16], [5, 2, 17], [5, 3, 18], [5, 4, 19], [5, 5, 20], [5, 6, 21], [5, 7, 1],
      [6, 0, 22], [6, 1, 23], [6, 2, 24], [6, 3, 25], [6, 4, 26], [6, 5, 27], [6, 6, 28], [6, 7, 1],
      [7, 0, 29], [7, 1, 30], [7, 2, 31], [7, 3, 32], [7, 4, 33], [7, 5, 34], [7, 6, 35], [7, 7, 36]]
cases.append(small("mediumSizeMatrix2", 8, 8, m5, 2, [32, 28, 56, 102, 32, 224, 352, 520], src=":345-418"))
# test_VectorProduct_largeMatrix (:420-495): 1051 x 1051 diagonal a_ii = i+1, x = 1 -> y_i = i+1
n = 1051
cases.append(dict(name="largeMatrixDiagonal", rows=n, columns=n, capacities=None,
                  entries=[[i, i, i + 1] for i in range(n)], x=[1] * n, y=[i + 1 for i in range(n)],
                  source=":420-455"))
# lower triangular, capacity i+1, a_ij = j+1 (j <= i), x = 1 -> y_i = (i+1)(i+2)/2
cases.append(dict(name="largeMatrixTriangular", rows=n, columns=n, capacities=None,
                  entries="lower_triangular_jplus1", x=[1] * n, y=[(i + 1) * (i + 2) // 2 for i in range(n)],
                  source=":457-495"))
# test_VectorProduct_longRowsMatrix (:497-533): 33 rows x c columns, a_rj = j + r, x = 1
for c in [64, 65, 128, 129, 256, 257, 512, 513, 1024, 1025, 2048, 2049, 3000]:
    cases.append(dict(name=f"longRowsMatrix_{c}", rows=33, columns=c, capacities=None, entries="long_rows_jplusr",
                      x=[1] * c, y=[c * (c - 1) // 2 + c * r for r in range(33)], source=":497-533"))

out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "spmv_kats.json")
with open(out, "w") as f:
    json.dump(dict(reference="src/UnitTests/Matrices/SparseMatrixVectorProductTest.hpp", cases=cases), f)
print("wrote", out, len(cases), "cases")
