"""Shared helpers for the tests (test infrastructure)."""
import numpy as np
import scipy.sparse as sp


def full_from_upper(rowOffs, colIdxs, vals, n=None):
    n = len(rowOffs) - 1 if n is None else n
    U = sp.csr_matrix((np.asarray(vals, float), np.asarray(colIdxs), np.asarray(rowOffs)), shape=(n, n))
    U = sp.triu(U)
    return sp.csr_matrix(U + sp.triu(U, 1).T)


def rel_residual(rowOffs, colIdxs, vals, x, b):
    A = full_from_upper(rowOffs, colIdxs, vals)
    return np.linalg.norm(A @ x - b) / max(np.linalg.norm(b), 1e-300)


def rel_err(x, ref):
    return np.linalg.norm(np.asarray(x) - np.asarray(ref)) / max(np.linalg.norm(ref), 1e-300)


# PardisoSolverTest.java:407-543 (upper CSR, 0-based)
P5_ROWS = np.array([0, 3, 6, 8, 10, 11], np.int32)
P5_COLS = np.array([0, 1, 2, 1, 2, 3, 2, 3, 3, 4, 4], np.int32)
P5_VALS = [
    (np.array([3, 1, 2, 0, 1, 2, 4, 1, 0, 6, 2], float),
     np.array([0.111111111111111, -0.88888888888888, 0.777777777777777, 0.555555555555555,
               0.833333333333333])),
    (np.array([3, 1, 2, 10, 1, 2, 4, 1, 10, 5, 2], float),
     np.array([0.6032064128257, -0.4368737474950, -0.1863727454910, 2.9759519038076,
               -4.9398797595190])),
    (np.array([5, 1, 2, 12, 1, 2, 4, 1, 9, 5, 2], float),
     np.array([0.1966035271065, -0.2482037883736, 0.1325930764206, 2.3246244284781,
               -3.3115610711953])),
]
P5_B = np.array([1, 2, 3, 4, 5], float)

# CuDssSolverTest.java:343-349 SPD5, 0-based upper CSR
SPD5_ROWS = np.array([0, 2, 4, 5, 7, 8], np.int32)
SPD5_COLS = np.array([0, 2, 1, 2, 2, 3, 4, 4], np.int32)
SPD5_VALS = np.array([4, 1, 3, 2, 5, 1, 1, 2], float)
SPD5_RHS = np.array([7, 12, 20, 9, 14], float)
SPD5_X = np.array([1, 2, 3, 4, 5], float)

# CuDssSolverTest.java:79-107 symmetric indefinite 3x3 (eigenvalues 3, 3, -1)
IND3_ROWS = np.array([0, 2, 3, 4], np.int32)
IND3_COLS = np.array([0, 1, 1, 2], np.int32)
IND3_VALS = np.array([1, 2, 1, 3], float)

# CuDssSolverTest.java:870-947 5x5 KKT with -0.1 regularisation
KKT5_ROWS = np.array([0, 3, 6, 9, 10, 11], np.int32)
KKT5_COLS = np.array([0, 2, 3, 1, 2, 4, 2, 3, 4, 3, 4], np.int32)
KKT5_VALS = np.array([4, 1, 1, 3, 2, 1, 5, 1, 1, -0.1, -0.1], float)
KKT5_B = np.array([1, 2, 3, 4, 5], float)
