"""KKT linear problems on disk, in the text format that KKTSolver.printLinearProblem
writes (KKTSolver.java:1746-1788) and the PARDISO test harness reads
(pardisoMklTest.cc:215-300).

Format (whitespace separated):
  [# comment line]
  [SYMMETRIC] n
  rowOffs[0..n]            one-based CSR row offsets (rowOffs[n] = nnz + 1)
  colIdxs[0..nnz)          one-based column indices
  vals[0..nnz)
  rhs[0..n)                bf followed by bg; the reader substitutes ones when the
                           right-hand side is missing or incomplete

SYMMETRIC means the upper triangle is stored (KKTSolver writes it when M is
partitioned upper triangular).  With omitLowerRightDiagonal the trailing numG
values (the diagonal entries of the constraint rows) are left out and the row
offsets are clamped to nnz + 1, as the reference writer does.

The reference writes with "%g" (6 significant digits, writeLinearProblem); this
module writes with "%.17g" by default so a round trip is exact.
"""
import numpy as np


def write_linear_problem(path, rowOffs, colIdxs, vals, bf, bg=None, symmetric=True,
                         fmt="%.17g", omit_lower_right_diagonal=False, num_g=0, comment=None):
    """Writes a 0-based upper CSR system (as this package's solvers take it) in the
    reference's one-based text format.  bf, bg: the right-hand side split as
    KKTSolver.printLinearProblem receives it (bg may be None or empty)."""
    ro = np.asarray(rowOffs, np.int64)
    ci = np.asarray(colIdxs, np.int64)
    va = np.asarray(vals, np.float64)
    n = len(ro) - 1
    nnz = int(ro[n])
    if omit_lower_right_diagonal:
        nnz -= int(num_g)
    bf = np.asarray(bf, np.float64).ravel()
    bg = np.zeros(0) if bg is None else np.asarray(bg, np.float64).ravel()
    with open(path, "w") as f:
        if comment:
            f.write("# %s\n" % comment.replace("\n", " "))
        if symmetric:
            f.write("SYMMETRIC ")
        f.write("%d\n" % n)
        f.write(" ".join(str(min(nnz + 1, int(v) + 1)) for v in ro) + "\n")
        f.write(" ".join(str(int(v) + 1) for v in ci[:nnz]) + "\n")
        f.write(" ".join(fmt % v for v in va[:nnz]) + "\n")
        f.write(" ".join(fmt % v for v in bf))
        f.write(" ")
        f.write(" ".join(fmt % v for v in bg))
        f.write("\n")


def read_linear_problem(path, add_missing_diagonal=False):
    """Reads the reference's text format.  Returns a dict with 0-based CSR arrays
    (rowOffs int32[n+1], colIdxs int32[nnz], vals float64[nnz]), rhs float64[n]
    (ones when absent or incomplete, as pardisoMklTest.cc does), 'symmetric' and
    'rhs_present'.  With add_missing_diagonal, rows of a symmetric (upper) matrix
    that store no diagonal entry get an explicit zero one (the constraint rows
    written with omitLowerRightDiagonal), so the pattern has a full diagonal."""
    with open(path) as f:
        text = f.read()
    lines = text.split("\n")
    if lines and lines[0].lstrip().startswith("#"):
        lines = lines[1:]
    tok = " ".join(lines).split()
    pos = 0
    symmetric = False
    if tok[pos] == "SYMMETRIC":
        symmetric = True
        pos += 1
    n = int(tok[pos])
    pos += 1
    ro1 = np.array([int(t) for t in tok[pos:pos + n + 1]], np.int64)
    if len(ro1) != n + 1:
        raise ValueError("row offsets incomplete")
    pos += n + 1
    nnz = int(ro1[n]) - 1
    ci1 = np.array([int(t) for t in tok[pos:pos + nnz]], np.int64)
    if len(ci1) != nnz:
        raise ValueError("column indices incomplete")
    pos += nnz
    vals = np.array([float(t) for t in tok[pos:pos + nnz]], np.float64)
    if len(vals) != nnz:
        raise ValueError("values incomplete")
    pos += nnz
    rest = tok[pos:pos + n]
    rhs_present = len(rest) == n
    rhs = np.array([float(t) for t in rest], np.float64) if rhs_present else np.ones(n)
    ro = (ro1 - 1).astype(np.int32)
    ci = (ci1 - 1).astype(np.int32)
    if add_missing_diagonal and symmetric:
        ro, ci, vals = _with_diagonal(ro, ci, vals)
    return {"n": n, "symmetric": symmetric, "rowOffs": ro, "colIdxs": ci, "vals": vals,
            "rhs": rhs, "rhs_present": rhs_present}


def _with_diagonal(ro, ci, vals):
    n = len(ro) - 1
    rows = []
    for i in range(n):
        c = ci[ro[i]:ro[i + 1]]
        v = vals[ro[i]:ro[i + 1]]
        if not np.any(c == i):
            c = np.concatenate([[i], c])
            v = np.concatenate([[0.0], v])
        order = np.argsort(c, kind="stable")
        rows.append((c[order], v[order]))
    nro = np.zeros(n + 1, np.int32)
    nro[1:] = np.cumsum([len(c) for c, _ in rows])
    nci = np.concatenate([c for c, _ in rows]).astype(np.int32) if n else np.zeros(0, np.int32)
    nva = np.concatenate([v for _, v in rows]) if n else np.zeros(0)
    return nro, nci, nva
