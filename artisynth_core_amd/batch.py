"""Batched KKT solves: many independent systems of uneven size factored and solved
together.

The systems are concatenated into one block-diagonal upper CSR.  Its assembly
forest (one tree per system) goes through the same level-scheduled kernels as a
single large system: every launch covers the fronts of all systems at the same
tree level, and the whole factor (and the whole triangular solve) is one
captured HIP graph launch.  Per-system results are slices of the batch vectors.
"""
import ctypes

import numpy as np

from . import generators as G
from ._lib import KKT_PREC_FP64, lib
from ._lib import ptr as nptr


class BatchedKKT:
    """Factor/solve a list of (rowOffs, colIdxs, vals, b) upper-CSR systems as one batch."""

    def __init__(self, systems, precision=KKT_PREC_FP64, ir_steps=1):
        self.L = lib()
        ro, ci, va, b, offs = G.block_diag_csr(systems)
        self.rowOffs, self.colIdxs, self.vals, self.b, self.offs = ro, ci, va, b, offs
        self.n = len(ro) - 1
        self.nnz = int(ro[-1])
        self.nsys = len(systems)
        self.val_offs = np.zeros(self.nsys + 1, np.int64)
        for s, sysd in enumerate(systems):
            self.val_offs[s + 1] = self.val_offs[s] + int(sysd[0][-1])
        h = self.L.kkt_create()
        if not h:
            raise RuntimeError("kkt_create failed")
        self.h = h
        self._check(self.L.kkt_set_precision(h, precision), "setPrecision")
        self._check(self.L.kkt_set_ir_steps(h, ir_steps), "setIrSteps")
        self._check(self.L.kkt_set_pattern(h, self.n, self.nnz, nptr(ro, ctypes.c_int),
                                           nptr(ci, ctypes.c_int), 1), "setPattern")
        self._check(self.L.kkt_analyze_values(h, nptr(va, ctypes.c_double)), "analyze")

    def _check(self, rc, op):
        if rc != 0:
            m = self.L.kkt_last_error(self.h)
            raise RuntimeError("batch %s failed (%d): %s" % (op, rc, m.decode() if m else ""))

    def factor_device(self, d_vals):
        self._check(self.L.kkt_factor_device(self.h, d_vals), "factor")

    def solve_device(self, d_b, d_x):
        self._check(self.L.kkt_solve_device(self.h, d_b, d_x), "solve")

    def factor(self, vals=None):
        v = np.ascontiguousarray(self.vals if vals is None else vals, np.float64)
        self._check(self.L.kkt_factor(self.h, nptr(v, ctypes.c_double)), "factor")

    def solve(self, b=None):
        bb = np.ascontiguousarray(self.b if b is None else b, np.float64)
        x = np.zeros(self.n)
        self._check(self.L.kkt_solve(self.h, nptr(bb, ctypes.c_double), nptr(x, ctypes.c_double)),
                    "solve")
        return x

    def split(self, x):
        return [x[self.offs[s]:self.offs[s + 1]] for s in range(self.nsys)]

    def per_system_residuals(self, x, systems):
        from scipy import sparse as sp
        out = []
        for s, (ro, ci, va, b) in enumerate(systems):
            n = len(ro) - 1
            U = sp.csr_matrix((va, ci, ro), shape=(n, n))
            A = U + sp.triu(U, 1).T
            xs = x[self.offs[s]:self.offs[s + 1]]
            out.append(float(np.linalg.norm(A @ xs - b) / max(np.linalg.norm(b), 1e-300)))
        return out

    def dispose(self):
        if getattr(self, "h", None):
            self.L.kkt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.dispose()
        except Exception:
            pass
