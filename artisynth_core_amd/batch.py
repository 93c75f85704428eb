"""Batched KKT solves: many independent systems of uneven size factored and solved
together, through the library's kkt_batch_* entry points (include/kkt_mi355x.h).

The library concatenates the systems into one block-diagonal upper CSR whose
assembly forest holds one tree per system, so every factor and triangular-solve
launch covers the fronts of all systems at one tree level, and the whole factor
(and the whole triangular solve) is one captured HIP graph launch.  Pivot
perturbation, inertia and the refinement stopping test are per system.  This
module only packs the per-system arrays and reads results back.
"""
import ctypes

import numpy as np

from ._lib import KKT_PREC_FP64, KktBatchSysInfo, KktInfo, lib
from ._lib import ptr as nptr


class BatchedKKT:
    """Factor/solve a list of (rowOffs, colIdxs, vals, b) upper-CSR systems as one batch."""

    def __init__(self, systems, precision=KKT_PREC_FP64, ir_steps=1, refine_tol=None):
        self.L = lib()
        self.nsys = len(systems)
        self.sys_n = np.array([len(s[0]) - 1 for s in systems], np.int32)
        self.rowOffs = np.ascontiguousarray(np.concatenate([np.asarray(s[0], np.int32) for s in systems]))
        self.colIdxs = np.ascontiguousarray(np.concatenate([np.asarray(s[1], np.int32) for s in systems]))
        self.vals = np.ascontiguousarray(np.concatenate([np.asarray(s[2], np.float64) for s in systems]))
        self.b = np.ascontiguousarray(np.concatenate([np.asarray(s[3], np.float64) for s in systems]))
        self.offs = np.concatenate([[0], np.cumsum(self.sys_n, dtype=np.int64)])
        self.n = int(self.offs[-1])
        self.nnz = len(self.vals)
        h = self.L.kkt_batch_create()
        if not h:
            raise RuntimeError("kkt_batch_create failed")
        self.h = h
        self._check(self.L.kkt_batch_set_precision(h, precision), "setPrecision")
        self._check(self.L.kkt_batch_set_ir_steps(h, ir_steps), "setIrSteps")
        if refine_tol is not None:
            self._check(self.L.kkt_batch_set_refine_tol(h, refine_tol[0], refine_tol[1]), "setRefineTol")
        self._check(self.L.kkt_batch_set_systems(h, self.nsys, nptr(self.sys_n, ctypes.c_int),
                                                 nptr(self.rowOffs, ctypes.c_int),
                                                 nptr(self.colIdxs, ctypes.c_int), 1), "setSystems")
        self._check(self.L.kkt_batch_analyze(h, nptr(self.vals, ctypes.c_double)), "analyze")

    def _check(self, rc, op):
        if rc != 0:
            m = self.L.kkt_batch_last_error(self.h)
            raise RuntimeError("batch %s failed (%d): %s" % (op, rc, m.decode() if m else ""))

    def factor_device(self, d_vals):
        self._check(self.L.kkt_batch_factor_device(self.h, d_vals), "factor")

    def solve_device(self, d_b, d_x):
        self._check(self.L.kkt_batch_solve_device(self.h, d_b, d_x), "solve")

    def autotune_solve(self, reps=3):
        """Times the solve shapes per tree level on the factored batch and keeps
        the fastest (kkt_batch_autotune_solve)."""
        self._check(self.L.kkt_batch_autotune_solve(self.h, int(reps)), "autotune")

    def factor(self, vals=None):
        v = np.ascontiguousarray(self.vals if vals is None else vals, np.float64)
        self._check(self.L.kkt_batch_factor(self.h, nptr(v, ctypes.c_double)), "factor")

    def solve(self, b=None):
        bb = np.ascontiguousarray(self.b if b is None else b, np.float64)
        x = np.zeros(self.n)
        self._check(self.L.kkt_batch_solve(self.h, nptr(bb, ctypes.c_double), nptr(x, ctypes.c_double)),
                    "solve")
        return x

    def info(self):
        inf = KktInfo()
        self._check(self.L.kkt_batch_get_info(self.h, ctypes.byref(inf)), "getInfo")
        return inf.as_dict()

    def system_info(self, q):
        si = KktBatchSysInfo()
        self._check(self.L.kkt_batch_system_info(self.h, int(q), ctypes.byref(si)), "systemInfo")
        return si.as_dict()

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
            self.L.kkt_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.dispose()
        except Exception:
            pass


class IterativeGroup:
    """Stale-factor iterative solves of many factored handles in lockstep
    (kkt_group_iterative_solve_device): each BiCGStab step is one launch for every
    handle still iterating, and each handle's result is bitwise the one its own
    kkt_iterative_solve_device call gives."""

    def __init__(self):
        self.L = lib()
        self.g = self.L.kkt_group_create()
        if not self.g:
            raise RuntimeError("kkt_group_create failed")

    def solve(self, handles, d_vals, d_b, d_x, tol_rel, max_iter):
        """handles: raw kkt_solver pointers; d_vals, d_b, d_x: device pointers (ints).
        Returns (rc, statuses): per handle the iterations or a negative status."""
        k = len(handles)
        PA = ctypes.c_void_p * k
        st = np.zeros(k, np.int32)
        rc = self.L.kkt_group_iterative_solve_device(
            self.g, k, PA(*handles), PA(*d_vals), PA(*d_b), PA(*d_x), float(tol_rel), int(max_iter),
            nptr(st, ctypes.c_int))
        return rc, st

    def solve_host(self, handles, vals, b, x, tol_rel, max_iter):
        """The same on host arrays (kkt_group_iterative_solve): vals, b, x are lists
        of float64 numpy arrays; x[i] is written for every positive status."""
        k = len(handles)
        PA = ctypes.c_void_p * k
        ins = [np.ascontiguousarray(v, np.float64) for v in vals]
        bs = [np.ascontiguousarray(v, np.float64) for v in b]
        for xi in x:
            if xi.dtype != np.float64 or not xi.flags.c_contiguous:
                raise ValueError("x arrays must be contiguous float64")
        st = np.zeros(k, np.int32)
        rc = self.L.kkt_group_iterative_solve(
            self.g, k, PA(*handles), PA(*[a.ctypes.data for a in ins]), PA(*[a.ctypes.data for a in bs]),
            PA(*[a.ctypes.data for a in x]), float(tol_rel), int(max_iter), nptr(st, ctypes.c_int))
        return rc, st

    def last_error(self):
        m = self.L.kkt_group_last_error(self.g)
        return m.decode() if m else None

    def dispose(self):
        if self.g:
            self.L.kkt_group_destroy(self.g)
            self.g = None

    def __del__(self):
        try:
            self.dispose()
        except Exception:
            pass
