"""Host-side mirror of ArtiSynth's direct-solver interfaces over the MI355X C ABI.

`MI355XSolver` implements the `maspack.solvers.DirectSolver` contract
(DirectSolver.java:26-254) with the semantics of the reference's GPU backend
`CuDssSolver` (CuDssSolver.java:49-962): 0-based CSR arrays, state machine
UNSET -> ANALYZED -> FACTORED, NumericalException on native failures,
ImproperStateException on misuse, iterativeSolve returning iterations or a
non-positive value meaning "refactor".

`KKTSolver` mirrors the equality-constrained part of maspack.solvers.KKTSolver
(KKTSolver.java:171-454 analyzeMG / getCRSValues, 889-1014 factorAndSolve with
estimateOptimalCount, 1912-1934 solveMG with the lambda negation).
"""
import ctypes
import enum
import time

import numpy as np

from . import _lib
from ._lib import KKT_PREC_FP32_REFINE, KKT_PREC_FP64, KktInfo, lib, ptr


class NumericalException(RuntimeError):
    """maspack.matrix.NumericalException"""


class ImproperStateException(RuntimeError):
    """maspack.util.ImproperStateException"""


class Matrix:
    """Type flags of maspack.matrix.Matrix (Matrix.java:51-66)."""
    INDEFINITE = 0x0
    SYMMETRIC = 0x1
    POSITIVE_DEFINITE = 0x2
    SPD = 0x3


class SparseSolverId(enum.Enum):
    """maspack.solvers.SparseSolverId (SparseSolverId.java:8-98) plus the new backend."""
    Pardiso = "Pardiso"
    Umfpack = "Umfpack"
    CuDss = "CuDss"
    MI355X = "MI355X"

    def createDirectSolver(self, **kw):
        if self is SparseSolverId.MI355X:
            return MI355XSolver(**kw)
        raise UnsupportedOperationException("solver %s is not available in this build" % self.value)

    def isCompatible(self, matrixType):
        # SparseSolverId.java:59-75: symmetric types only for this backend
        if self is SparseSolverId.MI355X:
            return (matrixType & Matrix.SYMMETRIC) != 0
        return False


class UnsupportedOperationException(RuntimeError):
    pass


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def upper_csr(M):
    """0-based upper-triangular CSR (rowOffs, colIdxs, vals) of a scipy sparse matrix,
    columns sorted, diagonal always stored (Partition.UpperTriangular export)."""
    import scipy.sparse as sp
    n = M.shape[0]
    U = sp.triu(sp.csr_matrix(M), format="csr")
    U = U + sp.csr_matrix((np.zeros(n), (np.arange(n), np.arange(n))), shape=(n, n))
    U = sp.csr_matrix(U)
    U.sort_indices()
    return _i32(U.indptr), _i32(U.indices), _f64(U.data)


def _mtype_flag(type_):
    if type_ & Matrix.SYMMETRIC:
        return _lib.KKT_MT_SPD if (type_ & Matrix.POSITIVE_DEFINITE) else _lib.KKT_MT_SYMMETRIC
    return _lib.KKT_MT_GENERAL


class MI355XSolver:
    """DirectSolver on one MI355X (CuDssSolver semantics)."""

    UNSET, ANALYZED, FACTORED = 0, 1, 2
    _timing = False

    def __init__(self, precision=KKT_PREC_FP64, ordering=1):
        self._L = lib()
        self._h = self._L.kkt_create()
        if not self._h:
            raise UnsupportedOperationException("MI355X solver initialization failed")
        self._state = self.UNSET
        self._size = 0
        self._numVals = 0
        self._matrix = None
        self._type = 0
        self.setPrecision(precision)
        self._check(self._L.kkt_set_ordering(self._h, ordering), "setOrdering")

    # ---- plumbing ----
    @staticmethod
    def isAvailable():
        try:
            lib()
            return True
        except Exception:
            return False

    @staticmethod
    def getVersion():
        return lib().kkt_version().decode()

    @classmethod
    def setTimingEnabled(cls, on):
        cls._timing = bool(on)
        lib().kkt_set_timing_enabled(1 if on else 0)

    def getLastErrorMessage(self):
        if not self._h:
            return None
        m = self._L.kkt_last_error(self._h)
        return m.decode() if m else None

    def _check(self, status, op):
        if status != 0:
            raise NumericalException("MI355X %s failed (status %d %s): %s" % (
                op, status, _lib.ERRORS.get(status, "?"), self.getLastErrorMessage()))

    def _alive(self):
        if not self._h:
            raise ImproperStateException("Solver disposed")

    def setPrecision(self, precision):
        self._alive()
        self._check(self._L.kkt_set_precision(self._h, precision), "setPrecision")

    def setRefinement(self, tol=1e-10, maxSteps=10):
        self._alive()
        self._check(self._L.kkt_set_refine_tol(self._h, tol, maxSteps), "setRefinement")

    def setIterativeRefinementSteps(self, n):
        self._alive()
        self._check(self._L.kkt_set_ir_steps(self._h, n), "setIterativeRefinementSteps")

    def getState(self):
        return self._state

    def info(self):
        self._alive()
        inf = KktInfo()
        self._check(self._L.kkt_get_info(self._h, ctypes.byref(inf)), "getInfo")
        return inf.as_dict()

    def getNumNegEigenvalues(self):
        return self.info()["nneg"]

    def getNumPosEigenvalues(self):
        return self.info()["npos"]

    def getNumPerturbedPivots(self):
        return self.info()["nperturbed"]

    def symbolic(self):
        """Host symbolic arrays: perm, sup_ptr, sparent, rows_ptr, rows."""
        inf = self.info()
        n, ns = inf["n"], inf["nsuper"]
        perm = np.zeros(n, np.int32)
        sup = np.zeros(ns + 1, np.int32)
        spar = np.zeros(ns, np.int32)
        rp = np.zeros(ns + 1, np.int64)
        self._check(self._L.kkt_get_symbolic(self._h, ptr(perm, ctypes.c_int), ptr(sup, ctypes.c_int),
                                             ptr(spar, ctypes.c_int), ptr(rp, ctypes.c_longlong),
                                             None), "getSymbolic")
        rows = np.zeros(max(int(rp[-1]), 1), np.int32)
        self._check(self._L.kkt_get_symbolic(self._h, None, None, None, None,
                                             ptr(rows, ctypes.c_int)), "getSymbolic")
        return {"perm": perm, "sup_ptr": sup, "sparent": spar, "rows_ptr": rp,
                "rows": rows[:int(rp[-1])]}

    def verifyDeviceStructure(self):
        self._alive()
        return self._L.kkt_verify_device_structure(self._h) == 0

    # ---- DirectSolver: Matrix-object entry points ----
    def analyze(self, *args):
        """analyze(M, size, type) with a scipy sparse M (DirectSolver.java:48), or the
        array form analyze(vals, colIdxs, rowOffs, size, type) (DirectSolver.java:173)."""
        if len(args) == 3:
            return self._analyze_matrix(*args)
        if len(args) == 5:
            return self.analyzeArrays(*args)
        raise TypeError("analyze(M,size,type) or analyze(vals,colIdxs,rowOffs,size,type)")

    def _analyze_matrix(self, M, size, type_):
        self._alive()
        if M.shape[0] != M.shape[1]:
            raise ValueError("Matrix is not square")
        if size < 0 or size > M.shape[0]:
            raise ValueError("Requested size %d is out of bounds" % size)
        if not (type_ & Matrix.SYMMETRIC):
            raise ValueError("MI355X solver handles symmetric matrices only")
        sub = M[:size, :size]
        ro, ci, va = upper_csr(sub)
        self._set_pattern_and_analyze(ro, ci, va, size, type_)
        self._matrix = (M, size)

    def analyzeArrays(self, vals, colIdxs, rowOffs, size, type_):
        """0-based arrays, as CuDssSolver.analyze(double[],int[],int[],int,int) (:389-403)."""
        self._alive()
        ro = _i32(rowOffs[:size + 1])
        ci = _i32(colIdxs[:ro[size]])
        va = None if vals is None else _f64(vals[:ro[size]])
        self._set_pattern_and_analyze(ro, ci, va, size, type_)
        self._matrix = None

    def _set_pattern_and_analyze(self, ro, ci, va, size, type_):
        self._ro, self._ci = ro, ci
        self._check(self._L.kkt_set_pattern(self._h, size, int(ro[size]), ptr(ro, ctypes.c_int),
                                            ptr(ci, ctypes.c_int), _mtype_flag(type_)), "setPattern")
        if va is not None:
            self._check(self._L.kkt_analyze_values(self._h, ptr(va, ctypes.c_double)), "analyze")
        else:
            self._check(self._L.kkt_analyze(self._h), "analyze")
        self._size = size
        self._numVals = int(ro[size])
        self._type = type_
        self._state = self.ANALYZED

    def _matrix_vals(self):
        M, size = self._matrix
        ro, ci, va = upper_csr(M[:size, :size])
        if len(va) != self._numVals or not np.array_equal(ci, self._ci):
            raise ValueError("matrix structure changed since analyze()")
        return va

    def factor(self, vals=None):
        """factor() (DirectSolver.java:58) or factor(double[] vals) (:191)."""
        self._alive()
        if self._state == self.UNSET:
            raise ImproperStateException("analyze() not previously called")
        if vals is None:
            if self._matrix is None:
                raise ImproperStateException(
                    "analyze() or analyzeAndFactor() not previously called")
            vals = self._matrix_vals()
        va = _f64(vals)
        if len(va) < self._numVals:
            raise ValueError("vals has length %d, expected %d" % (len(va), self._numVals))
        self._check(self._L.kkt_factor(self._h, ptr(va, ctypes.c_double)), "factor")
        self._state = self.FACTORED

    def analyzeAndFactor(self, M):
        self.analyze(M, M.shape[0], Matrix.SYMMETRIC)
        self.factor()

    def solve(self, x, b, nrhs=None):
        """solve(x, b) (DirectSolver.java:98/202) or solve(X, B, nrhs) column-major (:223).
        x is written in place and returned."""
        self._alive()
        if self._state != self.FACTORED:
            raise ImproperStateException("factor() not previously called")
        n = self._size
        if nrhs is None:
            bb = _f64(b)
            if bb.size < n:
                raise ValueError("Right-hand side b has size %d, expected at least %d" % (bb.size, n))
            xx = np.zeros(n)
            self._check(self._L.kkt_solve(self._h, ptr(bb[:n].copy(), ctypes.c_double),
                                          ptr(xx, ctypes.c_double)), "solve")
            x[:n] = xx
            return x
        if nrhs <= 0:
            raise ValueError("nrhs must be positive")
        B = _f64(b)
        if B.size < n * nrhs or x.size < n * nrhs:
            raise ValueError("X and B must have length >= n*nrhs (%d)" % (n * nrhs))
        XX = np.zeros(n * nrhs)
        self._check(self._L.kkt_solve_multi(self._h, nrhs, ptr(B[:n * nrhs].copy(), ctypes.c_double),
                                            ptr(XX, ctypes.c_double)), "solveMulti")
        x[:n * nrhs] = XX
        return x

    # ---- device-side value assembly (CuDssSolver.java:421-788) ----
    def _need_analyze(self):
        self._alive()
        if self._state == self.UNSET:
            raise ImproperStateException("analyze() not previously called")

    def clearDeviceValues(self):
        self._need_analyze()
        self._check(self._L.kkt_clear_device_values(self._h), "clearDeviceValues")

    def getDeviceValues(self, vals):
        self._need_analyze()
        out = np.zeros(self._numVals)
        self._check(self._L.kkt_get_device_values(self._h, ptr(out, ctypes.c_double)), "getDeviceValues")
        vals[:self._numVals] = out
        return vals

    def addDeviceValues(self, slots, vals, nvals, scale):
        self._need_analyze()
        sl, va = _i32(slots), _f64(vals)
        if nvals < 0 or nvals > sl.size or nvals > va.size:
            raise ValueError("nvals exceeds slot/value array length")
        self._check(self._L.kkt_add_device_values(self._h, ptr(sl, ctypes.c_int), ptr(va, ctypes.c_double),
                                                  int(nvals), float(scale)), "addDeviceValues")

    def addScaledDiagonal3DeviceValues(self, diagSlots, masses, nblocks, scale):
        self._need_analyze()
        sl, m = _i32(diagSlots), _f64(masses)
        if nblocks < 0 or nblocks > sl.size // 3 or nblocks > m.size:
            raise ValueError("nblocks exceeds diagonal-slot/mass array length")
        self._check(self._L.kkt_add_scaled_diagonal3_device_values(
            self._h, ptr(sl, ctypes.c_int), ptr(m, ctypes.c_double), int(nblocks), float(scale)),
            "addScaledDiagonal3DeviceValues")

    def addScaledBlock3DeviceValues(self, blockSlots, blockVals, blockScales, nblocks, scale):
        self._need_analyze()
        sl, bv, bs = _i32(blockSlots), _f64(blockVals), _f64(blockScales)
        if nblocks < 0 or nblocks > sl.size // 9 or nblocks > bv.size // 9 or nblocks > bs.size:
            raise ValueError("nblocks exceeds block slot/value/scale array length")
        self._check(self._L.kkt_add_scaled_block3_device_values(
            self._h, ptr(sl, ctypes.c_int), ptr(bv, ctypes.c_double), ptr(bs, ctypes.c_double),
            int(nblocks), float(scale)), "addScaledBlock3DeviceValues")

    def addMaterialStiffness3DeviceValues(self, blockSlots, gis, gjs, Ds, sigmas, dvs, nblocks, scale):
        self._need_analyze()
        arr = [_i32(blockSlots), _f64(gis), _f64(gjs), _f64(Ds), _f64(sigmas), _f64(dvs)]
        per = [9, 3, 3, 36, 6, 1]
        if nblocks < 0 or any(nblocks > a.size // k for a, k in zip(arr, per)):
            raise ValueError("nblocks exceeds material stiffness descriptor array length")
        sl, gi, gj, D, sg, dv = arr
        self._check(self._L.kkt_add_material_stiffness3_device_values(
            self._h, ptr(sl, ctypes.c_int), ptr(gi, ctypes.c_double), ptr(gj, ctypes.c_double),
            ptr(D, ctypes.c_double), ptr(sg, ctypes.c_double), ptr(dv, ctypes.c_double),
            int(nblocks), float(scale)), "addMaterialStiffness3DeviceValues")

    # element-batched forms (CuDssSolver.java:580-775): the same argument checks
    # as the reference, then the native call
    def addMaterialStiffness3ElementDeviceValues(self, elemNodeCounts, elemNodeOffsets,
                                                 elemPairOffsets, elemIpOffsets, elemGradOffsets,
                                                 pairNodeIdxs, blockSlots, nodeDims, nodeTransforms,
                                                 grads, Ds, sigmas, dvs, nelems, scale):
        self._need_analyze()
        nc, no, po, io, go, pn, bs, nd = (_i32(a) for a in (
            elemNodeCounts, elemNodeOffsets, elemPairOffsets, elemIpOffsets, elemGradOffsets,
            pairNodeIdxs, blockSlots, nodeDims))
        T, g, D, sg, dv = (_f64(a) for a in (nodeTransforms, grads, Ds, sigmas, dvs))
        if nelems < 0 or nelems > nc.size or any(nelems + 1 > a.size for a in (no, po, io, go)):
            raise ValueError("nelems exceeds compact material stiffness element arrays")
        nn, npr, nip, ng = int(no[nelems]), int(po[nelems]), int(io[nelems]), int(go[nelems])
        if (min(nn, npr, nip, ng) < 0 or npr > pn.size // 2 or npr > bs.size // 36 or nn > nd.size
                or nn > T.size // 18 or nip > D.size // 36 or nip > sg.size // 6 or nip > dv.size
                or ng > g.size // 3):
            raise ValueError("compact material stiffness descriptor arrays are inconsistent")
        I, F = ctypes.c_int, ctypes.c_double
        self._check(self._L.kkt_add_material_stiffness3_element_device_values(
            self._h, ptr(nc, I), ptr(no, I), ptr(po, I), ptr(io, I), ptr(go, I), ptr(pn, I), ptr(bs, I),
            ptr(nd, I), ptr(T, F), ptr(g, F), ptr(D, F), ptr(sg, F), ptr(dv, F), int(nelems),
            float(scale)), "addMaterialStiffness3ElementDeviceValues")

    def addLinearElasticStiffness3ElementDeviceValues(self, elemNodeCounts, elemPairOffsets,
                                                      elemIpOffsets, elemGradOffsets, pairNodeIdxs,
                                                      blockSlots, elemParams, grads, dvs, nelems, scale):
        self._need_analyze()
        nc, po, io, go, pn, bs = (_i32(a) for a in (
            elemNodeCounts, elemPairOffsets, elemIpOffsets, elemGradOffsets, pairNodeIdxs, blockSlots))
        ep, g, dv = (_f64(a) for a in (elemParams, grads, dvs))
        if nelems < 0 or nelems > nc.size or any(nelems + 1 > a.size for a in (po, io, go)):
            raise ValueError("nelems exceeds compact linear elastic element arrays")
        npr, nip, ng = int(po[nelems]), int(io[nelems]), int(go[nelems])
        if (min(npr, nip, ng) < 0 or npr > pn.size // 2 or npr > bs.size // 9
                or nelems > ep.size // 2 or nip > dv.size or ng > g.size // 3):
            raise ValueError("compact linear elastic descriptor arrays are inconsistent")
        I, F = ctypes.c_int, ctypes.c_double
        self._check(self._L.kkt_add_linear_elastic_stiffness3_element_device_values(
            self._h, ptr(nc, I), ptr(po, I), ptr(io, I), ptr(go, I), ptr(pn, I), ptr(bs, I),
            ptr(ep, F), ptr(g, F), ptr(dv, F), int(nelems), float(scale)),
            "addLinearElasticStiffness3ElementDeviceValues")

    def addLinearElasticStiffness3ElementGeometryDeviceValues(
            self, elemNodeCounts, elemNodeOffsets, elemPairOffsets, elemIpOffsets,
            elemNaturalGradOffsets, pairNodeIdxs, blockSlots, nodeDims, nodeTransforms, elemParams,
            elemNodePositions, naturalGrads, ipWeights, nelems, scale):
        self._need_analyze()
        nc, no, po, io, go, pn, bs, nd = (_i32(a) for a in (
            elemNodeCounts, elemNodeOffsets, elemPairOffsets, elemIpOffsets, elemNaturalGradOffsets,
            pairNodeIdxs, blockSlots, nodeDims))
        T, ep, X, g, w = (_f64(a) for a in (nodeTransforms, elemParams, elemNodePositions,
                                             naturalGrads, ipWeights))
        if nelems < 0 or nelems > nc.size or any(nelems + 1 > a.size for a in (no, po, io, go)):
            raise ValueError("nelems exceeds compact linear elastic geometry arrays")
        nn, npr, nip, ng = int(no[nelems]), int(po[nelems]), int(io[nelems]), int(go[nelems])
        if (min(nn, npr, nip, ng) < 0 or npr > pn.size // 2 or npr > bs.size // 36 or nn > nd.size
                or nn > T.size // 18 or nelems > ep.size // 2 or nn > X.size // 3
                or ng > g.size // 3 or nip > w.size):
            raise ValueError("compact linear elastic geometry descriptor arrays are inconsistent")
        I, F = ctypes.c_int, ctypes.c_double
        self._check(self._L.kkt_add_linear_elastic_stiffness3_element_geometry_device_values(
            self._h, ptr(nc, I), ptr(no, I), ptr(po, I), ptr(io, I), ptr(go, I), ptr(pn, I), ptr(bs, I),
            ptr(nd, I), ptr(T, F), ptr(ep, F), ptr(X, F), ptr(g, F), ptr(w, F), int(nelems),
            float(scale)), "addLinearElasticStiffness3ElementGeometryDeviceValues")

    def addDilationalStiffness3ElementDeviceValues(
            self, elemNodeCounts, elemNodeOffsets, elemPressureCounts, elemPairOffsets,
            elemConstraintOffsets, elemRinvOffsets, pairNodeIdxs, blockSlots, nodeDims,
            nodeTransforms, constraints, rinvs, nelems, scale):
        self._need_analyze()
        nc, no, pc, po, co, ro, pn, bs, nd = (_i32(a) for a in (
            elemNodeCounts, elemNodeOffsets, elemPressureCounts, elemPairOffsets,
            elemConstraintOffsets, elemRinvOffsets, pairNodeIdxs, blockSlots, nodeDims))
        T, C, R = (_f64(a) for a in (nodeTransforms, constraints, rinvs))
        if (nelems < 0 or nelems > nc.size or nelems > pc.size
                or any(nelems + 1 > a.size for a in (no, po, co, ro))):
            raise ValueError("nelems exceeds compact dilational element arrays")
        nn, npr, ncon, nr = int(no[nelems]), int(po[nelems]), int(co[nelems]), int(ro[nelems])
        if (min(nn, npr, ncon, nr) < 0 or npr > pn.size // 2 or npr > bs.size // 36 or nn > nd.size
                or nn > T.size // 18 or ncon > C.size // 3 or nr > R.size):
            raise ValueError("compact dilational descriptor arrays are inconsistent")
        I, F = ctypes.c_int, ctypes.c_double
        self._check(self._L.kkt_add_dilational_stiffness3_element_device_values(
            self._h, ptr(nc, I), ptr(no, I), ptr(pc, I), ptr(po, I), ptr(co, I), ptr(ro, I),
            ptr(pn, I), ptr(bs, I), ptr(nd, I), ptr(T, F), ptr(C, F), ptr(R, F), int(nelems),
            float(scale)), "addDilationalStiffness3ElementDeviceValues")

    def autotuneSolve(self, reps=3):
        """Picks the solve kernels' per-level lane shapes by timing them on the
        current factor (kkt_autotune_solve); returns getSolveShapes()."""
        self._check(self._L.kkt_autotune_solve(self._h, int(reps)), "autotuneSolve")
        return self.getSolveShapes()

    def getSolveShapes(self):
        """(forward, backward) elements per lane of every tree level."""
        nl = self.info()["nlevels"]
        f = np.zeros(max(nl, 1), np.int32)
        b = np.zeros(max(nl, 1), np.int32)
        self._check(self._L.kkt_get_solve_shapes(self._h, ptr(f, ctypes.c_int), ptr(b, ctypes.c_int)),
                    "getSolveShapes")
        return f[:nl], b[:nl]

    def getForwardForms(self):
        """Forward launches of every tree level: 1 (row blocks gather for
        themselves) or 2 (a gather launch, then the rows)."""
        nl = self.info()["nlevels"]
        f = np.zeros(max(nl, 1), np.int32)
        self._check(self._L.kkt_get_forward_forms(self._h, ptr(f, ctypes.c_int)), "getForwardForms")
        return f[:nl]

    def setForwardForms(self, forms):
        """Sets the forward form of every tree level (1 or 2 each); the solution
        is bitwise the same either way."""
        f = np.ascontiguousarray(forms, np.int32)
        if len(f) != self.info()["nlevels"]:
            raise ValueError("one form per tree level expected")
        self._check(self._L.kkt_set_forward_forms(self._h, ptr(f, ctypes.c_int)), "setForwardForms")

    def getForwardRowGroups(self):
        """Row groups per forward row block of every tree level (1, 2, 4 or 8)."""
        nl = self.info()["nlevels"]
        g = np.zeros(max(nl, 1), np.int32)
        self._check(self._L.kkt_get_forward_row_groups(self._h, ptr(g, ctypes.c_int)), "getForwardRowGroups")
        return g[:nl]

    def setMaxSupernode(self, cols):
        """Widest supernode (64..2048 columns) of the next analysis: wider means
        fewer tree levels (faster triangular solves) for more factor work."""
        self._check(self._L.kkt_set_max_supernode(self._h, int(cols)), "setMaxSupernode")

    def setIterativeHalf(self, on):
        """1: iterativeSolve's preconditioner reads a half-precision copy of an
        FP32 factor's solve arrays (FP32 arithmetic); 0 (default): the FP32 arrays.
        Applies from the next factor; info()["precond_half"] says which is used."""
        self._alive()
        self._check(self._L.kkt_set_iterative_half(self._h, 1 if on else 0), "setIterativeHalf")

    def setForwardRowGroups(self, groups):
        """Sets the row groups of every tree level; the solution is bitwise the same."""
        g = np.ascontiguousarray(groups, np.int32)
        if len(g) != self.info()["nlevels"]:
            raise ValueError("one value per tree level expected")
        self._check(self._L.kkt_set_forward_row_groups(self._h, ptr(g, ctypes.c_int)), "setForwardRowGroups")

    def factorDeviceValues(self):
        self._need_analyze()
        self._check(self._L.kkt_factor_device_values(self._h), "factorDeviceValues")
        self._state = self.FACTORED

    def multiply(self, x, y):
        self._alive()
        xx = _f64(x)
        yy = np.zeros(self._size)
        self._check(self._L.kkt_multiply(self._h, ptr(xx, ctypes.c_double), ptr(yy, ctypes.c_double)),
                    "multiply")
        y[:self._size] = yy
        return y

    def hasAutoIterativeSolving(self):
        return True

    def iterativeSolve(self, vals, x, b, tolExp):
        """CuDssSolver.iterativeSolve (:934-942): BiCGStab with the stale factor,
        tolRel = 10^-tolExp, maxIter = 32; returns iterations (>0) or <= 0."""
        self._alive()
        if self._state != self.FACTORED:
            return -1
        tolRel = 10.0 ** (-tolExp)
        va = _f64(vals)
        bb = _f64(b)
        xx = np.zeros(self._size)
        rc = self._L.kkt_iterative_solve(self._h, ptr(va, ctypes.c_double), ptr(bb, ctypes.c_double),
                                         ptr(xx, ctypes.c_double), tolRel, 32)
        if rc > 0:
            x[:self._size] = xx
        return rc

    def autoFactorAndSolve(self, x, b, tolExp):
        """CuDssSolver.autoFactorAndSolve (:866-913).  lastAutoIterations records the
        branch taken: the BiCGStab iterations with the stale factor, or 0 after a
        factor + solve."""
        self._alive()
        if self._state == self.UNSET or self._matrix is None:
            raise ImproperStateException(
                "analyze(Matrix) or analyzeAndFactor(Matrix) not previously called")
        self.lastAutoIterations = 0
        if tolExp <= 0 or self._state != self.FACTORED:
            self.factor()
            self.solve(x, b)
            return
        vals = self._matrix_vals()
        xx = np.zeros(self._size)
        rc = self.iterativeSolve(vals, xx, _f64(b)[:self._size], tolExp)
        if rc > 0:
            x[:self._size] = xx
            self.lastAutoIterations = rc
            return
        self.factor()
        self.solve(x, b)

    def dispose(self):
        if getattr(self, "_h", None):
            self._L.kkt_destroy(self._h)
            self._h = None
        self._state = self.UNSET

    def __del__(self):
        try:
            self.dispose()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# KKT layer
# ---------------------------------------------------------------------------
def kkt_upper_csr(M, GT, Rg=None):
    """CSR of [M G^T; G -Rg] (upper triangle for symmetric M), in the layout of
    KKTSolver.analyzeMG / getCRSValues (KKTSolver.java:171-454): rows < sizeM hold
    M's upper entries then G^T's entries (columns sizeM+j); row sizeM+i holds
    only the lower-right diagonal -Rg[i] (explicit 0 when Rg is None)."""
    import scipy.sparse as sp
    sizeM = M.shape[0]
    numG = 0 if GT is None else GT.shape[1]
    U = sp.csr_matrix(sp.triu(sp.csr_matrix(M)))
    U.sort_indices()
    if numG:
        G = sp.csr_matrix(GT)
        G.sort_indices()
    rows, cols, vals = [0], [], []
    for i in range(sizeM):
        c = U.indices[U.indptr[i]:U.indptr[i + 1]]
        v = U.data[U.indptr[i]:U.indptr[i + 1]]
        if numG:
            c = np.concatenate([c, sizeM + G.indices[G.indptr[i]:G.indptr[i + 1]]])
            v = np.concatenate([v, G.data[G.indptr[i]:G.indptr[i + 1]]])
        cols.append(c)
        vals.append(v)
        rows.append(rows[-1] + len(c))
    for i in range(numG):
        cols.append(np.array([sizeM + i]))
        vals.append(np.array([-(Rg[i] if Rg is not None else 0.0)]))
        rows.append(rows[-1] + 1)
    return (_i32(rows), _i32(np.concatenate(cols) if cols else []),
            _f64(np.concatenate(vals) if vals else []))


class KKTSolver:
    """Equality-constrained KKT solves on the MI355X backend (KKTSolver.java)."""

    def __init__(self, solverType=SparseSolverId.MI355X, precision=KKT_PREC_FP64):
        if solverType is not SparseSolverId.MI355X:
            raise ValueError("Solver type %s not supported" % solverType)
        self.solver = MI355XSolver(precision=precision)
        self._state = "UNSET"
        self.directCnt = 0
        self.directTimeMsec = 0.0
        self.iterativeCnt = 0
        self.iterativeTimeMsec = 0.0
        self.firstIterativeTimeMsec = 0.0
        self.lastSolveWasIterative = False

    def analyze(self, M, sizeM, GT, Rg, typeM):
        """analyzeMG (KKTSolver.java:171-288)."""
        ro, ci, va = kkt_upper_csr(M, GT, Rg)
        self.sizeM = sizeM
        self.numG = 0 if GT is None else GT.shape[1]
        self._ci = ci
        self._ro = ro
        self._va = va
        kktType = Matrix.SYMMETRIC if (typeM & Matrix.SYMMETRIC) else Matrix.INDEFINITE
        if not (kktType & Matrix.SYMMETRIC):
            raise ValueError("MI355X backend handles symmetric KKT systems only")
        self.solver.analyzeArrays(va, ci, ro, sizeM + self.numG, kktType)
        self.directCnt = 0
        self.directTimeMsec = 0.0
        self.iterativeCnt = 0
        self.iterativeTimeMsec = 0.0
        self._state = "ANALYZED"

    def _vals(self, M, GT, Rg):
        ro, ci, va = kkt_upper_csr(M, GT, Rg)
        if not np.array_equal(ci, self._ci):
            raise ValueError("KKT structure changed since analyze()")
        return va

    def factor(self, M, GT, Rg):
        """factorMG (KKTSolver.java:1958-1994)."""
        self._va = self._vals(M, GT, Rg)
        self.solver.factor(self._va)
        self._state = "FACTORED"

    def printLinearProblem(self, path, bf, bg, fmtStr="%.17g", omitLowerRightDiagonal=False):
        """KKTSolver.printLinearProblem (KKTSolver.java:1746-1788): the current KKT
        values (upper CSR, one-based on disk) and the right-hand side [bf; bg]."""
        from .kktio import write_linear_problem
        write_linear_problem(path, self._ro, self._ci, self._va, bf, bg, symmetric=True,
                             fmt=fmtStr, omit_lower_right_diagonal=omitLowerRightDiagonal,
                             num_g=self.numG)

    def writeLinearProblem(self, fileName, bf, bg):
        """KKTSolver.writeLinearProblem (:1790-1800), at full precision ("%.17g";
        the reference's "%g" keeps 6 digits)."""
        self.printLinearProblem(fileName, bf, bg)

    def solve(self, vel, lam, bm, bg):
        """solveMG (KKTSolver.java:1912-1934): solves with [bm; bg] and negates lam."""
        if self._state != "FACTORED":
            raise ImproperStateException("Factor has not been called")
        b = np.concatenate([_f64(bm), _f64(bg)])
        y = np.zeros(self.sizeM + self.numG)
        self.solver.solve(y, b)
        vel[:] = y[:self.sizeM]
        lam[:] = -y[self.sizeM:]

    def estimateOptimalCount(self):
        """KKTSolver.java:889-922."""
        ft = self.directTimeMsec / self.directCnt
        it0 = self.firstIterativeTimeMsec
        itAvg = self.iterativeTimeMsec / self.iterativeCnt
        if itAvg >= ft:
            return self.iterativeCnt + 1
        dt = 2 * (itAvg - it0) / self.iterativeCnt
        if dt <= 0:
            return self.iterativeCnt + 19
        return int(np.ceil(np.sqrt(2 * (ft - it0 + dt) / dt)))

    def factorAndSolve(self, M, GT, Rg, vel, lam, bm, bg, tolExp):
        """KKTSolver.factorAndSolve (:930-1014): stale-factor iterative solve while it
        pays, else a real factor + solve."""
        self.lastSolveWasIterative = False
        if self._state != "FACTORED":
            raise ImproperStateException("Factor has not been called")
        b = np.concatenate([_f64(bm), _f64(bg)])
        if (self.solver.hasAutoIterativeSolving() and self.directCnt > 0 and
                (self.iterativeCnt == 0 or self.iterativeCnt + 1 < self.estimateOptimalCount())):
            t0 = time.perf_counter()
            vals = self._vals(M, GT, Rg)
            y = np.zeros(self.sizeM + self.numG)
            st = self.solver.iterativeSolve(vals, y, b, tolExp)
            if st > 0:
                vel[:] = y[:self.sizeM]
                lam[:] = -y[self.sizeM:]
                self.iterativeTimeMsec += (time.perf_counter() - t0) * 1e3
                if self.iterativeCnt == 0:
                    self.firstIterativeTimeMsec = self.iterativeTimeMsec
                self.iterativeCnt += 1
                self.lastSolveWasIterative = True
                return st
        t0 = time.perf_counter()
        self.factor(M, GT, Rg)
        self.solve(vel, lam, bm, bg)
        self.directTimeMsec += (time.perf_counter() - t0) * 1e3
        self.directCnt += 1
        self.iterativeCnt = 0
        self.iterativeTimeMsec = 0.0
        return 0

    def dispose(self):
        self.solver.dispose()
