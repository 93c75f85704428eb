"""ctypes face of the CPU oracle (oracle/kkt_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product path never loads it.
Reference citations per function are in kkt_oracle.h.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libkkt_oracle.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle` (%s)" % _LIB_PATH)
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_spmv_upper.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.orc_spmv_upper.restype = None
        L.orc_symbolic.argtypes = [ctypes.c_int, _i32p, _i32p, _i32p, _i32p, _i32p]
        L.orc_symbolic.restype = ctypes.c_int
        L.orc_supernode_rows.argtypes = [ctypes.c_int, _i32p, _i32p, _i32p, ctypes.c_int,
                                         _i32p, _i64p, _i32p, ctypes.c_long]
        L.orc_supernode_rows.restype = ctypes.c_long
        L.orc_factor_create.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, _i32p, ctypes.c_int,
                                        _i32p, ctypes.c_int, ctypes.c_double]
        L.orc_factor_create.restype = ctypes.c_void_p
        L.orc_factor_free.argtypes = [ctypes.c_void_p]
        L.orc_factor_free.restype = None
        L.orc_factor_solve.argtypes = [ctypes.c_void_p, _f64p, _f64p]
        L.orc_factor_solve.restype = None
        L.orc_factor_info.argtypes = [ctypes.c_void_p, _i64p]
        L.orc_factor_info.restype = None
        L.orc_bicgstab.argtypes = [ctypes.c_void_p, ctypes.c_int, _i32p, _i32p, _f64p, _f64p,
                                   _f64p, ctypes.c_double, ctypes.c_int]
        L.orc_bicgstab.restype = ctypes.c_int
        L.orc_refine.argtypes = [ctypes.c_void_p, ctypes.c_int, _i32p, _i32p, _f64p, _f64p,
                                 _f64p, ctypes.c_double, ctypes.c_int]
        L.orc_refine.restype = ctypes.c_int
        _lib = L
    return _lib


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def spmv_upper(rowOffs, colIdxs, vals, x):
    n = len(rowOffs) - 1
    y = np.zeros(n)
    lib().orc_spmv_upper(n, _i32(rowOffs), _i32(colIdxs), _f64(vals), _f64(x), y)
    return y


def symbolic(rowOffs, colIdxs, perm):
    n = len(rowOffs) - 1
    parent = np.zeros(n, np.int32)
    cc = np.zeros(n, np.int32)
    rc = lib().orc_symbolic(n, _i32(rowOffs), _i32(colIdxs), _i32(perm), parent, cc)
    if rc != 0:
        raise RuntimeError("orc_symbolic failed")
    return parent, cc


def supernode_rows(rowOffs, colIdxs, perm, sup_ptr):
    n = len(rowOffs) - 1
    ns = len(sup_ptr) - 1
    cap = max(1024, 16 * n)
    while True:
        rp = np.zeros(ns + 1, np.int64)
        rows = np.zeros(cap, np.int32)
        t = lib().orc_supernode_rows(n, _i32(rowOffs), _i32(colIdxs), _i32(perm), ns,
                                     _i32(sup_ptr), rp, rows, cap)
        if t == -2:
            cap *= 4
            continue
        if t < 0:
            raise RuntimeError("orc_supernode_rows failed")
        return rp, rows[:t].copy()


class Factor:
    """FP64 multifrontal LDL^T with panel-restricted Bunch-Kaufman pivoting."""

    def __init__(self, rowOffs, colIdxs, vals, perm=None, sup_ptr=None, nb=64, pert=1e-8):
        self.rowOffs = _i32(rowOffs)
        self.colIdxs = _i32(colIdxs)
        self.vals = _f64(vals)
        self.n = n = len(rowOffs) - 1
        if perm is None:
            perm = np.arange(n, dtype=np.int32)
        if sup_ptr is None:
            sup_ptr = np.array([0, n], np.int32)
        self._h = lib().orc_factor_create(n, self.rowOffs, self.colIdxs, self.vals, _i32(perm),
                                          len(sup_ptr) - 1, _i32(sup_ptr), nb, pert)
        if not self._h:
            raise RuntimeError("oracle factorization failed")

    def solve(self, b):
        x = np.zeros(self.n)
        lib().orc_factor_solve(self._h, _f64(b), x)
        return x

    def info(self):
        a = np.zeros(5, np.int64)
        lib().orc_factor_info(self._h, a)
        return {"npos": int(a[0]), "nneg": int(a[1]), "nperturbed": int(a[2]),
                "n2x2": int(a[3]), "nnzL": int(a[4])}

    def bicgstab(self, b, tolRel, maxIter=32, vals=None):
        x = np.zeros(self.n)
        v = self.vals if vals is None else _f64(vals)
        it = lib().orc_bicgstab(self._h, self.n, self.rowOffs, self.colIdxs, v, _f64(b), x,
                                tolRel, maxIter)
        return it, x

    def refine(self, b, tol=1e-8, maxSteps=20, x0=None):
        x = self.solve(b) if x0 is None else _f64(x0).copy()
        steps = lib().orc_refine(self._h, self.n, self.rowOffs, self.colIdxs, self.vals,
                                 _f64(b), x, tol, maxSteps)
        return steps, x

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_factor_free(h)
            self._h = None


# ---------------------------------------------------------------------------
# Device-side value assembly (test restatement)
# ---------------------------------------------------------------------------
def voigt_B(g):
    """6x3 strain-displacement matrix of a shape-function gradient g, Voigt strain
    order xx, yy, zz, xy, yz, xz with engineering shears (FemUtilities.java:78-124)."""
    gx, gy, gz = g
    return np.array([[gx, 0, 0], [0, gy, 0], [0, 0, gz],
                     [gy, gx, 0], [0, gz, gy], [gz, 0, gx]], float)


def material_stiffness3(gi, gj, D, sig, dv):
    """FemUtilities.addMaterialStiffness(K, gi, D, sig, gj, dv)
    (FemUtilities.java:61-66, 78-124, 131-146): K = dv (Bi^T D Bj + (gi^T S gj) I),
    D 6x6 row-major, sig = [s00, s11, s22, s01, s12, s02]."""
    D = np.asarray(D, float).reshape(6, 6)
    s = np.asarray(sig, float)
    S = np.array([[s[0], s[3], s[5]], [s[3], s[1], s[4]], [s[5], s[4], s[2]]])
    gi = np.asarray(gi, float)
    gj = np.asarray(gj, float)
    return dv * (voigt_B(gi).T @ D @ voigt_B(gj) + (gi @ S @ gj) * np.eye(3))


def isotropic_D(lam, mu):
    """Linear isotropic tangent in the same Voigt convention (engineering shears)."""
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[0, 0] = D[1, 1] = D[2, 2] = lam + 2 * mu
    D[3, 3] = D[4, 4] = D[5, 5] = mu
    return D


def scatter_add(vals, slots, contrib, scale):
    """vals[slots[e]] += scale * contrib[e] for slots >= 0, in index order."""
    slots = np.asarray(slots)
    keep = slots >= 0
    np.add.at(vals, slots[keep], scale * np.asarray(contrib, float)[keep])
    return vals


# ---- element-batched assembly forms (cudssBridgeKernels.cu:336-739) ----
# Each returns (slots, contrib): one entry per target-block position of every
# (integration point, node pair), in element / point / pair / row-major order,
# which scatter_add then applies.  The reference adds them with atomicAdd.

def reduce_block3(K, Ti, rdim, Tj, cdim):
    """T_i K T_j^T with T row-major (dim x 3) in an 18-entry slot
    (scatterReduced3Block, cudssBridgeKernels.cu:336-367)."""
    A = np.asarray(Ti, float)[:3 * rdim].reshape(rdim, 3)
    B = np.asarray(Tj, float)[:3 * cdim].reshape(cdim, 3)
    return A @ np.asarray(K, float).reshape(3, 3) @ B.T


def linear_elastic_D(E, nu):
    """fillLinearElasticD (cudssBridgeKernels.cu:451-468)."""
    a = E / (1.0 + nu)
    D = np.zeros((6, 6))
    D[:3, :3] = nu / (1.0 - 2.0 * nu) * a
    D[0, 0] = D[1, 1] = D[2, 2] = (1.0 - nu) / (1.0 - 2.0 * nu) * a
    D[3, 3] = D[4, 4] = D[5, 5] = 0.5 * a
    return D


def _emit36(out_s, out_c, blockSlots, pidx, R):
    rd, cd = R.shape
    for r in range(rd):
        for c in range(cd):
            out_s.append(int(blockSlots[36 * pidx + r * cd + c]))
            out_c.append(R[r, c])


def material_stiffness3_element(elemNodeCounts, elemNodeOffsets, elemPairOffsets, elemIpOffsets,
                                elemGradOffsets, pairNodeIdxs, blockSlots, nodeDims, nodeTransforms,
                                grads, Ds, sigmas, dvs, nelems):
    """addMaterialStiffness3ElementKernel (cudssBridgeKernels.cu:381-449)."""
    S, C = [], []
    for e in range(nelems):
        nn, n0 = elemNodeCounts[e], elemNodeOffsets[e]
        for ip in range(elemIpOffsets[e], elemIpOffsets[e + 1]):
            gb = elemGradOffsets[e] + (ip - elemIpOffsets[e]) * nn
            for p in range(elemPairOffsets[e], elemPairOffsets[e + 1]):
                i, j = pairNodeIdxs[2 * p], pairNodeIdxs[2 * p + 1]
                K = material_stiffness3(grads[3 * (gb + i):3 * (gb + i) + 3],
                                        grads[3 * (gb + j):3 * (gb + j) + 3],
                                        Ds[36 * ip:36 * ip + 36], sigmas[6 * ip:6 * ip + 6], dvs[ip])
                R = reduce_block3(K, nodeTransforms[18 * (n0 + i):18 * (n0 + i) + 18], nodeDims[n0 + i],
                                  nodeTransforms[18 * (n0 + j):18 * (n0 + j) + 18], nodeDims[n0 + j])
                _emit36(S, C, blockSlots, p, R)
    return np.array(S, np.int64), np.array(C, float)


def linear_elastic_stiffness3_element(elemNodeCounts, elemPairOffsets, elemIpOffsets, elemGradOffsets,
                                      pairNodeIdxs, blockSlots, elemParams, grads, dvs, nelems):
    """addLinearElasticStiffness3ElementKernel (cudssBridgeKernels.cu:470-512):
    9 slots per pair, no node reduction."""
    S, C = [], []
    z = np.zeros(6)
    for e in range(nelems):
        nn = elemNodeCounts[e]
        D = linear_elastic_D(elemParams[2 * e], elemParams[2 * e + 1])
        for ip in range(elemIpOffsets[e], elemIpOffsets[e + 1]):
            gb = elemGradOffsets[e] + (ip - elemIpOffsets[e]) * nn
            for p in range(elemPairOffsets[e], elemPairOffsets[e + 1]):
                i, j = pairNodeIdxs[2 * p], pairNodeIdxs[2 * p + 1]
                K = material_stiffness3(grads[3 * (gb + i):3 * (gb + i) + 3],
                                        grads[3 * (gb + j):3 * (gb + j) + 3], D, z, dvs[ip])
                S.extend(int(s) for s in blockSlots[9 * p:9 * p + 9])
                C.extend(K.ravel())
    return np.array(S, np.int64), np.array(C, float)


def linear_elastic_stiffness3_element_geometry(elemNodeCounts, elemNodeOffsets, elemPairOffsets,
                                               elemIpOffsets, elemNaturalGradOffsets, pairNodeIdxs,
                                               blockSlots, nodeDims, nodeTransforms, elemParams,
                                               elemNodePositions, naturalGrads, ipWeights, nelems):
    """addLinearElasticStiffness3ElementGeometryKernel (cudssBridgeKernels.cu:531-640):
    J = sum_n x_n (dN_n/dxi)^T, g = J^-T dN/dxi, dv = det(J) w."""
    S, C = [], []
    z = np.zeros(6)
    X = np.asarray(elemNodePositions, float).reshape(-1, 3)
    NG = np.asarray(naturalGrads, float).reshape(-1, 3)
    for e in range(nelems):
        nn, n0 = elemNodeCounts[e], elemNodeOffsets[e]
        D = linear_elastic_D(elemParams[2 * e], elemParams[2 * e + 1])
        for ip in range(elemIpOffsets[e], elemIpOffsets[e + 1]):
            gb = elemNaturalGradOffsets[e] + (ip - elemIpOffsets[e]) * nn
            J = X[n0:n0 + nn].T @ NG[gb:gb + nn]
            Jit = np.linalg.inv(J).T
            dv = np.linalg.det(J) * ipWeights[ip]
            for p in range(elemPairOffsets[e], elemPairOffsets[e + 1]):
                i, j = pairNodeIdxs[2 * p], pairNodeIdxs[2 * p + 1]
                K = material_stiffness3(Jit @ NG[gb + i], Jit @ NG[gb + j], D, z, dv)
                R = reduce_block3(K, nodeTransforms[18 * (n0 + i):18 * (n0 + i) + 18], nodeDims[n0 + i],
                                  nodeTransforms[18 * (n0 + j):18 * (n0 + j) + 18], nodeDims[n0 + j])
                _emit36(S, C, blockSlots, p, R)
    return np.array(S, np.int64), np.array(C, float)


def dilational_stiffness3_element(elemNodeCounts, elemNodeOffsets, elemPressureCounts, elemPairOffsets,
                                  elemConstraintOffsets, elemRinvOffsets, pairNodeIdxs, blockSlots,
                                  nodeDims, nodeTransforms, constraints, rinvs, nelems):
    """addDilationalStiffness3ElementKernel (cudssBridgeKernels.cu:658-717):
    K_ij = sum_ab c_{i,a} Rinv(a,b) c_{j,b}^T; elements with np <= 0 add nothing."""
    S, C = [], []
    CV = np.asarray(constraints, float).reshape(-1, 3)
    for e in range(nelems):
        npr, n0 = elemPressureCounts[e], elemNodeOffsets[e]
        if npr <= 0:
            continue
        c0, r0 = elemConstraintOffsets[e], elemRinvOffsets[e]
        Rinv = np.asarray(rinvs[r0:r0 + npr * npr], float).reshape(npr, npr)
        for p in range(elemPairOffsets[e], elemPairOffsets[e + 1]):
            i, j = pairNodeIdxs[2 * p], pairNodeIdxs[2 * p + 1]
            Ci = CV[c0 + i * npr:c0 + i * npr + npr]  # (np, 3)
            Cj = CV[c0 + j * npr:c0 + j * npr + npr]
            K = Ci.T @ Rinv @ Cj
            R = reduce_block3(K, nodeTransforms[18 * (n0 + i):18 * (n0 + i) + 18], nodeDims[n0 + i],
                              nodeTransforms[18 * (n0 + j):18 * (n0 + j) + 18], nodeDims[n0 + j])
            _emit36(S, C, blockSlots, p, R)
    return np.array(S, np.int64), np.array(C, float)
