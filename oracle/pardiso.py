"""ctypes harness for MKL PARDISO with the reference's settings -- TEST
INFRASTRUCTURE and CPU BASELINE ONLY.

Used by tests/golden/make_pardiso.py (fixtures written in the build container)
and by bench.py's cpu_baseline leg; the product path never loads it.  MKL is a
third-party library that is never vendored: it is opened with dlopen from the
host's install ($MKLROOT/lib, $MKL_RT, or /opt/conda/lib), and when it is absent
`available()` says so and nothing else runs.

What it drives (reference file:line):
  - Pardiso4 constructor defaults: pardisoinit, then iparm[0]=1, iparm[1]=3
    (parallel METIS), iparm[7]=1 refinement step (REAL_SYMMETRIC), iparm[9]=8
    (pivot perturbation 1e-8), iparm[10]=iparm[12]=1 (scaling, weighted
    matching), iparm[17]=-1, iparm[20]=1 (Bunch-Kaufman 1x1/2x2)
    (pardisoMkl.cc:82-110, 190-393; mtype=-2, maxfct=mnum=1, msglvl=0,
    pardisoMkl.cc:36-78);
  - setMatrix = phase 11 with iparm[26]=0 (pardisoMkl.cc:440-511), nnz(L) in
    iparm[17];
  - factorMatrix = phase 22 (pardisoMkl.cc:536-588): perturbed pivots
    iparm[13], inertia iparm[21] (positive) / iparm[22] (negative);
  - solveMatrix = phase 33, iparm[3]=0 (pardisoMkl.cc:650-676), refinement
    steps taken in iparm[6];
  - iterativeSolve = phase 33 with iparm[3]=10*tolExp+2 on new values, the
    stale factor as CGS preconditioner; returns iparm[19] (pardisoMkl.cc:704-750);
  - releaseMatrix = phase -1.
Indices are shifted to 1-based, as KKTSolver does for PARDISO
(KKTSolver.java:256-268); the upper triangle is passed, as for mtype=-2.
"""
import ctypes
import os

import numpy as np

_CANDIDATES = []
if os.environ.get("MKL_RT"):
    _CANDIDATES.append(os.environ["MKL_RT"])
if os.environ.get("MKLROOT"):
    _CANDIDATES += [os.path.join(os.environ["MKLROOT"], "lib", "intel64", "libmkl_rt.so"),
                    os.path.join(os.environ["MKLROOT"], "lib", "libmkl_rt.so")]
_CANDIDATES += ["/opt/conda/lib/libmkl_rt.so", "/opt/intel/oneapi/mkl/latest/lib/intel64/libmkl_rt.so",
                "libmkl_rt.so"]

_lib = None
_path = None
_why = None

REAL_SYMMETRIC_INDEF = -2


def _load():
    global _lib, _path, _why
    if _lib is not None or _why is not None:
        return _lib
    errs = []
    for c in _CANDIDATES:
        if os.path.sep in c and not os.path.exists(c):
            continue
        try:
            L = ctypes.CDLL(c)
        except OSError as e:
            errs.append("%s: %s" % (c, e))
            continue
        if not hasattr(L, "pardiso") or not hasattr(L, "pardisoinit"):
            errs.append("%s: no pardiso symbols" % c)
            continue
        i64 = ctypes.POINTER(ctypes.c_int)
        L.pardisoinit.argtypes = [ctypes.c_void_p, i64, i64]
        L.pardisoinit.restype = None
        L.pardiso.argtypes = [ctypes.c_void_p, i64, i64, i64, i64, i64, ctypes.c_void_p, i64, i64,
                              i64, i64, i64, i64, ctypes.c_void_p, ctypes.c_void_p, i64]
        L.pardiso.restype = None
        L.MKL_Set_Num_Threads.argtypes = [ctypes.c_int]
        L.MKL_Set_Num_Threads.restype = None
        L.MKL_Get_Max_Threads.argtypes = []
        L.MKL_Get_Max_Threads.restype = ctypes.c_int
        _lib, _path = L, c
        return L
    _why = "MKL PARDISO unavailable (" + ("; ".join(errs) if errs else "no libmkl_rt.so found") + ")"
    return None


def available():
    """(True, path and version) when MKL PARDISO can be loaded, else (False, reason)."""
    L = _load()
    if L is None:
        return False, _why
    buf = ctypes.create_string_buffer(256)
    L.MKL_Get_Version_String(buf, 256)
    return True, "%s (%s)" % (buf.value.decode(errors="replace").strip(), _path)


def set_threads(n):
    L = _load()
    if L is None:
        raise RuntimeError(_why)
    L.MKL_Set_Num_Threads(int(n))


def max_threads():
    L = _load()
    if L is None:
        raise RuntimeError(_why)
    return int(L.MKL_Get_Max_Threads())


class PardisoError(RuntimeError):
    pass


class Pardiso:
    """One Pardiso4-like instance on a 0-based upper CSR (rowOffs, colIdxs, vals)."""

    def __init__(self, rowOffs, colIdxs, vals, mtype=REAL_SYMMETRIC_INDEF):
        L = _load()
        if L is None:
            raise RuntimeError(_why)
        self.L = L
        ro = np.asarray(rowOffs, np.int64)
        self.n = n = len(ro) - 1
        self.nnz = int(ro[-1])
        self.ia = np.ascontiguousarray(ro + 1, np.int32)
        self.ja = np.ascontiguousarray(np.asarray(colIdxs, np.int64) + 1, np.int32)
        self.a = np.ascontiguousarray(vals, np.float64).copy()
        self.mtype = ctypes.c_int(mtype)
        self.pt = (ctypes.c_void_p * 64)()
        self.iparm = (ctypes.c_int * 64)()
        L.pardisoinit(self.pt, ctypes.byref(self.mtype), self.iparm)
        ip = self.iparm
        ip[0] = 1
        ip[1] = 3      # METIS_REORDER_PARALLEL
        ip[2] = 0
        ip[3] = 0
        ip[4] = 0
        ip[5] = 0
        ip[6] = 0
        ip[7] = 1      # one refinement step (REAL_SYMMETRIC default)
        ip[9] = 8      # pivot perturbation 1e-8
        ip[10] = 1     # scaling
        ip[12] = 1     # weighted matching
        ip[13] = 0
        ip[17] = -1
        ip[18] = -1
        ip[19] = 0
        ip[20] = 1     # Bunch-Kaufman 1x1 and 2x2 pivots
        ip[26] = 0     # no matrix checking
        self.maxfct = ctypes.c_int(1)
        self.mnum = ctypes.c_int(1)
        self.nrhs = ctypes.c_int(1)
        self.msglvl = ctypes.c_int(0)
        self.perm = (ctypes.c_int * 1)()
        self._analyzed = False
        self.phase(11)
        self._analyzed = True
        self.nnzL = int(self.iparm[17])

    def phase(self, ph, b=None, x=None, nrhs=1):
        err = ctypes.c_int(0)
        php = ctypes.c_int(ph)
        nn = ctypes.c_int(self.n)
        nr = ctypes.c_int(nrhs)
        dummy = np.zeros(1)
        bp = (b if b is not None else dummy).ctypes.data_as(ctypes.c_void_p)
        xp = (x if x is not None else dummy).ctypes.data_as(ctypes.c_void_p)
        self.L.pardiso(self.pt, ctypes.byref(self.maxfct), ctypes.byref(self.mnum), ctypes.byref(self.mtype),
                       ctypes.byref(php), ctypes.byref(nn), self.a.ctypes.data_as(ctypes.c_void_p),
                       self.ia.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                       self.ja.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), self.perm, ctypes.byref(nr),
                       self.iparm, ctypes.byref(self.msglvl), bp, xp, ctypes.byref(err))
        return err.value

    def factor(self, vals=None):
        """Phase 22 (pardisoMkl.cc:536-588)."""
        if vals is not None:
            self.a[:] = np.asarray(vals, np.float64)
        self.iparm[3] = 0
        self.iparm[26] = 0
        e = self.phase(22)
        if e != 0:
            raise PardisoError("PARDISO numeric factorization failed: %d" % e)

    def solve(self, b):
        """Phase 33 with the configured refinement (pardisoMkl.cc:650-676)."""
        b = np.ascontiguousarray(b, np.float64).copy()
        x = np.zeros(self.n)
        self.iparm[5] = 0
        self.iparm[7] = 1
        self.iparm[3] = 0
        e = self.phase(33, b, x)
        if e != 0:
            raise PardisoError("PARDISO solve failed: %d" % e)
        return x

    def iterative_solve(self, vals, b, tol_exp=10):
        """Phase 33 with iparm[3] = 10 tolExp + 2 (preconditioned CGS on the new
        values with the stale factor; pardisoMkl.cc:704-750).  Returns (iparm[19], x):
        > 0 converged in that many iterations; <= 0 the caller refactors."""
        self.a[:] = np.asarray(vals, np.float64)
        b = np.ascontiguousarray(b, np.float64).copy()
        x = np.zeros(self.n)
        self.iparm[5] = 0
        self.iparm[7] = 1
        self.iparm[26] = 0
        self.iparm[3] = 10 * tol_exp + 2
        self.phase(33, b, x)
        it = int(self.iparm[19])
        self.iparm[3] = 0
        return it, x

    def info(self):
        ip = self.iparm
        return {"nperturbed": int(ip[13]), "nnzL": self.nnzL, "npos": int(ip[21]), "nneg": int(ip[22]),
                "refine_steps": int(ip[6]), "mflops": int(ip[18])}

    def release(self):
        if self._analyzed:
            self.phase(-1)
            self._analyzed = False

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass
