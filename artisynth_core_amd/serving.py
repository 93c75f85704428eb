"""Serving loop: step-solve requests from many concurrent simulation instances.

- Every instance owns one solver handle (its own HIP stream) whose symbolic
  analysis and factors stay resident in HBM between requests.
- Instances are admitted by an HBM budget: the bytes an instance needs are known
  from its symbolic analysis (fronts, panel workspace, structure, vectors); an
  instance that does not fit waits until idle instances are evicted (least
  recently used first).
- Requests enter one FIFO queue; `nworkers` threads take the oldest request whose
  instance is admitted and idle, so requests that arrive while others run start as
  soon as a worker and memory free up (continuous batching), and the device work
  of different instances overlaps on their streams.  Requests of one instance
  are executed in submission order, one at a time, so an instance's results do
  not depend on what else is in the mix.
- A request is one KKT step solve with KKTSolver.factorAndSolve's policy: BiCGStab
  with the resident (stale) factor while it converges within `max_iter`, else a
  refactor and a refined direct solve (KKTSolver.java:930-1014).
"""
import collections
import ctypes
import threading
import time

import numpy as np

from ._lib import KKT_PREC_FP32_REFINE, KktInfo, lib
from ._lib import ptr as nptr


class Instance:
    def __init__(self, iid, rowOffs, colIdxs, vals0, precision):
        self.iid = iid
        self.ro = np.ascontiguousarray(rowOffs, np.int32)
        self.ci = np.ascontiguousarray(colIdxs, np.int32)
        self.vals0 = np.ascontiguousarray(vals0, np.float64)
        self.n = len(self.ro) - 1
        self.nnz = int(self.ro[-1])
        self.precision = precision
        self.h = None
        self.bytes = 0
        self.busy = False
        self.factored = False
        self.last_used = 0.0
        self.nrequests = 0


class Request:
    def __init__(self, iid, vals, b):
        self.iid = iid
        self.vals = np.ascontiguousarray(vals, np.float64)
        self.b = np.ascontiguousarray(b, np.float64)
        self.x = None
        self.kind = None
        self.iters = 0
        self.error = None
        self.t_submit = time.perf_counter()
        self.t_start = None
        self.t_done = None
        self.done = threading.Event()


class SolveServer:
    def __init__(self, budget_bytes=200 << 30, nworkers=8, precision=KKT_PREC_FP32_REFINE,
                 tol_exp=10, max_iter=30, refine_tol=1e-10):
        self.L = lib()
        self.budget = int(budget_bytes)
        self.used = 0
        self.precision = precision
        self.tol = 10.0 ** (-tol_exp)
        self.max_iter = max_iter
        self.refine_tol = refine_tol
        self.instances = {}
        self.queue = collections.deque()
        self.cv = threading.Condition()
        self.stop = False
        self.admissions = 0
        self.evictions = 0
        self.workers = [threading.Thread(target=self._worker, daemon=True) for _ in range(nworkers)]
        for w in self.workers:
            w.start()

    # ---- instances ----
    def register(self, iid, rowOffs, colIdxs, vals0, precision=None):
        with self.cv:
            self.instances[iid] = Instance(iid, rowOffs, colIdxs, vals0,
                                           self.precision if precision is None else precision)

    def _admit(self, inst):
        """Analyze and allocate an instance (called with the lock released)."""
        L = self.L
        h = L.kkt_create()
        L.kkt_set_precision(h, inst.precision)
        L.kkt_set_refine_tol(h, self.refine_tol, 10)
        rc = L.kkt_set_pattern(h, inst.n, inst.nnz, nptr(inst.ro, ctypes.c_int),
                               nptr(inst.ci, ctypes.c_int), 1)
        if rc == 0:
            rc = L.kkt_analyze_values(h, nptr(inst.vals0, ctypes.c_double))
        if rc != 0:
            msg = L.kkt_last_error(h)
            L.kkt_destroy(h)
            raise RuntimeError("admission of instance %r failed: %s" % (inst.iid, msg))
        return h

    def _front_bytes(self, h, inst):
        info = KktInfo()
        self.L.kkt_get_info(h, ctypes.byref(info))
        esz = 4 if inst.precision == KKT_PREC_FP32_REFINE else 8
        # fronts hold sum m^2 elements; front_bytes is reported for the factor precision
        fb = info.front_bytes if info.precision == inst.precision else info.front_bytes * esz // 8
        return fb + inst.n * 64 * esz + 14 * inst.n * 8 + 6 * inst.n * esz + inst.nnz * 48

    def _evict_one(self, need):
        idle = [i for i in self.instances.values() if i.h is not None and not i.busy]
        if not idle:
            return False
        victim = min(idle, key=lambda i: i.last_used)
        self.L.kkt_destroy(victim.h)
        victim.h = None
        victim.factored = False
        self.used -= victim.bytes
        victim.bytes = 0
        self.evictions += 1
        return True

    # ---- requests ----
    def submit(self, iid, vals, b):
        r = Request(iid, vals, b)
        with self.cv:
            self.queue.append(r)
            self.cv.notify()
        return r

    def _next(self):
        # oldest request whose instance is idle (keeps per-instance order)
        blocked = set()
        for idx, r in enumerate(self.queue):
            if r.iid in blocked:
                continue
            inst = self.instances[r.iid]
            if inst.busy:
                blocked.add(r.iid)
                continue
            del self.queue[idx]
            return r, inst
        return None, None

    def _worker(self):
        while True:
            with self.cv:
                while True:
                    if self.stop:
                        return
                    r, inst = self._next()
                    if r is not None:
                        inst.busy = True
                        break
                    self.cv.wait(0.05)
            try:
                self._run(r, inst)
            except Exception as e:  # reported on the request
                r.error = str(e)
            finally:
                with self.cv:
                    inst.busy = False
                    inst.last_used = time.perf_counter()
                    self.cv.notify_all()
                r.t_done = time.perf_counter()
                r.done.set()

    def _ensure_resident(self, inst):
        if inst.h is not None:
            return
        h = self._admit(inst)
        need = self._front_bytes(h, inst)
        with self.cv:
            while self.used + need > self.budget:
                if not self._evict_one(need):
                    # nothing idle to evict: wait for memory to free up
                    self.cv.wait(0.05)
            self.used += need
            inst.bytes = need
            inst.h = h
            inst.factored = False
            self.admissions += 1

    def _run(self, r, inst):
        r.t_start = time.perf_counter()
        self._ensure_resident(inst)
        L, h = self.L, inst.h
        x = np.zeros(inst.n)
        rc = -1
        if inst.factored:
            rc = L.kkt_iterative_solve(h, nptr(r.vals, ctypes.c_double), nptr(r.b, ctypes.c_double),
                                       nptr(x, ctypes.c_double), self.tol, self.max_iter)
        if rc > 0:
            r.kind, r.iters = "iterative", rc
        else:
            st = L.kkt_factor(h, nptr(r.vals, ctypes.c_double))
            if st == 0:
                st = L.kkt_solve(h, nptr(r.b, ctypes.c_double), nptr(x, ctypes.c_double))
            if st != 0:
                m = L.kkt_last_error(h)
                raise RuntimeError("solve failed (%d): %s" % (st, m.decode() if m else ""))
            inst.factored = True
            r.kind = "direct"
        inst.nrequests += 1
        r.x = x

    def shutdown(self):
        with self.cv:
            self.stop = True
            self.cv.notify_all()
        for w in self.workers:
            w.join(timeout=30)
        for inst in self.instances.values():
            if inst.h is not None:
                self.L.kkt_destroy(inst.h)
                inst.h = None


def latency_stats(requests):
    lat = np.array([(r.t_done - r.t_submit) * 1e3 for r in requests if r.t_done is not None])
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "n": int(len(lat))}
