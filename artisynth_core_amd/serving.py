"""Serving loop: step-solve requests from many concurrent simulation instances,
through the library's kkt_serve_* entry points (include/kkt_mi355x.h; the loop,
its worker threads, admission and eviction are C++ in csrc/serve.cpp).

- Every instance owns one solver handle (its own HIP stream) whose symbolic
  analysis and factors stay resident in HBM between requests.
- Instances are admitted by an HBM budget: after an instance's analysis, what
  its handle will hold once factored is reserved before its fronts are
  allocated, evicting idle instances least recently used first; an instance
  larger than the whole budget fails its requests.
- Requests enter one FIFO queue; `nworkers` threads take the oldest request
  whose instance is idle (continuous batching: arrivals start as soon as a
  worker and memory are free; instances' device work overlaps on their
  streams).  An instance's requests run in submission order, one at a time.
- A request is one KKT step with the iteration-limit part of
  KKTSolver.factorAndSolve's policy (KKTSolver.java:930-1014): BiCGStab with the
  resident factor while it converges within `max_iter`, else a refactor and a
  refined direct solve (the timing-based early refactor of estimateOptimalCount
  is not applied: a request's time in a shared server includes other
  instances' work).
- With `group_runners` > 0, requests of resident, factored, idle
  instances are solved together (up to `group_max` per call) by group runner
  threads, each BiCGStab step one launch for all of them; a request that does
  not converge there is refactored and solved directly by a worker.  Results
  are bitwise those of the one-instance-at-a-time path.
This module only marshals arrays and results.
"""
import ctypes

import numpy as np

from ._lib import KKT_PREC_FP32_REFINE, KktServeConfig, KktServeResult, KktServeStats, lib
from ._lib import ptr as nptr

KINDS = {1: "iterative", 2: "direct"}


class Request:
    """One submitted step; wait() blocks until it completes.  Times are ms on the
    server's clock (since its creation)."""

    def __init__(self, srv, rid, iid, x, keep):
        self._srv, self.id, self.iid, self.x = srv, rid, iid, x
        self._keep = keep  # arrays the request reads until it completes
        self.kind = self.iters = self.error = None
        self.status = None
        self.rel_residual = None
        self.t_submit = self.t_start = self.t_done = None

    def wait(self):
        if self.status is None:
            res = KktServeResult()
            self.status = self._srv.L.kkt_serve_wait(self._srv.h, self.id, ctypes.byref(res))
            self.kind = KINDS.get(res.kind)
            self.iters = res.iterations
            self.rel_residual = res.rel_residual
            self.t_submit, self.t_start, self.t_done = res.submit_ms, res.start_ms, res.done_ms
            self.error = res.message.decode(errors="replace") if self.status else None
            self._keep = None
        return self


class SolveServer:
    def __init__(self, budget_bytes=200 << 30, nworkers=12, precision=KKT_PREC_FP32_REFINE,
                 tol_exp=10, max_iter=30, refine_tol=1e-10, group_runners=None, group_max=None):
        self.L = lib()
        cfg = KktServeConfig()
        self.L.kkt_serve_default_config(ctypes.byref(cfg))
        cfg.budget_bytes = int(budget_bytes)
        cfg.nworkers = int(nworkers)
        cfg.precision = int(precision)
        cfg.tol_exp = int(tol_exp)
        cfg.max_iter = int(max_iter)
        cfg.refine_tol = float(refine_tol)
        if group_runners is not None:
            cfg.group_runners = int(group_runners)
        if group_max is not None:
            cfg.group_max = int(group_max)
        self.h = self.L.kkt_serve_create(ctypes.byref(cfg))
        if not self.h:
            raise RuntimeError("kkt_serve_create failed (bad configuration)")
        self.sizes = {}

    def _err(self, rc, op):
        m = self.L.kkt_serve_last_error(self.h)
        raise RuntimeError("serve %s failed (%d): %s" % (op, rc, m.decode() if m else ""))

    def register(self, iid, rowOffs, colIdxs, vals0, mtype=1):
        ro = np.ascontiguousarray(rowOffs, np.int32)
        ci = np.ascontiguousarray(colIdxs, np.int32)
        v0 = np.ascontiguousarray(vals0, np.float64)
        n = len(ro) - 1
        rc = self.L.kkt_serve_register(self.h, int(iid), n, int(ro[n]), nptr(ro, ctypes.c_int),
                                       nptr(ci, ctypes.c_int), nptr(v0, ctypes.c_double), int(mtype))
        if rc:
            self._err(rc, "register")
        self.sizes[iid] = n

    def submit(self, iid, vals, b):
        """Host arrays: copied by the server at submission."""
        v = np.ascontiguousarray(vals, np.float64)
        bb = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.sizes[iid])
        rid = self.L.kkt_serve_submit(self.h, int(iid), nptr(v, ctypes.c_double), nptr(bb, ctypes.c_double),
                                      nptr(x, ctypes.c_double))
        if rid <= 0:
            self._err(rid, "submit")
        return Request(self, rid, iid, x, None)

    def submit_device(self, iid, d_vals, d_b, d_x, keep=None):
        """Device pointers, read in place until the request completes."""
        rid = self.L.kkt_serve_submit_device(self.h, int(iid), d_vals, d_b, d_x)
        if rid <= 0:
            self._err(rid, "submit")
        return Request(self, rid, iid, None, keep)

    def stats(self):
        st = KktServeStats()
        self.L.kkt_serve_get_stats(self.h, ctypes.byref(st))
        return {f: getattr(st, f) for f, _ in st._fields_}

    def shutdown(self):
        if getattr(self, "h", None):
            self.L.kkt_serve_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.shutdown()
        except Exception:
            pass


def latency_stats(requests):
    lat = np.array([r.t_done - r.t_submit for r in requests if r.t_done is not None])
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "n": int(len(lat))}
