"""Benchmark: KKT solves/s at <= 1e-8 relative residual on the constrained hex beam.

Workload (one GPU): the 87,567-DOF neo-Hookean hex beam 100x16x16 with 867
attachment constraint rows (n = 88,434), a 20-step ConstrainedBackwardEuler
sequence (h = 0.005, seeded tip loads).  One "step" = one KKT solve of that
sequence with the reference's hybrid policy (KKTSolver.factorAndSolve): the
factor of an earlier step is reused by BiCGStab (stale-factor preconditioner,
tolExp = 10) while estimateOptimalCount says the iterative steps still pay
against a refactor (from the measured step times) and BiCGStab converges within
30 iterations; otherwise the matrix is refactored and solved directly with
refinement.  The FP32-MFMA + FP64-refinement path is the headline `value`; the
same sequence with FP64 fronts is timed in the same run.

Inputs (all 20 value arrays and right-hand sides) are resident in HBM before the
timed region; the library is driven through its device-pointer entry points.

After the headline sequence the default run also measures the two other
workloads in their own timed regions and adds them to its line: "batch" (256
uneven FP64 systems, one batched factor + solve against one handle per system)
and "serve" (64 instances, 1k-90k DOF, Poisson requests at 2,000/s through the
serving loop, 200 GB budget), plus "serve_budget_12gb", the same stream under a
12 GB budget (below the instances' resident total, so admission evicts and
re-admits); --no-mix skips them.

  python bench.py [--gpus 1] [--steps 20] [--warmup 3]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAKS = {"hbm_GBs": 8000.0, "fp64_TFLOPs": 78.6, "fp32_TFLOPs": 157.3}
TIMING_CLASSES = ["assemble", "panel_diag", "panel_trsm", "panel_update", "solve_forward",
                  "solve_backward", "spmv", "vector"]


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def _structure(L, h, with_parent=False):
    """Supernode sizes k, r and tree levels of the handle's analysis (and the
    assembly-tree parents)."""
    from artisynth_core_amd._lib import KktInfo
    from artisynth_core_amd._lib import ptr as nptr
    inf = KktInfo()
    L.kkt_get_info(h, ctypes.byref(inf))
    ns = inf.nsuper
    sup = np.zeros(ns + 1, np.int32)
    rp = np.zeros(ns + 1, np.int64)
    par = np.zeros(ns, np.int32)
    L.kkt_get_symbolic(h, None, nptr(sup, ctypes.c_int), nptr(par, ctypes.c_int), nptr(rp, ctypes.c_longlong), None)
    k = np.diff(sup).astype(np.int64)
    r = np.diff(rp).astype(np.int64)
    lev = np.zeros(ns, np.int64)
    for s_ in range(ns):  # children precede parents; level = longest path from a leaf
        if par[s_] >= 0:
            lev[par[s_]] = max(lev[par[s_]], lev[s_] + 1)
    assert lev.max() + 1 == inf.nlevels
    if with_parent:
        return k, r, lev, inf, par
    return k, r, lev, inf


def bwd_sweep_bytes(L, h, w, wg=None):
    """Algorithmic HBM bytes of one backward sweep (all k_bwd_cols launches of a
    triangular solve): per supernode the lower copy of G once (k(k-1)/2 + k r
    entries, wg bytes each: the front precision, or 2 for the half-precision
    copy), its backward vector once (k + r entries), x written for its k pivots
    and pushed to the r row entries it fills in the descendants (w bytes per
    vector entry), and the indices (pivot permutation and push ranges, k each;
    push positions, r)."""
    wg = w if wg is None else wg
    k, r, lev, inf = _structure(L, h)
    k, r = k.astype(np.float64), r.astype(np.float64)
    return float(np.sum((k * (k - 1) / 2 + k * r) * wg + (2 * k + 2 * r) * w + (r + 2 * k) * 4))


def fwd_sweep_bytes(L, h, w, fwd_epl, forms, groups, wg=None):
    """Algorithmic HBM bytes of one forward sweep (all k_fwd_level launches of a
    triangular solve), per supernode (k pivots, r rows below): the upper copy of
    G once (k(k-1)/2 entries of X, unit diagonal implicit, and k r of Z), the
    right-hand side at the pivots (k) and y written for them (k), its r update
    entries written once and read once by the parent, the pivot permutation (k)
    and the rows' slot positions in the parent (r), 4 bytes each; w = bytes per
    vector entry, wg per entry of G (the front precision, or 2 for the
    half-precision copy).
    Returned separately, the layout's extra reads: the zero entries of each dense
    update-slot block (nch x m entries read, of which only the children's rows
    were written) and, in the one-launch form, every further row block of a large
    front gathering the front's pivot part again (k positions: right-hand side,
    nch slots and the permutation; a block covers `groups` row groups of its
    level); in the two-launch form, the gathered pivot part written once and read
    by every row block."""
    wg = w if wg is None else wg
    k, r, lev, inf, par = _structure(L, h, with_parent=True)
    m = k + r
    kf, rf, mf = k.astype(np.float64), r.astype(np.float64), m.astype(np.float64)
    nch = np.zeros(len(k), np.float64)
    rch = np.zeros(len(k), np.float64)  # update entries the children write into the block
    for s_ in range(len(k)):
        if par[s_] >= 0:
            nch[par[s_]] += 1
            rch[par[s_]] += rf[s_]
    alg = float(np.sum((kf * (kf - 1) / 2 + kf * rf) * wg + (2 * kf + 2 * rf) * w + (kf + rf) * 4.0))
    red = float(np.sum((nch * mf - rch) * w))
    for s_ in range(len(k)):
        if k[s_] <= 32 and m[s_] <= 256:  # small fronts: gather and rows in one workgroup
            continue
        L_ = lev[s_]
        lg = 0
        while lg < 6 and fwd_epl[L_] * (1 << lg) < k[s_]:
            lg += 1
        if forms[L_] != 1:
            nblk = (m[s_] + 4 * (64 >> lg) - 1) // (4 * (64 >> lg))
            red += (nblk + 1) * kf[s_] * w
            continue
        rg = int(groups[L_])
        while rg > 1 and rg * 4 * (64 >> lg) > 1024:
            rg //= 2
        rows_per_block = 4 * (64 >> lg) * rg
        nblk = (m[s_] + rows_per_block - 1) // rows_per_block
        red += (nblk - 1) * kf[s_] * ((1 + nch[s_]) * w + 4.0)
    return alg, red


def spmv_bytes(n, nnz_upper, mode_resid=True):
    """SpMV of refinement / BiCGStab on the full symmetric CSR the library keeps:
    values and column indices of every stored entry (nnz_full = 2 nnz_upper - n
    with the diagonal once), row offsets, x read, y written, b read in the
    residual form; and the minimum of upper storage for comparison."""
    nnz_full = 2 * nnz_upper - n
    full = nnz_full * 12 + (n + 1) * 4 + (3 if mode_resid else 2) * n * 8
    upper = nnz_upper * 12 + (n + 1) * 4 + 2 * n * 8
    return float(full), float(upper), nnz_full


class FactorAndSolvePolicy:
    """KKTSolver.factorAndSolve's choice between the stale-factor iterative solve
    and a factor + solve (KKTSolver.java:889-1014), on wall-clock step times:
    after a factor, iterative steps are tried while their count stays below
    estimateOptimalCount (the step count that minimises the average time per step
    when iterative steps grow linearly more expensive after a factor); a failed
    iterative attempt (more than max_iter iterations) also leads to a factor.
    Direct times accumulate over the whole sequence, iterative times since the
    last factor, as in the reference."""

    def __init__(self):
        self.direct_ms, self.direct_cnt = 0.0, 0
        self.iter_ms, self.first_iter_ms, self.iter_cnt = 0.0, 0.0, 0

    def estimate_optimal_count(self):
        ft = self.direct_ms / self.direct_cnt
        it0 = self.first_iter_ms
        it_avg = self.iter_ms / self.iter_cnt
        if it_avg >= ft:
            return self.iter_cnt + 1
        dt = 2.0 * (it_avg - it0) / self.iter_cnt
        if dt <= 0:
            return self.iter_cnt + 19
        return int(np.ceil(np.sqrt(2.0 * (ft - it0 + dt) / dt)))

    def try_iterative(self):
        return self.direct_cnt > 0 and (self.iter_cnt == 0 or self.iter_cnt + 1 < self.estimate_optimal_count())

    def iterative_done(self, ms):
        self.iter_ms += ms
        if self.iter_cnt == 0:
            self.first_iter_ms = self.iter_ms
        self.iter_cnt += 1

    def direct_done(self, ms):
        self.direct_ms += ms
        self.direct_cnt += 1
        self.iter_cnt, self.iter_ms = 0, 0.0


def run_sequence(L, h, vals_d, b_d, x_d, steps, nsys, tol_exp=10, max_iter=30):
    """Runs `steps` KKT solves over the systems 0..nsys-1 (cyclic) with
    KKTSolver.factorAndSolve's policy (FactorAndSolvePolicy, a fresh one per
    sequence); returns per-step latencies (ms), kinds and iteration counts."""
    import torch
    lat, kinds, iters = [], [], []
    tol = 10.0 ** (-tol_exp)
    pol = FactorAndSolvePolicy()
    for i in range(steps):
        k = i % nsys
        vp = vals_d[k].data_ptr()
        bp = b_d[k].data_ptr()
        xp = x_d[i % len(x_d)].data_ptr()
        t0 = time.perf_counter()
        rc = -1
        if pol.try_iterative():
            rc = L.kkt_iterative_solve_device(h, vp, bp, xp, tol, max_iter)  # returns after its last copy
        if rc > 0:
            pol.iterative_done((time.perf_counter() - t0) * 1e3)
            kinds.append("iterative")
            iters.append(rc)
        else:
            t1 = time.perf_counter()
            st = L.kkt_factor_device(h, vp)
            if st != 0:
                raise RuntimeError("factor failed: %s" % L.kkt_last_error(h))
            st = L.kkt_solve_device(h, bp, xp)
            if st != 0:
                raise RuntimeError("solve failed: %s" % L.kkt_last_error(h))
            torch.cuda.synchronize()
            pol.direct_done((time.perf_counter() - t1) * 1e3)
            kinds.append("direct")
            iters.append(0)
        lat.append((time.perf_counter() - t0) * 1e3)
    torch.cuda.synchronize()
    return lat, kinds, iters


def _timing(L, h):
    ms = (ctypes.c_double * 8)()
    cnt = (ctypes.c_longlong * 8)()
    uf = ctypes.c_double()
    L.kkt_timing_get(h, ms, cnt, ctypes.byref(uf))
    return [ms[c] for c in range(8)], [int(cnt[c]) for c in range(8)], uf.value


def kernel_pass(L, h, vals_ptr, b_ptr, x_ptr):
    """After the timed region: per-launch HIP-event times (events on the solver's
    stream, eager launches) of one factor's kernel classes; the average kernel
    duration of the forward sweep, the backward sweep and the residual SpMV from
    20 graph replays each (one sweep / one SpMV per replay); and the factor's wall
    time as it runs in the sequence (graph replay, no events)."""
    import torch
    from artisynth_core_amd._lib import KktInfo
    # sweep and SpMV kernel durations: graph replays between two HIP events on
    # the solver's stream (no per-launch events, so they match a kernel trace).
    # The SpMV first, on the sequence's last values and factor (20 us per launch
    # at config L); replayed right after the refactor below, on step 0's values,
    # the same launch measured 26-27 us, for a reason not identified (the values
    # hold no subnormals); the iterative steps that dominate the sequence run on
    # the later values
    kt = {}

    def time_kernels(what, key):
        ms = ctypes.c_double()
        cnt = ctypes.c_int()
        assert L.kkt_time_kernels(h, what, 20, ctypes.byref(ms), ctypes.byref(cnt)) == 0, L.kkt_last_error(h)
        kt[key] = (ms.value, cnt.value)
    time_kernels(2, "spmv")
    L.kkt_timing_reset(h, 0x8F)  # factor classes + vector ops
    assert L.kkt_factor_device(h, vals_ptr) == 0, L.kkt_last_error(h)
    fms, fcnt, uflops = _timing(L, h)
    L.kkt_timing_reset(h, 0)
    time_kernels(0, "fwd")
    time_kernels(1, "bwd")
    inf = KktInfo()
    L.kkt_get_info(h, ctypes.byref(inf))
    if inf.precond_half:  # the iterative solve's sweeps on the half-precision copy
        time_kernels(3, "fwd16")
        time_kernels(4, "bwd16")
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        assert L.kkt_factor_device(h, vals_ptr) == 0
    torch.cuda.synchronize()
    factor_ms = (time.perf_counter() - t0) * 1e3 / reps
    return {"factor_kernel_ms": {TIMING_CLASSES[c]: fms[c] for c in (0, 1, 2, 3, 7)},
            "factor_kernel_launches": {TIMING_CLASSES[c]: fcnt[c] for c in (0, 1, 2, 3, 7)},
            "update_flops": uflops, "factor_ms": factor_ms,
            "fwd_ms": kt["fwd"][0] * kt["fwd"][1], "fwd_launches": kt["fwd"][1],
            "bwd_ms": kt["bwd"][0] * kt["bwd"][1], "bwd_launches": kt["bwd"][1],
            "spmv_ms": kt["spmv"][0], "spmv_launches": 1, "sweeps": 1,
            "fwd16_ms": kt["fwd16"][0] * kt["fwd16"][1] if "fwd16" in kt else 0.0,
            "fwd16_launches": kt["fwd16"][1] if "fwd16" in kt else 0,
            "bwd16_ms": kt["bwd16"][0] * kt["bwd16"][1] if "bwd16" in kt else 0.0,
            "bwd16_launches": kt["bwd16"][1] if "bwd16" in kt else 0}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nsys", type=int, default=20, help="systems in the step sequence")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mix", action="store_true",
                    help="bench: skip the 256-system batch and the 64-instance serving runs that the "
                         "default run adds to its line (batch / serve objects)")
    ap.add_argument("--grid", default="100x16x16")
    ap.add_argument("--mode", default="bench",
                    choices=["bench", "pmc", "batch", "serve", "precision", "multi", "stream", "levels", "group"],
                    help="pmc: one FP32 factor + solve and a 256 MiB copy, for counter runs; "
                         "batch: the 256-system mix, one batched launch vs per-system launches; "
                         "precision: the ~210k-DOF Mooney-Rivlin tet box, direct FP32-MFMA+refine "
                         "vs FP64 solves; multi: block solves of --nrhs right-hand sides on the "
                         "beam's first KKT, batched vs one at a time; stream: streaming read and copy "
                         "rates of one launch over 4 MB - 1 GB with cold caches (the achievable HBM ceiling); "
                         "levels: per precision one factor of the beam's first KKT, the solve-shape autotuner and "
                         "four stale-factor iterative solves on the second (graphs on), for a kernel trace of "
                         "the sweeps level by level; group: one stale-factor iterative step of every "
                         "serving instance, as one group solve vs one handle at a time vs --workers threads")
    ap.add_argument("--nrhs", type=int, default=64, help="multi: right-hand sides per block solve")
    ap.add_argument("--tet", default="121x24x24", help="precision: Delaunay lattice points")
    ap.add_argument("--iterative-half", type=int, default=0, choices=[0, 1],
                    help="1: the FP32 path's iterative preconditioner reads the half-precision copy of the "
                         "solve arrays (kkt_set_iterative_half); 0: the FP32 arrays")
    ap.add_argument("--max-super", type=int, default=2048,
                    help="widest supernode of the analysis (64..2048; kkt_set_max_supernode)")
    ap.add_argument("--oracle", action="store_true",
                    help="precision: also run the CPU oracle on step 0 (minutes at full size)")
    ap.add_argument("--nbatch", type=int, default=256)
    ap.add_argument("--ninst", type=int, default=64)
    ap.add_argument("--nreq", type=int, default=2000)
    ap.add_argument("--rate", type=float, default=2000.0, help="serving: offered requests/s")
    ap.add_argument("--workers", type=int, default=12)
    ap.add_argument("--group-runners", type=int, default=2,
                    help="serve: group runner threads (iterative steps of idle factored instances solved "
                         "together); 0: every request on a worker")
    ap.add_argument("--group-max", type=int, default=64, help="serve: requests per group solve")
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--eps", type=float, default=5e-3,
                    help="serving: per-step relative drift of every matrix entry (random walk)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested); run with --gpus 1" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from artisynth_core_amd import KKT_PREC_FP32_REFINE, KKT_PREC_FP64
    from artisynth_core_amd import generators as G
    from artisynth_core_amd._lib import KktInfo, lib
    from artisynth_core_amd._lib import ptr as nptr

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    nx, ny, nz = (int(v) for v in args.grid.split("x"))
    if args.mode == "pmc":
        return pmc_run(nx, ny, nz)
    if args.mode == "batch":
        return batch_run(args.nbatch, args.steps, args.warmup)
    if args.mode == "serve":
        return serve_run(args.ninst, args.nreq, args.rate, args.workers, args.budget_gb, args.eps,
                         group_runners=args.group_runners, group_max=args.group_max)
    if args.mode == "precision":
        return precision_run(tuple(int(v) for v in args.tet.split("x")), min(args.nsys, 4),
                             args.warmup, args.oracle, args.max_super)
    if args.mode == "multi":
        return multi_run(nx, ny, nz, args.nrhs, max(args.steps // 4, 2), args.warmup)
    if args.mode == "stream":
        return stream_run()
    if args.mode == "group":
        return group_run(args.ninst, args.eps, args.steps, args.workers)
    if args.mode == "levels":
        return levels_run(nx, ny, nz, args.iterative_half)
    L = lib()
    # setup (untimed): integrate the ConstrainedBackwardEuler sequence with the
    # FP64 GPU path, so each step's system is assembled at the state the previous
    # steps produced
    t0 = time.time()
    integ = {}

    def fp64_solve(beam, vals, b):
        if "h" not in integ:
            hh = L.kkt_create()
            L.kkt_set_precision(hh, KKT_PREC_FP64)
            L.kkt_set_ir_steps(hh, 1)
            ro_ = np.ascontiguousarray(beam.rowOffs, np.int32)
            ci_ = np.ascontiguousarray(beam.colIdxs, np.int32)
            assert L.kkt_set_pattern(hh, beam.n, beam.nnz, nptr(ro_, ctypes.c_int), nptr(ci_, ctypes.c_int), 1) == 0
            v_ = np.ascontiguousarray(vals)
            assert L.kkt_analyze_values(hh, nptr(v_, ctypes.c_double)) == 0, L.kkt_last_error(hh)
            integ["h"] = hh
        hh = integ["h"]
        v_ = np.ascontiguousarray(vals)
        b_ = np.ascontiguousarray(b)
        x_ = np.zeros(beam.n)
        assert L.kkt_factor(hh, nptr(v_, ctypes.c_double)) == 0, L.kkt_last_error(hh)
        assert L.kkt_solve(hh, nptr(b_, ctypes.c_double), nptr(x_, ctypes.c_double)) == 0, L.kkt_last_error(hh)
        return x_

    beam, V, B, tipdefl = G.large_beam_dynamics(fp64_solve, nsteps=args.nsys, nx=nx, ny=ny, nz=nz,
                                                width=0.16 * ny / 16)
    L.kkt_destroy(integ.pop("h"))
    log("integrated %d ConstrainedBackwardEuler steps n=%d nnz=%d in %.1fs (tip displacement %.1f mm)" % (
        len(V), beam.n, beam.nnz, time.time() - t0, 1e3 * tipdefl[-1]))
    n, nnz = beam.n, beam.nnz
    ro = np.ascontiguousarray(beam.rowOffs, np.int32)
    ci = np.ascontiguousarray(beam.colIdxs, np.int32)
    vals_d = [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in V]
    b_d = [torch.from_numpy(np.ascontiguousarray(b)).to(dev) for b in B]
    x_d = [torch.zeros(n, dtype=torch.float64, device=dev) for _ in range(args.steps)]
    torch.cuda.synchronize()

    results = {}
    for prec, name in ((KKT_PREC_FP32_REFINE, "fp32_mfma_refine"), (KKT_PREC_FP64, "fp64")):
        h = L.kkt_create()
        L.kkt_set_precision(h, prec)
        L.kkt_set_refine_tol(h, 1e-10, 10)
        L.kkt_set_iterative_half(h, args.iterative_half)
        if prec == KKT_PREC_FP64:
            L.kkt_set_ir_steps(h, 1)  # one refinement step (PARDISO iparm[7]=1)
        t0 = time.time()
        rc = L.kkt_set_pattern(h, n, nnz, nptr(ro, ctypes.c_int), nptr(ci, ctypes.c_int), 1)
        assert rc == 0, L.kkt_last_error(h)
        v0 = np.ascontiguousarray(V[0])
        rc = L.kkt_analyze_values(h, nptr(v0, ctypes.c_double))
        assert rc == 0, L.kkt_last_error(h)
        t_an = time.time() - t0
        # warmup (untimed) and the solve-shape autotuner on the warmup's factor, then
        # the timed sequence from a fresh factor
        run_sequence(L, h, vals_d, b_d, x_d, args.warmup, len(V))
        inf0 = KktInfo()
        L.kkt_get_info(h, ctypes.byref(inf0))
        t_tune = time.time()
        rc = L.kkt_autotune_solve(h, 3)
        assert rc == 0, L.kkt_last_error(h)
        t_tune = time.time() - t_tune
        shp_f = np.zeros(inf0.nlevels, np.int32)
        shp_b = np.zeros(inf0.nlevels, np.int32)
        shp_n = np.zeros(inf0.nlevels, np.int32)
        L.kkt_get_solve_shapes(h, nptr(shp_f, ctypes.c_int), nptr(shp_b, ctypes.c_int))
        L.kkt_get_forward_forms(h, nptr(shp_n, ctypes.c_int))
        shp_g = np.zeros(inf0.nlevels, np.int32)
        L.kkt_get_forward_row_groups(h, nptr(shp_g, ctypes.c_int))
        log("%s: autotuned solve shapes (elements per lane, levels 0..%d) fwd %s bwd %s, forward launches "
            "per level %s, row groups %s, in %.2f s" % (name, inf0.nlevels - 1, shp_f.tolist(), shp_b.tolist(),
                                                       shp_n.tolist(), shp_g.tolist(), t_tune))
        # the timed region: no kernel timing (factors and solves replay their graphs)
        L.kkt_timing_reset(h, 0)
        run_sequence(L, h, vals_d, b_d, x_d, 1, len(V))  # graphs captured for the re-planned solve
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        lat, kinds, iters = run_sequence(L, h, vals_d, b_d, x_d, args.steps, len(V))
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        inf = KktInfo()
        L.kkt_get_info(h, ctypes.byref(inf))
        # after the timed region: per-launch kernel times on the solver's stream
        x_t = torch.zeros(n, dtype=torch.float64, device=dev)  # not one of the verified solutions
        kp = kernel_pass(L, h, vals_d[0].data_ptr(), b_d[0].data_ptr(), x_t.data_ptr())
        w = 4 if prec == KKT_PREC_FP32_REFINE else 8
        fwd_bytes, fwd_red = fwd_sweep_bytes(L, h, w, shp_f, shp_n, shp_g)
        bwd_bytes = bwd_sweep_bytes(L, h, w)
        fwd16_bytes, fwd16_red = fwd_sweep_bytes(L, h, w, shp_f, shp_n, shp_g, wg=2)
        bwd16_bytes = bwd_sweep_bytes(L, h, w, wg=2)
        # verify every step's solution (host, after timing)
        import scipy.sparse as sp
        res = []
        for i in range(args.steps):
            k = i % len(V)
            x = x_d[i].cpu().numpy()
            U = sp.csr_matrix((V[k], ci, ro), shape=(n, n))
            A = U + sp.triu(U, 1).T
            res.append(float(np.linalg.norm(A @ x - B[k]) / np.linalg.norm(B[k])))
        results[name] = {
            "solves_per_s": args.steps / elapsed, "elapsed_s": elapsed, "lat_ms": lat,
            "kinds": kinds, "iters": iters, "rel_residual": res, "analyze_s": t_an,
            "kp": kp, "info": inf.as_dict(), "fwd_sweep_bytes": fwd_bytes, "fwd_redundant_bytes": fwd_red,
            "bwd_sweep_bytes": bwd_bytes, "nlevels": inf.nlevels,
            "fwd16_sweep_bytes": fwd16_bytes, "fwd16_redundant_bytes": fwd16_red, "bwd16_sweep_bytes": bwd16_bytes,
            "solve_shapes": {"fwd": shp_f.tolist(), "bwd": shp_b.tolist(), "fwd_launches": shp_n.tolist(),
                             "fwd_row_groups": shp_g.tolist(), "tune_s": t_tune},
            "x": [x_d[i].cpu().numpy() for i in range(min(args.steps, len(V)))],
        }
        log("%s: %.2f solves/s, p50 %.2f ms, p99 %.2f ms, kinds %s, iters %s, max rel resid %.2e" % (
            name, args.steps / elapsed, np.percentile(lat, 50), np.percentile(lat, 99),
            "".join("D" if k == "direct" else "i" for k in kinds), iters, max(res)))
        L.kkt_destroy(h)

    r32, r64 = results["fp32_mfma_refine"], results["fp64"]
    fwd = [float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(r32["x"], r64["x"])]
    kp = r32["kp"]
    # dominant kernel of the FP32 sequence: the forward triangular sweep
    # (k_fwd_level<float>, HBM/latency bound); algorithmic bytes per launch = the
    # sweep's bytes / its launches, average launch time from the HIP events
    def hbm_roof(kernel, sweep_bytes, ms, launches, sweeps, traffic_key):
        if launches <= 0 or ms <= 0:
            return None
        bpl = sweep_bytes * sweeps / launches
        avg = ms / launches
        ach = bpl / (avg * 1e-3) / 1e9
        return {"bound": "hbm", "kernel": kernel, "achieved": ach, "peak": PEAKS["hbm_GBs"], "unit": "GB/s",
                "frac": ach / PEAKS["hbm_GBs"],
                "traffic": profile_traffic(traffic_key, r32["info"]["nlevels"], launches / sweeps),
                "avg_launch_ms": avg, "launches": launches, "bytes_per_launch": bpl}
    # with the half-precision copy, the iterative steps' preconditioner sweeps
    # (most of the sequence) read it: the headline entry is that forward sweep
    # and the FP32 arrays' sweeps (direct solves, refinement) are reported beside
    roof_fwd32 = hbm_roof("k_fwd_level<float>", r32["fwd_sweep_bytes"], kp["fwd_ms"], kp["fwd_launches"],
                          kp["sweeps"], "k_fwd_level<float>")
    roof_fwd32["redundant_gather_bytes_per_launch"] = r32["fwd_redundant_bytes"] * kp["sweeps"] / kp["fwd_launches"]
    roof_bwd32 = hbm_roof("k_bwd_cols<float>", r32["bwd_sweep_bytes"], kp["bwd_ms"], kp["bwd_launches"],
                          kp["sweeps"], "k_bwd_cols<float>")
    roof_fwd16 = roof_bwd16 = None
    if kp["fwd16_launches"]:
        roof_fwd16 = hbm_roof("k_fwd_level<float, _Float16>", r32["fwd16_sweep_bytes"], kp["fwd16_ms"],
                              kp["fwd16_launches"], kp["sweeps"], "k_fwd_level<float, _Float16>")
        roof_fwd16["redundant_gather_bytes_per_launch"] = (r32["fwd16_redundant_bytes"] * kp["sweeps"] /
                                                           kp["fwd16_launches"])
        roof_bwd16 = hbm_roof("k_bwd_cols<float, _Float16>", r32["bwd16_sweep_bytes"], kp["bwd16_ms"],
                              kp["bwd16_launches"], kp["sweeps"], "k_bwd_cols<float, _Float16>")
    roofline = roof_fwd16 or roof_fwd32
    roof_bwd = roof_bwd16 or roof_bwd32
    sp_full, sp_upper, nnz_full = spmv_bytes(n, nnz)
    roof_spmv = None
    if kp["spmv_launches"] > 0:
        avg = kp["spmv_ms"] / kp["spmv_launches"]
        ach = sp_full / (avg * 1e-3) / 1e9
        roof_spmv = {"bound": "hbm", "kernel": "k_spmv", "achieved": ach, "peak": PEAKS["hbm_GBs"], "unit": "GB/s",
                     "frac": ach / PEAKS["hbm_GBs"], "avg_launch_ms": avg, "launches": kp["spmv_launches"],
                     "bytes_per_launch": sp_full, "upper_storage_min_bytes": sp_upper,
                     "frac_of_upper_min_rate": sp_upper / (avg * 1e-3) / 1e9 / PEAKS["hbm_GBs"]}
    # the factor: whole-factor rate against the FP32 MFMA peak, and the Schur update
    fl = r32["info"]["factor_flops"]
    fac_tf = fl / (kp["factor_ms"] * 1e-3) / 1e12
    roof_factor = {"bound": "mfma", "kernel": "whole FP32 factor (graph replay)", "achieved": fac_tf,
                   "peak": PEAKS["fp32_TFLOPs"], "unit": "TFLOP/s", "frac": fac_tf / PEAKS["fp32_TFLOPs"],
                   "factor_ms": kp["factor_ms"], "flops": fl}
    upd_ms = kp["factor_kernel_ms"]["panel_update"]
    upd_n = max(kp["factor_kernel_launches"]["panel_update"], 1)
    upd_tf = kp["update_flops"] / (upd_ms * 1e-3) / 1e12 if upd_ms > 0 else 0.0
    roof_update = {"bound": "mfma", "kernel": "k_panel_update<float>", "achieved": upd_tf,
                   "peak": PEAKS["fp32_TFLOPs"], "unit": "TFLOP/s", "frac": upd_tf / PEAKS["fp32_TFLOPs"],
                   "avg_launch_ms": upd_ms / upd_n, "launches": upd_n, "flops_per_launch": kp["update_flops"] / upd_n}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(beam, V, B, r64["x"], r32["x"])

    lat = np.array(r32["lat_ms"])

    def iter_ms(r):
        # wall time of the steps BiCGStab finished, per iteration (each step's
        # start-up solve and residual included): the cost of an iteration apart
        # from how many iterations the rounding of a run leads to
        it = [(t, c) for t, k, c in zip(r["lat_ms"], r["kinds"], r["iters"]) if k == "iterative"]
        return sum(t for t, c in it) / max(1, sum(c for t, c in it))

    line = {
        "metric": "KKT solves/s at <=1e-8 rel resid, FP32-MFMA+refine vs FP64 LDLT; p50/p99 lat",
        "value": r32["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * r32["elapsed_s"] / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64",
        "data": "synthetic (seeded neo-Hookean hex beam; the %d-step ConstrainedBackwardEuler run integrated "
                "during setup with the FP64 GPU solve)" % len(V),
        "config": {"workload": "constrained hex beam %s, n=%d (87,567 DOF + 867 constraints), %d integrated "
                               "ConstrainedBackwardEuler steps (h=0.005, gravity + seeded tip loads), hybrid "
                               "stale-factor BiCGStab (<=30 it)" % (args.grid, n, len(V)),
                   "n": n, "nnz_upper": nnz, "precision": "fp32 fronts on MFMA + fp64 refinement"},
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "max_rel_residual": max(r32["rel_residual"]),
        "fp64_value": r64["solves_per_s"], "speedup_vs_fp64": r32["solves_per_s"] / r64["solves_per_s"],
        "fp64_p50_ms": float(np.percentile(r64["lat_ms"], 50)),
        "fp64_p99_ms": float(np.percentile(r64["lat_ms"], 99)),
        "fp64_max_rel_residual": max(r64["rel_residual"]),
        "max_fwd_err_vs_fp64": max(fwd),
        "step_kinds": "".join("D" if k == "direct" else "i" for k in r32["kinds"]),
        "bicgstab_iters": r32["iters"], "fp64_bicgstab_iters": r64["iters"],
        "ms_per_bicgstab_iter": iter_ms(r32), "fp64_ms_per_bicgstab_iter": iter_ms(r64),
        "step_ms": [round(t, 3) for t in r32["lat_ms"]], "fp64_step_ms": [round(t, 3) for t in r64["lat_ms"]],
        "tip_displacement_mm": [round(1e3 * t, 3) for t in tipdefl],
        "factor_gflop": fl / 1e9, "nnzL": r32["info"]["nnzL"],
        "n2x2": r32["info"]["n2x2"], "pivoted_blocks": r32["info"]["pivoted_blocks"],
        "solve_shapes_fp32": r32["solve_shapes"],
        "factor_ms_fp32": kp["factor_ms"], "factor_ms_fp64": r64["kp"]["factor_ms"],
        "kernel_ms_fp32": r32["kp"]["factor_kernel_ms"], "kernel_ms_fp64": r64["kp"]["factor_kernel_ms"],
        "roofline": roofline, "roofline_bwd": roof_bwd, "roofline_spmv": roof_spmv,
        "roofline_fwd_fp32": roof_fwd32 if roof_fwd16 else None, "roofline_bwd_fp32": roof_bwd32 if roof_fwd16 else None,
        "precond_half": bool(r32["info"]["precond_half"]),
        "roofline_factor": roof_factor, "roofline_factor_update": roof_update, "cpu_baseline": cpu,
    }
    if not args.no_mix:
        # the batching and serving workloads, measured in the same run (after the
        # headline sequence; their own timed regions)
        bl = batch_run(args.nbatch, args.steps, args.warmup, emit=False)
        line["batch"] = {k: bl[k] for k in ("nsys", "n_total", "batched_ms", "per_system_ms", "speedup",
                                            "batched_systems_per_s", "max_rel_residual",
                                            "max_diff_alone_vs_batch", "dtype")}
        keys = ("ninst", "nreq", "rate_target", "offered_rate_achieved", "workers", "budget_gb",
                "throughput_req_s", "p50_ms", "p99_ms", "errors", "direct", "iterative",
                "max_rel_residual_sample", "resident_gb", "evictions", "admissions_timed", "queue_max",
                "concurrency", "total_dof", "inputs", "group_runners", "group_calls", "group_requests",
                "mean_iterations", "releases")
        sl = serve_run(args.ninst, args.nreq, args.rate, args.workers, args.budget_gb, args.eps, emit=False,
                       group_runners=args.group_runners, group_max=args.group_max)
        line["serve"] = {k: sl[k] for k in keys}
        # the same stream with a budget below the instances' resident total, so
        # admission evicts idle instances and re-admits them (refactor) under load
        sl = serve_run(args.ninst, args.nreq, args.rate, args.workers, 12.0, args.eps, emit=False,
                       group_runners=args.group_runners, group_max=args.group_max)
        line["serve_budget_12gb"] = {k: sl[k] for k in keys}
    print(json.dumps(line), flush=True)


def profile_traffic(kernel, nlevels, launches_per_sweep):
    """HBM bytes of `kernel` per launch of this run's plan, from the committed
    counter profile (profiles/traffic_hexbeam_100x16x16.json: FETCH_SIZE +
    WRITE_SIZE per dispatch from two rocprofv3 --pmc runs of `bench.py --mode
    pmc`, each scaled by the factor the same counter needs on a 256 MiB device
    copy).  The profile runs the default solve plan (one launch per tree level),
    so its bytes per sweep (per-dispatch bytes x dispatches / sweeps, one
    k_bwd_cols dispatch per level and sweep) are spread over this run's
    launches per sweep.  None if absent."""
    tf = os.path.join(ROOT, "profiles", "traffic_hexbeam_100x16x16.json")
    if not os.path.exists(tf):
        return None
    d = json.load(open(tf)).get("kernels", {})
    k, b = d.get(kernel), d.get("k_bwd_cols<float>")
    if k and kernel.endswith("_Float16>"):  # the iterative solve's dispatches: per its own backward ones
        b = d.get("k_bwd_cols<float, _Float16>")
    if not k or not b or nlevels <= 0 or launches_per_sweep <= 0:
        return None
    sweeps = b["launches"] / float(nlevels)
    return k["bytes_per_launch"] * k["launches"] / sweeps / launches_per_sweep


def stream_run(reps=5):
    """Achievable HBM read rate of one launch at the sizes of the solve's tree
    levels and beyond: the streaming probe (artisynth_core_amd/csrc/probe.hip,
    16-byte loads, 1-16 in flight per lane, one pass over the buffer) launched
    alone with cold caches -- before every timed launch the probe reads a
    separate 1 GiB buffer (a read, not a fill, so no dirty lines are written back
    during the timed launch).  Times from HIP events around each launch on
    torch's current stream (where the probe runs); a rocprofv3 kernel trace of
    this mode gives the kernel durations without launch overhead."""
    import torch
    P = ctypes.CDLL(os.path.join(ROOT, "artisynth_core_amd", "libkkt_probe.so"))
    P.kkt_probe_stream_read.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int, ctypes.c_void_p,
                                        ctypes.c_void_p]
    P.kkt_probe_stream_read.restype = ctypes.c_int
    st = torch.cuda.current_stream().cuda_stream
    flush = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
    sink = torch.zeros((1 << 30) // 4096 + 16, dtype=torch.float32, device="cuda")
    out = []
    for mb in (4, 8, 16, 32, 64, 128, 256, 512, 1024):
        a = torch.ones(mb << 18, dtype=torch.float32, device="cuda")
        r = {"MB": mb}
        for g in (1, 4, 8, 16):
            ts = []
            for _ in range(reps):
                assert P.kkt_probe_stream_read(flush.data_ptr(), flush.numel() * 4, 8, sink.data_ptr(), st) == 0
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                assert P.kkt_probe_stream_read(a.data_ptr(), mb << 20, g, sink.data_ptr(), st) == 0
                e1.record()
                torch.cuda.synchronize()
                ts.append(e0.elapsed_time(e1))
            t = float(np.median(ts))
            r["read_g%d_ms" % g] = t
            r["read_g%d_GBs" % g] = (mb << 20) / (t * 1e-3) / 1e9
        out.append(r)
        log("stream %5d MB: read GB/s with 1/4/8/16 loads in flight per lane: %s" % (
            mb, " / ".join("%.0f" % r["read_g%d_GBs" % g] for g in (1, 4, 8, 16))))
        del a
    print(json.dumps({"mode": "stream", "peak_GBs": PEAKS["hbm_GBs"], "results": out}), flush=True)


def pmc_run(nx, ny, nz):
    """Counter-collection workload: one FP32 factor + solve of step 0, then a plain
    256 MiB device copy whose FETCH_SIZE / WRITE_SIZE calibrate the counters."""
    import torch
    from artisynth_core_amd import KKT_PREC_FP32_REFINE
    from artisynth_core_amd import generators as G
    from artisynth_core_amd.solvers import Matrix, MI355XSolver
    beam, V, B = G.large_beam_sequence(nsteps=1, nx=nx, ny=ny, nz=nz, width=0.16 * ny / 16)
    s = MI355XSolver(precision=KKT_PREC_FP32_REFINE)
    s._L.kkt_set_graphs(s._h, 0)  # kernels dispatched one by one for the counters
    s.setIterativeHalf(True)  # counters for both the FP32 arrays' and the half copy's sweeps
    s.analyzeArrays(V[0], beam.colIdxs, beam.rowOffs, beam.n, Matrix.SYMMETRIC)
    s.factor(V[0])
    x = np.zeros(beam.n)
    s.solve(x, B[0])
    # a few stale-factor iterations: the preconditioner's sweeps on the
    # half-precision copy
    x2 = np.zeros(beam.n)
    s.iterativeSolve(V[0] * np.random.default_rng(5).uniform(0.999, 1.001, V[0].shape), x2, B[0], 10)
    a = torch.empty(256 << 20, dtype=torch.uint8, device="cuda")
    b = torch.ones(256 << 20, dtype=torch.uint8, device="cuda")
    a.copy_(b)
    torch.cuda.synchronize()
    print(json.dumps({"mode": "pmc", "copy_bytes": 256 << 20, "n": beam.n,
                      "info": s.info()}), flush=True)
    s.dispose()


def levels_run(nx, ny, nz, half):
    """Kernel-trace workload for the per-level sweep times: for each precision one
    factor of the beam's first KKT system, the solve-shape autotuner on it, then
    four stale-factor iterative solves of the second system (BiCGStab, tolExp 10)
    replaying the tuned graphs, as in the bench sequence."""
    import torch
    from artisynth_core_amd import KKT_PREC_FP32_REFINE, KKT_PREC_FP64
    from artisynth_core_amd import generators as G
    from artisynth_core_amd.solvers import Matrix, MI355XSolver
    beam, V, B = G.large_beam_sequence(nsteps=2, nx=nx, ny=ny, nz=nz, width=0.16 * ny / 16)
    out = {"mode": "levels", "n": beam.n}
    for prec, name in ((KKT_PREC_FP32_REFINE, "fp32_mfma_refine"), (KKT_PREC_FP64, "fp64")):
        s = MI355XSolver(precision=prec)
        s.setIterativeHalf(bool(half))
        s.analyzeArrays(V[0], beam.colIdxs, beam.rowOffs, beam.n, Matrix.SYMMETRIC)
        s.factor(V[0])
        s.autotuneSolve(3)
        x = np.zeros(beam.n)
        its = [int(s.iterativeSolve(V[1], x, B[1], 10)) for _ in range(4)]
        torch.cuda.synchronize()
        out[name] = {"nlevels": s.info()["nlevels"], "iterations": its, "shapes": s.getSolveShapes()}
        log("%s: %d levels, iterations %s" % (name, out[name]["nlevels"], its))
        s.dispose()
    print(json.dumps(out, default=lambda o: o.tolist() if hasattr(o, "tolist") else str(o)), flush=True)


def multi_run(nx, ny, nz, nrhs, reps, warmup):
    """Block solves with many right-hand sides (the LCP callers' solveMG of
    KKTSolver.java:1537-1588) on the hex beam's first KKT system: kkt_solve_multi_device
    (eight right-hand sides per launch sequence, refinement per column) against the
    same right-hand sides solved by one kkt_solve_device call each, for both factor
    precisions.  Every column is checked equal to its single solve, and residuals
    are recomputed on the host."""
    import torch
    import scipy.sparse as sp
    from artisynth_core_amd import KKT_PREC_FP32_REFINE, KKT_PREC_FP64
    from artisynth_core_amd import generators as G
    from artisynth_core_amd._lib import lib
    from artisynth_core_amd._lib import ptr as nptr
    dev = torch.device("cuda:0")
    t0 = time.time()
    beam, V, Bs = G.large_beam_sequence(nsteps=1, nx=nx, ny=ny, nz=nz, width=0.16 * ny / 16)
    n, nnz = beam.n, beam.nnz
    log("generated n=%d nnz=%d in %.1fs" % (n, nnz, time.time() - t0))
    ro = np.ascontiguousarray(beam.rowOffs, np.int32)
    ci = np.ascontiguousarray(beam.colIdxs, np.int32)
    v0 = np.ascontiguousarray(V[0])
    U = sp.csr_matrix((v0, ci, ro), shape=(n, n))
    A = U + sp.triu(U, 1).T
    rng = np.random.default_rng(7)
    Bh = rng.uniform(-1.0, 1.0, (nrhs, n))
    Bh[0] = Bs[0]
    dB = torch.from_numpy(np.ascontiguousarray(Bh.ravel())).to(dev)
    dXm = torch.zeros(n * nrhs, dtype=torch.float64, device=dev)
    dXs = torch.zeros(n * nrhs, dtype=torch.float64, device=dev)
    dv = torch.from_numpy(v0).to(dev)
    L = lib()
    out = {}
    for prec, name in ((KKT_PREC_FP32_REFINE, "fp32_mfma_refine"), (KKT_PREC_FP64, "fp64")):
        h = L.kkt_create()
        L.kkt_set_precision(h, prec)
        L.kkt_set_refine_tol(h, 1e-10, 10)
        if prec == KKT_PREC_FP64:
            L.kkt_set_ir_steps(h, 1)
        assert L.kkt_set_pattern(h, n, nnz, nptr(ro, ctypes.c_int), nptr(ci, ctypes.c_int), 1) == 0
        assert L.kkt_analyze_values(h, nptr(v0, ctypes.c_double)) == 0, L.kkt_last_error(h)
        assert L.kkt_factor_device(h, dv.data_ptr()) == 0, L.kkt_last_error(h)

        def multi():
            assert L.kkt_solve_multi_device(h, nrhs, dB.data_ptr(), dXm.data_ptr()) == 0, L.kkt_last_error(h)

        def single():
            for j in range(nrhs):
                assert L.kkt_solve_device(h, dB.data_ptr() + 8 * n * j, dXs.data_ptr() + 8 * n * j) == 0

        for _ in range(max(warmup, 1)):
            multi()
            single()
        torch.cuda.synchronize()
        t = {}
        for key, fn in (("multi", multi), ("single", single)):
            t0 = time.perf_counter()
            for _ in range(reps):
                fn()
            torch.cuda.synchronize()
            t[key] = (time.perf_counter() - t0) / reps
        Xm = dXm.cpu().numpy().reshape(nrhs, n)
        Xs = dXs.cpu().numpy().reshape(nrhs, n)
        equal = bool(np.array_equal(Xm, Xs))
        res = max(float(np.linalg.norm(A @ Xm[j] - Bh[j]) / np.linalg.norm(Bh[j])) for j in range(nrhs))
        out[name] = {"rhs_per_s": nrhs / t["multi"], "single_rhs_per_s": nrhs / t["single"],
                     "speedup": t["single"] / t["multi"], "ms_per_block": t["multi"] * 1e3,
                     "equal_to_single": equal, "max_rel_residual": res}
        log("%s: %d right-hand sides in %.2f ms batched (%.0f RHS/s) vs %.2f ms one by one "
            "(%.0f RHS/s): %.2fx, equal to single %s, max resid %.2e"
            % (name, nrhs, t["multi"] * 1e3, nrhs / t["multi"], t["single"] * 1e3, nrhs / t["single"],
               t["single"] / t["multi"], equal, res))
        L.kkt_destroy(h)
    print(json.dumps({"metric": "multi-RHS block solves/s (right-hand sides per second)",
                      "value": out["fp32_mfma_refine"]["rhs_per_s"], "unit": "RHS/s",
                      "higher_is_better": True, "n_gpus": 1, "reps": reps, "warmup": warmup,
                      "dtype": "fp32+fp64", "data": "synthetic (seeded neo-Hookean hex beam, random RHS)",
                      "config": {"workload": "block solve of %d RHS on the constrained hex beam %dx%dx%d, n=%d"
                                 % (nrhs, nx, ny, nz, n)},
                      "results": out}), flush=True)


def batch_run(nsys, reps, warmup, emit=True):
    """The batching workload: nsys independent KKT systems (DOF ~ U[300,4000], rigid bodies + tet
    blocks + contact equalities, FP64).  One batched factor + solve (one captured graph
    each) against the same systems factored and solved one handle at a time."""
    import torch
    from artisynth_core_amd import KKT_PREC_FP64
    from artisynth_core_amd import generators as G
    from artisynth_core_amd._lib import lib
    from artisynth_core_amd._lib import ptr as nptr
    from artisynth_core_amd.batch import BatchedKKT
    t0 = time.time()
    systems = G.batch_mix(nsys)
    log("generated %d systems (n %d..%d) in %.1fs" % (nsys, min(len(s[0]) - 1 for s in systems),
                                                    max(len(s[0]) - 1 for s in systems), time.time() - t0))
    dev = torch.device("cuda:0")
    t0 = time.time()
    B = BatchedKKT(systems, precision=KKT_PREC_FP64, ir_steps=1)
    t_an = time.time() - t0
    dv = torch.from_numpy(B.vals).to(dev)
    db = torch.from_numpy(B.b).to(dev)
    dx = torch.zeros(B.n, dtype=torch.float64, device=dev)
    for _ in range(max(1, warmup)):
        B.factor_device(dv.data_ptr())
        B.solve_device(db.data_ptr(), dx.data_ptr())
    torch.cuda.synchronize()
    lat = []
    for _ in range(reps):
        t0 = time.perf_counter()
        B.factor_device(dv.data_ptr())
        B.solve_device(db.data_ptr(), dx.data_ptr())
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - t0) * 1e3)
    xb = dx.cpu().numpy()
    res = B.per_system_residuals(xb, systems)
    # per-system launches
    L = lib()
    hs, dvs, dbs, dxs = [], [], [], []
    for ro, ci, va, b in systems:
        h = L.kkt_create()
        L.kkt_set_precision(h, KKT_PREC_FP64)
        L.kkt_set_ir_steps(h, 1)
        ro32 = np.ascontiguousarray(ro, np.int32)
        ci32 = np.ascontiguousarray(ci, np.int32)
        va64 = np.ascontiguousarray(va, np.float64)
        assert L.kkt_set_pattern(h, len(ro) - 1, int(ro[-1]), nptr(ro32, ctypes.c_int),
                                 nptr(ci32, ctypes.c_int), 1) == 0
        assert L.kkt_analyze_values(h, nptr(va64, ctypes.c_double)) == 0
        hs.append(h)
        dvs.append(torch.from_numpy(va64).to(dev))
        dbs.append(torch.from_numpy(np.ascontiguousarray(b, np.float64)).to(dev))
        dxs.append(torch.zeros(len(b), dtype=torch.float64, device=dev))
    for q, h in enumerate(hs):
        assert L.kkt_factor_device(h, dvs[q].data_ptr()) == 0
        assert L.kkt_solve_device(h, dbs[q].data_ptr(), dxs[q].data_ptr()) == 0
    torch.cuda.synchronize()
    lat1 = []
    for _ in range(max(1, reps // 4)):
        t0 = time.perf_counter()
        for q, h in enumerate(hs):
            L.kkt_factor_device(h, dvs[q].data_ptr())
            L.kkt_solve_device(h, dbs[q].data_ptr(), dxs[q].data_ptr())
        torch.cuda.synchronize()
        lat1.append((time.perf_counter() - t0) * 1e3)
    # same result alone and in the batch (default solve shapes on both sides)
    diff = 0.0
    for q in range(nsys):
        xa = dxs[q].cpu().numpy()
        xs = xb[B.offs[q]:B.offs[q + 1]]
        diff = max(diff, float(np.linalg.norm(xa - xs) / max(np.linalg.norm(xa), 1e-300)))
    # then the autotuner on both sides (the batch's levels hold every system's
    # fronts; each per-system handle is tuned on its own), timed again
    t0 = time.perf_counter()
    B.autotune_solve(3)
    t_tune = time.perf_counter() - t0
    for h in hs:
        assert L.kkt_autotune_solve(h, 3) == 0, L.kkt_last_error(h)
    latt = []
    for _ in range(reps):
        t0 = time.perf_counter()
        B.factor_device(dv.data_ptr())
        B.solve_device(db.data_ptr(), dx.data_ptr())
        torch.cuda.synchronize()
        latt.append((time.perf_counter() - t0) * 1e3)
    rest = B.per_system_residuals(dx.cpu().numpy(), systems)
    lat1t = []
    for _ in range(max(1, reps // 4)):
        t0 = time.perf_counter()
        for q, h in enumerate(hs):
            L.kkt_factor_device(h, dvs[q].data_ptr())
            L.kkt_solve_device(h, dbs[q].data_ptr(), dxs[q].data_ptr())
        torch.cuda.synchronize()
        lat1t.append((time.perf_counter() - t0) * 1e3)
    for h in hs:
        L.kkt_destroy(h)
    inf = B.info()
    t_b, t_1 = float(np.median(latt)), float(np.median(lat1t))
    line = {"mode": "batch", "nsys": nsys, "n_total": B.n, "nnz_total": B.nnz,
            "batched_ms": t_b, "batched_p99_ms": float(np.percentile(latt, 99)),
            "per_system_ms": t_1, "speedup": t_1 / t_b,
            "batched_systems_per_s": nsys / (t_b * 1e-3),
            "per_system_systems_per_s": nsys / (t_1 * 1e-3),
            "solve_shapes": "autotuned (batch %.2f s); default shapes: batched %.2f ms, per system %.2f ms" % (
                t_tune, float(np.median(lat)), float(np.median(lat1))),
            "batched_default_ms": float(np.median(lat)), "per_system_default_ms": float(np.median(lat1)),
            "max_rel_residual": max(max(res), max(rest)), "max_diff_alone_vs_batch": diff,
            "factor_gflop": inf["factor_flops"] / 1e9, "nlevels": inf["nlevels"],
            "analyze_s": t_an, "dtype": "f64"}
    log("batch: %.2f ms per batch of %d vs %.2f ms one by one (%.1fx; autotuned shapes; default %.2f vs %.2f), "
        "max resid %.2e" % (t_b, nsys, t_1, t_1 / t_b, float(np.median(lat)), float(np.median(lat1)),
                            line["max_rel_residual"]))
    if emit:
        print(json.dumps(line), flush=True)
    B.dispose()
    return line


def group_run(ninst, eps, reps, nworkers, emit=True):
    """One stale-factor iterative step of each of ninst serving instances (FP32
    factor + FP64 BiCGStab to 1e-10, each instance's values drifted by up to eps
    relatively), three ways: all of them in one group solve
    (kkt_group_iterative_solve_device), one handle at a time, and nworkers host
    threads each running its share one handle at a time (the serving loop's
    concurrency).  Same results on every path (bitwise; checked)."""
    import threading
    import torch
    from artisynth_core_amd import KKT_PREC_FP32_REFINE
    from artisynth_core_amd import generators as G
    from artisynth_core_amd._lib import lib
    from artisynth_core_amd._lib import ptr as nptr
    from artisynth_core_amd.batch import IterativeGroup
    t0 = time.time()
    mix = G.serving_mix(ninst)
    log("generated %d instances (n %d..%d) in %.1fs" % (
        ninst, min(len(m[2]) - 1 for m in mix), max(len(m[2]) - 1 for m in mix), time.time() - t0))
    L = lib()
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(77)
    hs, dv, db, dx = [], [], [], []
    t0 = time.time()
    for kind, nd, ro, ci, va, b in mix:
        h = L.kkt_create()
        ro32 = np.ascontiguousarray(ro, np.int32)
        ci32 = np.ascontiguousarray(ci, np.int32)
        va64 = np.ascontiguousarray(va, np.float64)
        assert L.kkt_set_precision(h, KKT_PREC_FP32_REFINE) == 0
        assert L.kkt_set_pattern(h, len(ro) - 1, int(ro[-1]), nptr(ro32, ctypes.c_int),
                                 nptr(ci32, ctypes.c_int), 1) == 0
        assert L.kkt_analyze_values(h, nptr(va64, ctypes.c_double)) == 0, L.kkt_last_error(h)
        assert L.kkt_factor(h, nptr(va64, ctypes.c_double)) == 0, L.kkt_last_error(h)
        hs.append(h)
        dv.append(torch.from_numpy(va64 * (1.0 + eps * rng.uniform(-1, 1, len(va64)))).to(dev))
        db.append(torch.from_numpy(np.ascontiguousarray(b, np.float64)).to(dev))
        dx.append(torch.zeros(len(b), dtype=torch.float64, device=dev))
    torch.cuda.synchronize()
    log("analysed + factored %d handles in %.1fs" % (ninst, time.time() - t0))
    tol, mi = 1e-10, 30

    def single(qs, out):
        for q in qs:
            out[q] = L.kkt_iterative_solve_device(hs[q], dv[q].data_ptr(), db[q].data_ptr(), dx[q].data_ptr(),
                                                  tol, mi)

    its = [0] * ninst
    single(range(ninst), its)  # warm (graphs captured)
    x1 = [t.cpu().numpy() for t in dx]
    grp = IterativeGroup()
    ptrs = ([t.data_ptr() for t in dv], [t.data_ptr() for t in db], [t.data_ptr() for t in dx])
    rc, st = grp.solve(hs, *ptrs, tol, mi)
    assert rc == 0, grp.last_error()
    same = list(st) == its and all(np.array_equal(x1[q], dx[q].cpu().numpy()) for q in range(ninst))

    def timed(fn):
        out = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            out.append((time.perf_counter() - t0) * 1e3)
        return float(np.median(out))

    t_grp = timed(lambda: grp.solve(hs, *ptrs, tol, mi))
    t_one = timed(lambda: single(range(ninst), [0] * ninst))

    def threaded():
        out = [0] * ninst
        th = [threading.Thread(target=single, args=(range(w, ninst, nworkers), out)) for w in range(nworkers)]
        for t in th:
            t.start()
        for t in th:
            t.join()

    t_thr = timed(threaded)
    grp.dispose()
    for h in hs:
        L.kkt_destroy(h)
    line = {"mode": "group", "ninst": ninst, "eps": eps, "total_dof": int(sum(len(m[2]) - 1 for m in mix)),
            "iterations_min": int(min(its)), "iterations_max": int(max(its)),
            "iterations_mean": float(np.mean(its)), "unconverged": int(sum(1 for i in its if i <= 0)),
            "group_ms": t_grp, "one_by_one_ms": t_one, "threads_ms": t_thr, "threads": nworkers,
            "group_solves_per_s": ninst / (t_grp * 1e-3), "one_by_one_solves_per_s": ninst / (t_one * 1e-3),
            "threads_solves_per_s": ninst / (t_thr * 1e-3), "bitwise_same": bool(same)}
    log("group: %d stale-factor solves (iterations %d..%d) in %.2f ms as one group, %.2f ms one by one, "
        "%.2f ms on %d threads; bitwise same %s" % (ninst, min(its), max(its), t_grp, t_one, t_thr, nworkers, same))
    if emit:
        print(json.dumps(line), flush=True)
    return line


def serve_run(ninst, nreq, rate, nworkers, budget_gb, eps=5e-3, emit=True, group_runners=2, group_max=64):
    """The serving workload: ninst instances (1k-90k DOF, beam/tet/rigid), a Poisson
    stream of step requests at `rate` req/s, HBM-budget admission, FP32-MFMA + FP64
    refinement (1e-10 relative residual), through the C++ serving loop
    (kkt_serve_*).  Each instance's values drift as a seeded random walk (every
    entry moves by up to eps relatively per step), so the resident factor goes
    stale and some steps refactor.  All request inputs are generated in HBM before
    the clock starts and submitted as device arrays (kkt_serve_submit_device); the
    submission loop only sleeps and submits.  Latency = submission to completion
    (server clock); p50/p99, achieved offered rate, queue depth, admissions and
    evictions are reported.  With group_runners > 0 the iterative steps of idle,
    factored instances are solved together in group solves (reported as
    group_calls / group_requests)."""
    import torch
    import scipy.sparse as sp
    from artisynth_core_amd import generators as G
    from artisynth_core_amd.serving import SolveServer, latency_stats
    t0 = time.time()
    mix = G.serving_mix(ninst)
    log("generated %d instances (n %d..%d, total %d) in %.1fs" % (
        ninst, min(len(m[2]) - 1 for m in mix), max(len(m[2]) - 1 for m in mix),
        sum(len(m[2]) - 1 for m in mix), time.time() - t0))
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(2000)
    gaps = rng.exponential(1.0 / rate, nreq)
    who = rng.integers(0, ninst, nreq)
    # inputs in HBM: per instance its base values and right-hand side; per request
    # the step's values (random walk) and a solution buffer
    t0 = time.time()
    base = [torch.from_numpy(np.ascontiguousarray(m[4])).to(dev) for m in mix]
    rhs = [torch.from_numpy(np.ascontiguousarray(m[5])).to(dev) for m in mix]
    drift = [torch.zeros_like(v) for v in base]
    gen = torch.Generator(device=dev)
    gen.manual_seed(2000)
    vals, xs, steps = [], [], np.zeros(ninst, np.int64)
    for q in range(nreq):
        i = int(who[q])
        drift[i] += eps * (2.0 * torch.rand(base[i].shape, generator=gen, device=dev, dtype=torch.float64) - 1.0)
        vals.append(base[i] * (1.0 + drift[i]))
        xs.append(torch.zeros(len(mix[i][5]), dtype=torch.float64, device=dev))
        steps[i] += 1
    torch.cuda.synchronize()
    log("request inputs in HBM: %.1f GB in %.1fs" % (sum(v.numel() for v in vals) * 8 / 1e9, time.time() - t0))
    srv = SolveServer(budget_bytes=int(budget_gb * (1 << 30)), nworkers=nworkers, group_runners=group_runners,
                      group_max=group_max)
    for i, (kind, nd, ro, ci, va, b) in enumerate(mix):
        srv.register(i, ro, ci, va)
    t0 = time.time()
    warm_x = [torch.zeros(len(m[5]), dtype=torch.float64, device=dev) for m in mix]
    warm = [srv.submit_device(i, base[i].data_ptr(), rhs[i].data_ptr(), warm_x[i].data_ptr()) for i in range(ninst)]
    for r in warm:
        r.wait()
    errs = [r.error for r in warm if r.error]
    assert not errs, errs[:3]
    st0 = srv.stats()
    log("admitted + first factor of %d instances in %.1fs (%.1f GB resident)" % (
        ninst, time.time() - t0, st0["bytes_resident"] / 2 ** 30))
    reqs = []
    st_start = srv.stats()
    t_start = time.perf_counter()
    t_next = t_start
    for q in range(nreq):
        t_next += gaps[q]
        now = time.perf_counter()
        if t_next > now:
            time.sleep(t_next - now)
        i = int(who[q])
        reqs.append(srv.submit_device(i, vals[q].data_ptr(), rhs[i].data_ptr(), xs[q].data_ptr()))
    t_sub = time.perf_counter()
    for r in reqs:
        r.wait()
    t_end = time.perf_counter()
    st = srv.stats()
    errs = [r.error for r in reqs if r.error]
    lat = latency_stats(reqs)
    # residual check on a sample of the results (host)
    res = []
    for q in range(0, nreq, max(1, nreq // 100)):
        i = int(who[q])
        kind, nd, ro, ci, va, b = mix[i]
        n = len(ro) - 1
        U = sp.csr_matrix((vals[q].cpu().numpy(), ci, ro), shape=(n, n))
        A = U + sp.triu(U, 1).T
        x = xs[q].cpu().numpy()
        res.append(float(np.linalg.norm(A @ x - b) / np.linalg.norm(b)))
    kinds = [r.kind for r in reqs]
    its = [r.iters for r in reqs if r.kind == "iterative"]
    offered = nreq / (t_sub - t_start)
    line = {"mode": "serve", "ninst": ninst, "nreq": nreq, "rate_target": rate,
            "offered_rate_achieved": offered, "workers": nworkers, "budget_gb": budget_gb,
            "throughput_req_s": nreq / (t_end - t_start), "p50_ms": lat["p50_ms"], "p99_ms": lat["p99_ms"],
            "mean_ms": lat["mean_ms"], "errors": len(errs), "direct": kinds.count("direct"),
            "iterative": kinds.count("iterative"), "max_rel_residual_sample": max(res),
            "resident_gb": st["bytes_resident"] / 2 ** 30, "peak_gb": st["bytes_peak"] / 2 ** 30,
            "admissions": st["admissions"], "evictions": st["evictions"], "queue_max": st["queue_max"],
            "releases": st["releases"] - st_start["releases"],
            "concurrency": (st["service_ms"] - st_start["service_ms"]) / (1e3 * (t_end - t_start)),
            "admissions_timed": st["admissions"] - st_start["admissions"],
            "group_runners": group_runners, "group_max": group_max,
            "mean_iterations": float(np.mean(its)) if its else 0.0,
            "group_calls": st["group_calls"] - st_start["group_calls"],
            "group_requests": st["group_requests"] - st_start["group_requests"],
            "total_dof": int(sum(len(m[2]) - 1 for m in mix)), "inputs": "device (HBM), drift eps %g" % eps}
    log("serve: offered %.0f req/s (target %.0f), served %.0f req/s, p50 %.2f ms, p99 %.2f ms, direct %d, "
        "releases %d, evictions %d, group solves %d of %.1f requests on average, max resid %.2e" % (
            offered, rate, line["throughput_req_s"], lat["p50_ms"], lat["p99_ms"], line["direct"], line["releases"],
            line["evictions"],
            line["group_calls"], line["group_requests"] / max(1, line["group_calls"]), max(res)))
    if emit:
        print(json.dumps(line), flush=True)
    srv.shutdown()
    del vals, xs, base, rhs, drift
    torch.cuda.empty_cache()
    return line


def precision_run(dims, nsteps, warmup, with_oracle, max_super=2048):
    """The precision workload: seeded Delaunay tet box (0.1 x 0.02 x 0.02 m, ~210k DOF at the
    default lattice), Mooney-Rivlin c10 = 1e3, c01 = 0, kappa = 1e5, attached end.
    Every step is a direct solve: FP32 fronts on MFMA with FP64 refinement to the
    metric's 1e-8 relative residual, against FP64 fronts with one refinement step;
    inputs resident in HBM.  Per solve: relative residual (host FP64 SpMV), forward
    error of the FP32 path against the FP64 path, refinement steps, FP64 fallbacks.
    One extra factor per precision is timed per kernel class."""
    import torch
    import scipy.sparse as sp
    from artisynth_core_amd import KKT_PREC_FP32_REFINE, KKT_PREC_FP64
    from artisynth_core_amd import generators as G
    from artisynth_core_amd._lib import KktInfo, lib
    from artisynth_core_amd._lib import ptr as nptr
    dev = torch.device("cuda:0")
    t0 = time.time()
    body, V, B = G.tet_precision_sequence(nsteps=nsteps, dims=dims)
    n, nnz = body.n, body.nnz
    log("generated %d tet systems n=%d nnz=%d (%d tets) in %.1fs" % (
        len(V), n, nnz, len(body.elems), time.time() - t0))
    ro = np.ascontiguousarray(body.rowOffs, np.int32)
    ci = np.ascontiguousarray(body.colIdxs, np.int32)
    vals_d = [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in V]
    b_d = [torch.from_numpy(np.ascontiguousarray(b)).to(dev) for b in B]
    x_d = torch.zeros(n, dtype=torch.float64, device=dev)
    L = lib()
    out = {}
    for prec, name in ((KKT_PREC_FP32_REFINE, "fp32_mfma_refine"), (KKT_PREC_FP64, "fp64")):
        h = L.kkt_create()
        L.kkt_set_precision(h, prec)
        L.kkt_set_refine_tol(h, 1e-8, 10)
        assert L.kkt_set_max_supernode(h, max_super) == 0, L.kkt_last_error(h)
        if prec == KKT_PREC_FP64:
            L.kkt_set_ir_steps(h, 1)
        ta = time.time()
        assert L.kkt_set_pattern(h, n, nnz, nptr(ro, ctypes.c_int), nptr(ci, ctypes.c_int), 1) == 0
        v0 = np.ascontiguousarray(V[0])
        assert L.kkt_analyze_values(h, nptr(v0, ctypes.c_double)) == 0, L.kkt_last_error(h)
        t_an = time.time() - ta
        for _ in range(max(warmup, 1)):
            assert L.kkt_factor_device(h, vals_d[0].data_ptr()) == 0, L.kkt_last_error(h)
            assert L.kkt_solve_device(h, b_d[0].data_ptr(), x_d.data_ptr()) == 0
        # the solve-shape autotuner on the warm-up factor (untimed setup, as in the
        # sequence bench)
        assert L.kkt_autotune_solve(h, 3) == 0, L.kkt_last_error(h)
        torch.cuda.synchronize()
        tf, ts, xs, info = [], [], [], []
        for k in range(len(V)):
            t1 = time.perf_counter()
            assert L.kkt_factor_device(h, vals_d[k].data_ptr()) == 0, L.kkt_last_error(h)
            t2 = time.perf_counter()
            assert L.kkt_solve_device(h, b_d[k].data_ptr(), x_d.data_ptr()) == 0, L.kkt_last_error(h)
            torch.cuda.synchronize()
            t3 = time.perf_counter()
            tf.append(t2 - t1)
            ts.append(t3 - t2)
            xs.append(x_d.cpu().numpy().copy())
            inf = KktInfo()
            L.kkt_get_info(h, ctypes.byref(inf))
            info.append(inf.as_dict())
        # one more factor with per-class kernel timing (eager launches)
        L.kkt_timing_reset(h, 0x0F)
        assert L.kkt_factor_device(h, vals_d[0].data_ptr()) == 0
        ms = (ctypes.c_double * 8)()
        cnt = (ctypes.c_longlong * 8)()
        uflops = ctypes.c_double()
        L.kkt_timing_get(h, ms, cnt, ctypes.byref(uflops))
        L.kkt_timing_reset(h, 0)
        kern = {TIMING_CLASSES[c]: ms[c] for c in range(4)}
        res = []
        for k in range(len(V)):
            U = sp.csr_matrix((V[k], ci, ro), shape=(n, n))
            A = U + sp.triu(U, 1).T
            res.append(float(np.linalg.norm(A @ xs[k] - B[k]) / np.linalg.norm(B[k])))
        fl = info[0]["factor_flops"]
        out[name] = {"factor_ms": [1e3 * t for t in tf], "solve_ms": [1e3 * t for t in ts],
                     "solves_per_s": len(V) / (sum(tf) + sum(ts)),
                     "factor_tflops": fl / (np.median(tf)) / 1e12,
                     "rel_residual": res, "refine_steps": [i["refine_steps"] for i in info],
                     "fp64_fallbacks": info[-1]["fp64_fallbacks"], "analyze_s": t_an,
                     "npos": info[0]["npos"], "nneg": info[0]["nneg"],
                     "nperturbed": info[0]["nperturbed"], "n2x2": info[0]["n2x2"],
                     "pivoted_blocks": info[0]["pivoted_blocks"], "factor_kernel_ms": kern,
                     "update_tflops": uflops.value / (kern["panel_update"] * 1e-3) / 1e12
                     if kern["panel_update"] > 0 else None, "x": xs}
        log("%s: %.2f direct solves/s, factor %.1f ms (%.1f TFLOP/s), solve+refine %.1f ms, "
            "max rel resid %.2e, refine steps %s" % (
                name, out[name]["solves_per_s"], 1e3 * np.median(tf), out[name]["factor_tflops"],
                1e3 * np.median(ts), max(res), out[name]["refine_steps"]))
        L.kkt_destroy(h)
    r32, r64 = out["fp32_mfma_refine"], out["fp64"]
    fwd = [float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(r32["x"], r64["x"])]
    line = {"mode": "precision", "workload": "Delaunay tet box %s lattice, n=%d, nnz_upper=%d, "
                                            "%d tets, Mooney-Rivlin c10=1e3 kappa=1e5" % (
                                                "x".join(map(str, dims)), n, nnz, len(body.elems)),
            "n": n, "nnz_upper": nnz, "steps": len(V), "max_supernode": max_super,
            "nlevels": info[0]["nlevels"],
            "factor_gflop": info[0]["factor_flops"] / 1e9, "nnzL": info[0]["nnzL"],
            "fp32": {k: v for k, v in r32.items() if k != "x"},
            "fp64": {k: v for k, v in r64.items() if k != "x"},
            "speedup_vs_fp64": r32["solves_per_s"] / r64["solves_per_s"],
            "fwd_err_fp32_vs_fp64": fwd}
    if with_oracle:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle as O
        from symbolic_helpers import product_symbolic
        S = product_symbolic(ro, ci, V[0])
        t1 = time.perf_counter()
        f = O.Factor(ro, ci, V[0], perm=S["perm"], sup_ptr=S["sup_ptr"])
        _, xo = f.refine(B[0], tol=1e-10, maxSteps=1)
        line["oracle_step0"] = {
            "seconds": time.perf_counter() - t1, "cores": 1,
            "fp64_err": float(np.linalg.norm(r64["x"][0] - xo) / np.linalg.norm(xo)),
            "fp32_err": float(np.linalg.norm(r32["x"][0] - xo) / np.linalg.norm(xo))}
    log("precision: FP32-MFMA+refine %.2fx FP64, fwd err vs FP64 %s" % (
        line["speedup_vs_fp64"], ["%.1e" % e for e in fwd]))
    print(json.dumps(line), flush=True)


def cpu_baseline(beam, V, B, x64, x32):
    """CPU baseline on the host cores of the machine running the bench: MKL PARDISO with the reference's
    settings, dlopened from the host's MKL install and driven through the same
    step sequence with KKTSolver.factorAndSolve's policy (oracle/pardiso_run.py,
    a child process: PARDISO fails in a process that has loaded the GPU
    libraries), on all the threads this process may use and on one thread.
    When MKL is absent, the oracle (one thread, one direct solve of step 0) is
    reported instead, and the sample says so."""
    import subprocess
    import tempfile
    tmp = tempfile.mkdtemp(prefix="kkt_cpu_")
    inp, outp = os.path.join(tmp, "in.npz"), os.path.join(tmp, "out.json")
    try:
        np.savez(inp, rowOffs=np.asarray(beam.rowOffs, np.int32), colIdxs=np.asarray(beam.colIdxs, np.int32),
                 V=np.stack(V), B=np.stack(B), X64=np.stack(x64), X32=np.stack(x32))
        r = subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "pardiso_run.py"), inp, outp],
                           capture_output=True, text=True, timeout=600)
        if r.returncode != 0 or not os.path.exists(outp):
            return {"value": None, "unit": "solves/s", "cores": None, "kind": "reference",
                    "sample": "PARDISO run failed (rc %d): %s" % (r.returncode, r.stderr[-500:])}
        d = json.load(open(outp))
    finally:
        for f in (inp, outp):
            if os.path.exists(f):
                os.remove(f)
        os.rmdir(tmp)
    if not d.get("available"):
        return cpu_baseline_oracle(beam, V[0], B[0], x64[0], x32[0], d.get("why", "MKL PARDISO unavailable"))
    out = {"value": d.pop("solves_per_s"), "unit": "solves/s", "cores": d["threads"], "kind": "reference",
           "sample": "MKL PARDISO (reference settings: mtype=-2, METIS, Bunch-Kaufman, perturbation 1e-8, "
                     "scaling + matching, 1 refinement step) on the same %d-step sequence with the factorAndSolve "
                     "policy (phase 22+33, then phase-33 CGS with the stale factor), %d repetitions on %d threads; "
                     "20 direct factor+solve repetitions; one thread: 2 direct + 5 hybrid steps"
                     % (d["steps"], d["reps"], d["threads"])}
    out.update(d)
    return out


def cpu_baseline_oracle(beam, vals, b, x64, x32, why):
    """The CPU oracle (FP64 multifrontal LDL^T, one thread) on the first step of the
    same sequence: analysis excluded, factor + solve + one refinement step timed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    try:
        import oracle as O
        from symbolic_helpers import product_symbolic
        S = product_symbolic(beam.rowOffs, beam.colIdxs, vals)
        t0 = time.perf_counter()
        f = O.Factor(beam.rowOffs, beam.colIdxs, vals, perm=S["perm"], sup_ptr=S["sup_ptr"])
        steps, xo = f.refine(b, tol=1e-10, maxSteps=1)
        dt = time.perf_counter() - t0
        return {"value": 1.0 / dt, "unit": "solves/s", "cores": 1, "kind": "port",
                "sample": "%s; the oracle instead: one direct solve (factor + solve + 1 refinement step) of "
                          "step 0 of the same n=%d sequence, %.1f s" % (why, beam.n, dt),
                "gpu_fp64_err_vs_oracle": float(np.linalg.norm(x64 - xo) / np.linalg.norm(xo)),
                "gpu_fp32_err_vs_oracle": float(np.linalg.norm(x32 - xo) / np.linalg.norm(xo))}
    except Exception as e:  # report, never fall back to anything
        return {"value": None, "unit": "solves/s", "cores": 1, "kind": "port",
                "sample": "%s; oracle failed: %s" % (why, e)}


if __name__ == "__main__":
    main()
