"""CPU baseline harness: MKL PARDISO (oracle/pardiso.py, the reference's settings)
driven through a KKT step sequence the way KKTSolver.factorAndSolve drives it --
CPU BASELINE ONLY, run by bench.py's cpu_baseline leg as a child process.

It runs in a process of its own because MKL PARDISO fails (error -2 at the first
factorization, nnz(L) = -1 after the analysis) in a process that has loaded the
GPU framework's libraries first; a child that loads only numpy, scipy and MKL is
unaffected.

    python oracle/pardiso_run.py IN.npz OUT.json

IN.npz (written with numpy, no pickles): rowOffs, colIdxs (0-based upper CSR),
V (steps x nnz values), B (steps x n right-hand sides), X64, X32 (the GPU
solutions of the first steps, for the error columns).  Timed on the threads of
OMP_NUM_THREADS (else the affinity mask) and on one thread.

Policy (KKTSolver.java:930-1014 with PardisoSolver.iterativeSolve,
PardisoSolver.java:1697-1737): step 0 factors (phase 22) and solves (phase 33,
one refinement step); each later step first runs phase 33 with iparm[3] = 102 on
its new values (CGS preconditioned by the stale factor, pardisoMkl.cc:704-750)
while KKTSolver's estimateOptimalCount (KKTSolver.java:889-922, on this run's own
step times) says the iterative steps still pay, and refactors + solves when it
does not, or when CGS returns <= 0 or needed more than MAX_ITER iterations -- the
GPU sequence's rule; the failed attempt's time counts.
"""
import json
import os
import sys
import time

import numpy as np
import scipy.sparse as sp

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import pardiso as P  # noqa: E402

MAX_ITER = 30


def main(inp, outp):
    ok, what = P.available()
    if not ok:
        json.dump({"available": False, "why": what}, open(outp, "w"))
        return
    d = np.load(inp, allow_pickle=False)
    ro, ci, V, B = d["rowOffs"], d["colIdxs"], d["V"], d["B"]
    X64, X32 = d["X64"], d["X32"]
    n = len(ro) - 1
    nthreads = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
    nsteps = len(V)

    def resid(k, x):
        U = sp.csr_matrix((V[k], ci, ro), shape=(n, n))
        A = U + sp.triu(U, 1).T
        return float(np.linalg.norm(A @ x - B[k]) / np.linalg.norm(B[k]))

    def sequence(p):
        lat, kinds, iters, xs = [], [], [], []
        # KKTSolver.factorAndSolve's state (KKTSolver.java:883-887, 930-1014)
        direct_ms, direct_cnt, iter_ms, first_iter_ms, iter_cnt = 0.0, 0, 0.0, 0.0, 0

        def optimal_count():  # estimateOptimalCount, KKTSolver.java:889-922
            ft, it_avg = direct_ms / direct_cnt, iter_ms / iter_cnt
            if it_avg >= ft:
                return iter_cnt + 1
            dt = 2.0 * (it_avg - first_iter_ms) / iter_cnt
            if dt <= 0:
                return iter_cnt + 19
            return int(np.ceil(np.sqrt(2.0 * (ft - first_iter_ms + dt) / dt)))

        for k in range(nsteps):
            t0 = time.perf_counter()
            it = 0
            if direct_cnt > 0 and (iter_cnt == 0 or iter_cnt + 1 < optimal_count()):
                it, x = p.iterative_solve(V[k], B[k], 10)
            if 0 < it <= MAX_ITER:
                kinds.append("i")
                iter_ms += (time.perf_counter() - t0) * 1e3
                if iter_cnt == 0:
                    first_iter_ms = iter_ms
                iter_cnt += 1
            else:
                t1 = time.perf_counter()
                p.factor(V[k])
                x = p.solve(B[k])
                direct_ms += (time.perf_counter() - t1) * 1e3
                direct_cnt += 1
                iter_cnt, iter_ms = 0, 0.0
                kinds.append("D")
            lat.append((time.perf_counter() - t0) * 1e3)
            iters.append(int(it))
            xs.append(x)
        return lat, "".join(kinds), iters, xs

    P.set_threads(nthreads)
    t0 = time.perf_counter()
    p = P.Pardiso(ro, ci, V[0])
    analyze_s = time.perf_counter() - t0
    reps = 3
    seq_lat, seq_el, seq_kinds = [], [], ""
    for r in range(reps):
        t0 = time.perf_counter()
        lat, kinds, iters, xs = sequence(p)
        seq_el.append(time.perf_counter() - t0)
        seq_lat += lat
        seq_kinds += kinds
    direct = []
    for r in range(20):
        t0 = time.perf_counter()
        p.factor(V[0])
        p.solve(B[0])
        direct.append((time.perf_counter() - t0) * 1e3)
    inf = p.info()
    p.release()
    res = max(resid(k, xs[k]) for k in range(nsteps))
    err = lambda X: max(float(np.linalg.norm(X[k] - xs[k]) / np.linalg.norm(xs[k])) for k in range(len(X)))
    # one thread: direct factor + solve, and the first five hybrid steps
    P.set_threads(1)
    p1 = P.Pardiso(ro, ci, V[0])
    d1, h1 = [], []
    for r in range(2):
        t0 = time.perf_counter()
        p1.factor(V[0])
        p1.solve(B[0])
        d1.append((time.perf_counter() - t0) * 1e3)
    for k in range(1, min(6, nsteps)):
        t0 = time.perf_counter()
        p1.iterative_solve(V[k], B[k], 10)
        h1.append((time.perf_counter() - t0) * 1e3)
    p1.release()
    hyb = [l for l, c in zip(seq_lat, seq_kinds) if c == "i"]
    out = {
        "available": True, "engine": what, "threads": nthreads,
        "solves_per_s": nsteps / float(np.median(seq_el)), "reps": reps, "steps": nsteps,
        "step_kinds": kinds, "cgs_iters": iters,
        "p50_ms": float(np.percentile(seq_lat, 50)), "p99_ms": float(np.percentile(seq_lat, 99)),
        "direct_solves_per_s": 1e3 / float(np.median(direct)),
        "direct_p50_ms": float(np.percentile(direct, 50)), "direct_p99_ms": float(np.percentile(direct, 99)),
        "hybrid_step_p50_ms": float(np.percentile(hyb, 50)) if hyb else None,
        "analyze_s": analyze_s, "nnzL": inf["nnzL"], "npos": inf["npos"], "nneg": inf["nneg"],
        "nperturbed": inf["nperturbed"], "max_rel_residual": res,
        "gpu_fp64_err_vs_pardiso": err(X64), "gpu_fp32_err_vs_pardiso": err(X32),
        "one_thread": {"direct_ms": d1, "hybrid_step_ms": h1,
                       "sequence_solves_per_s_est": nsteps / (1e-3 * (np.median(d1) + (nsteps - 1) * np.median(h1)))},
    }
    json.dump(out, open(outp, "w"))


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2])
