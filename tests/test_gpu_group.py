"""GPU tests of the group iterative solve (kkt_group_iterative_solve_device):
the stale-factor BiCGStab of many handles in lockstep, one launch per step for
all of them.  Each handle's solution and iteration count must equal, bitwise,
what its own kkt_iterative_solve_device call gives on an identical handle (same
pattern, factor values and plan), for uneven sizes and structures, both factor
precisions, handles converging at different iterations, and limits and
argument errors."""
import ctypes
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from artisynth_core_amd import KKT_PREC_FP32_REFINE, KKT_PREC_FP64
from artisynth_core_amd import generators as G
from artisynth_core_amd._lib import lib
from artisynth_core_amd._lib import ptr as nptr
from artisynth_core_amd.batch import IterativeGroup

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


def _systems():
    """Uneven instances of the serving mix's three generators, with a step's
    perturbed values (so BiCGStab iterates) and its right-hand side."""
    out = []
    for q, (kind, nd) in enumerate([("beam", 900), ("tet", 1500), ("rigid", 600), ("beam", 3000),
                                    ("tet", 700), ("rigid", 2500), ("beam", 5000)]):
        ro, ci, va, b = G.serving_instance(kind, nd, 100 + q)
        rng = np.random.default_rng(7 + q)
        eps = 1e-3 * (1 + q % 3)  # different drifts: different iteration counts
        v1 = va * (1.0 + eps * rng.uniform(-1, 1, va.shape))
        out.append((np.asarray(ro, np.int32), np.asarray(ci, np.int32), np.asarray(va, np.float64), v1,
                    np.asarray(b, np.float64)))
    return out


def _handle(L, ro, ci, va, prec):
    h = L.kkt_create()
    assert L.kkt_set_precision(h, prec) == 0
    assert L.kkt_set_pattern(h, len(ro) - 1, int(ro[-1]), nptr(ro, ctypes.c_int), nptr(ci, ctypes.c_int), 1) == 0
    assert L.kkt_analyze_values(h, nptr(va, ctypes.c_double)) == 0, L.kkt_last_error(h)
    assert L.kkt_factor(h, nptr(va, ctypes.c_double)) == 0, L.kkt_last_error(h)
    return h


def _run(prec, max_iter=30, tol=1e-10):
    L = lib()
    dev = torch.device("cuda:0")
    syss = _systems()
    ha, hb, dv, db, xa, xb = [], [], [], [], [], []
    try:
        for ro, ci, va, v1, b in syss:
            ha.append(_handle(L, ro, ci, va, prec))
            hb.append(_handle(L, ro, ci, va, prec))
            dv.append(torch.from_numpy(v1).to(dev))
            db.append(torch.from_numpy(b).to(dev))
            xa.append(torch.zeros(len(b), dtype=torch.float64, device=dev))
            xb.append(torch.full((len(b),), 7.0, dtype=torch.float64, device=dev))
        # the library's streams do not wait on torch's: the copies and fills must be done
        torch.cuda.synchronize()
        single = [L.kkt_iterative_solve_device(ha[q], dv[q].data_ptr(), db[q].data_ptr(), xa[q].data_ptr(),
                                               tol, max_iter) for q in range(len(syss))]
        grp = IterativeGroup()
        try:
            rc, st = grp.solve(hb, [t.data_ptr() for t in dv], [t.data_ptr() for t in db],
                               [t.data_ptr() for t in xb], tol, max_iter)
        finally:
            grp.dispose()
        torch.cuda.synchronize()
        return syss, single, rc, st, [t.cpu().numpy() for t in xa], [t.cpu().numpy() for t in xb], ha, hb
    except Exception:
        for h in ha + hb:
            L.kkt_destroy(h)
        raise


def _free(ha, hb):
    L = lib()
    for h in ha + hb:
        L.kkt_destroy(h)


@pytest.mark.parametrize("prec", [KKT_PREC_FP32_REFINE, KKT_PREC_FP64])
def test_group_equals_single_handle_solves(gpu, prec):
    _say("group iterative solve vs single handles, precision %d" % prec)
    syss, single, rc, st, xa, xb, ha, hb = _run(prec)
    try:
        assert rc == 0
        assert list(st) == single, (list(st), single)
        assert len(set(single)) > 1  # the handles leave the group at different iterations
        for q, (ro, ci, va, v1, b) in enumerate(syss):
            assert single[q] > 0
            assert np.array_equal(xa[q], xb[q]), q
            n = len(ro) - 1
            U = sp.csr_matrix((v1, ci, ro), shape=(n, n))
            A = U + sp.triu(U, 1).T
            assert np.linalg.norm(A @ xb[q] - b) / np.linalg.norm(b) <= 1e-9
    finally:
        _free(ha, hb)


def test_group_iteration_limit_matches_single(gpu):
    _say("group iterative solve, iteration limit 2")
    syss, single, rc, st, xa, xb, ha, hb = _run(KKT_PREC_FP32_REFINE, max_iter=2)
    try:
        assert rc == 0
        assert list(st) == single
        assert any(s == -13 for s in single) and any(s > 0 for s in single), single
        L = lib()
        for q in range(len(syss)):
            if single[q] > 0:
                assert np.array_equal(xa[q], xb[q])
            else:
                assert L.kkt_last_error(hb[q]) == L.kkt_last_error(ha[q])
    finally:
        _free(ha, hb)


def test_group_argument_errors(gpu):
    _say("group iterative solve, argument errors")
    L = lib()
    dev = torch.device("cuda:0")
    syss = _systems()[:2]
    hs = [_handle(L, ro, ci, va, p) for (ro, ci, va, v1, b), p in zip(syss, [KKT_PREC_FP64, KKT_PREC_FP32_REFINE])]
    unf = L.kkt_create()
    grp = IterativeGroup()
    try:
        dv = [torch.from_numpy(s[3]).to(dev) for s in syss]
        db = [torch.from_numpy(s[4]).to(dev) for s in syss]
        dx = [torch.zeros(len(s[4]), dtype=torch.float64, device=dev) for s in syss]
        torch.cuda.synchronize()
        args = ([t.data_ptr() for t in dv], [t.data_ptr() for t in db], [t.data_ptr() for t in dx])
        # an FP64 and an FP32 handle: two lockstep groups, each as its own solve
        ref = [L.kkt_iterative_solve_device(hs[q], args[0][q], args[1][q], args[2][q], 1e-10, 30) for q in range(2)]
        torch.cuda.synchronize()
        x_ref = [t.cpu().numpy() for t in dx]
        for t in dx:
            t.zero_()
        torch.cuda.synchronize()
        rc, st = grp.solve(hs, *args, 1e-10, 30)
        assert rc == 0 and list(st) == ref and min(ref) > 0, (rc, list(st), ref)
        for q in range(2):
            assert np.array_equal(dx[q].cpu().numpy(), x_ref[q]), q
        rc, _ = grp.solve([hs[0], hs[0]], [args[0][0]] * 2, [args[1][0]] * 2, [args[2][0]] * 2, 1e-10, 30)
        assert rc == -14 and "twice" in grp.last_error()
        rc, _ = grp.solve([unf], args[0][:1], args[1][:1], args[2][:1], 1e-10, 30)
        assert rc == -8
        rc, _ = grp.solve([hs[0]], args[0][:1], args[1][:1], args[2][:1], 1e-10, 0)
        assert rc == -14
    finally:
        grp.dispose()
        L.kkt_destroy(unf)
        for h in hs:
            L.kkt_destroy(h)


def test_serving_group_runners_equal_workers(gpu):
    # the serving loop with group runners (device requests of factored instances
    # solved together) gives bitwise the results of the loop without them: the
    # same step kinds, iteration counts and solutions per instance and step; with
    # max_iter 3 some group solves do not converge and their requests are
    # refactored and solved directly by a worker
    _say("serving: group runners vs workers only")
    from artisynth_core_amd.serving import SolveServer
    dev = torch.device("cuda:0")
    kinds = ("beam", "tet", "rigid")
    inst = [G.serving_instance(kinds[i % 3], 700 + 400 * i, 300 + i) for i in range(9)]
    steps = 4
    seq = {i: [(G.perturbed_values(inst[i][2], s_, 40 + i, eps=5e-3), inst[i][3]) for s_ in range(steps)]
           for i in range(len(inst))}
    dv = {(i, s_): torch.from_numpy(np.ascontiguousarray(seq[i][s_][0])).to(dev)
          for i in range(len(inst)) for s_ in range(steps)}
    db = {i: torch.from_numpy(np.ascontiguousarray(inst[i][3], np.float64)).to(dev) for i in range(len(inst))}
    torch.cuda.synchronize()
    out = {}
    for runners in (0, 2):
        dx = {k: torch.zeros(len(inst[k[0]][3]), dtype=torch.float64, device=dev) for k in dv}
        torch.cuda.synchronize()
        srv = SolveServer(nworkers=3, max_iter=3, group_runners=runners, group_max=8)
        try:
            for i in range(len(inst)):
                srv.register(i, inst[i][0], inst[i][1], inst[i][2])
            rs = []
            for s_ in range(steps):
                for i in range(len(inst)):
                    rs.append(((i, s_), srv.submit_device(i, dv[(i, s_)].data_ptr(), db[i].data_ptr(),
                                                          dx[(i, s_)].data_ptr())))
            res = {}
            for k, r in rs:
                r.wait()
                assert r.error is None, (k, r.error)
                res[k] = (r.kind, r.iters)
            st = srv.stats()
        finally:
            srv.shutdown()
        torch.cuda.synchronize()
        out[runners] = (res, {k: t.cpu().numpy() for k, t in dx.items()}, st)
    (r0, x0, st0), (r2, x2, st2) = out[0], out[2]
    assert st0["group_calls"] == 0
    assert st2["group_calls"] > 0 and st2["group_requests"] >= sum(1 for k in r2 if r2[k][0] == "iterative")
    later = [r2[k][0] for k in r2 if k[1] > 0]
    assert "iterative" in later and "direct" in later, later  # both outcomes of a group solve occur
    for k in r0:
        assert r0[k] == r2[k], (k, r0[k], r2[k])
        assert np.array_equal(x0[k], x2[k]), k
        i, s_ = k
        ro, ci = inst[i][0], inst[i][1]
        n = len(ro) - 1
        U = sp.csr_matrix((seq[i][s_][0], ci, ro), shape=(n, n))
        A = U + sp.triu(U, 1).T
        b = inst[i][3]
        assert np.linalg.norm(A @ x2[k] - b) / np.linalg.norm(b) <= 1e-8, k


@pytest.mark.parametrize("prec", [KKT_PREC_FP32_REFINE, KKT_PREC_FP64])
def test_release_factor_keeps_the_analysis(gpu, prec):
    # kkt_release_factor frees the factor's device memory and leaves the handle
    # analyzed: solves are refused until the next factor, which needs no new
    # analysis and gives bitwise the solution the first factor gave
    _say("release factor, precision %d" % prec)
    from artisynth_core_amd._lib import KktInfo
    L = lib()
    ro, ci, va, v1, b = _systems()[3]
    h = _handle(L, ro, ci, va, prec)
    try:
        x1 = np.zeros(len(b))
        x2 = np.zeros(len(b))
        assert L.kkt_solve(h, nptr(b, ctypes.c_double), nptr(x1, ctypes.c_double)) == 0
        before, after = KktInfo(), KktInfo()
        assert L.kkt_get_info(h, ctypes.byref(before)) == 0
        assert L.kkt_release_factor(h) == 0
        assert L.kkt_get_info(h, ctypes.byref(after)) == 0
        assert after.device_bytes < before.device_bytes - after.front_bytes // 2
        assert after.nsuper == before.nsuper and after.nnzL == before.nnzL  # the analysis stays
        assert L.kkt_solve(h, nptr(b, ctypes.c_double), nptr(x2, ctypes.c_double)) == -8  # not factored
        assert L.kkt_factor(h, nptr(va, ctypes.c_double)) == 0, L.kkt_last_error(h)
        assert L.kkt_solve(h, nptr(b, ctypes.c_double), nptr(x2, ctypes.c_double)) == 0
        assert np.array_equal(x1, x2)
        assert L.kkt_release_factor(h) == 0 and L.kkt_release_factor(h) == 0  # idempotent
    finally:
        L.kkt_destroy(h)
    u = L.kkt_create()
    try:
        assert L.kkt_release_factor(u) == -8  # no analysis yet
    finally:
        L.kkt_destroy(u)


def test_group_host_arrays_equal_single_handle_solves(gpu):
    # kkt_group_iterative_solve (host arrays, staged through each handle's own
    # device buffers) gives bitwise what kkt_iterative_solve gives per handle
    _say("group iterative solve on host arrays vs single handles")
    L = lib()
    syss = _systems()
    ha = [_handle(L, ro, ci, va, KKT_PREC_FP32_REFINE) for ro, ci, va, v1, b in syss]
    hb = [_handle(L, ro, ci, va, KKT_PREC_FP32_REFINE) for ro, ci, va, v1, b in syss]
    grp = IterativeGroup()
    try:
        xa = [np.zeros(len(s[4])) for s in syss]
        xb = [np.full(len(s[4]), 7.0) for s in syss]
        single = [L.kkt_iterative_solve(ha[q], nptr(s[3], ctypes.c_double), nptr(s[4], ctypes.c_double),
                                        nptr(xa[q], ctypes.c_double), 1e-10, 30) for q, s in enumerate(syss)]
        rc, st = grp.solve_host(hb, [s[3] for s in syss], [s[4] for s in syss], xb, 1e-10, 30)
        assert rc == 0, grp.last_error()
        assert list(st) == single and min(single) > 0, (list(st), single)
        for q in range(len(syss)):
            assert np.array_equal(xa[q], xb[q]), q
        # an iteration limit: the handles that stop keep their x untouched, as the single call does
        xc = [np.full(len(s[4]), 7.0) for s in syss]
        rc, st = grp.solve_host(hb, [s[3] for s in syss], [s[4] for s in syss], xc, 1e-10, 2)
        assert rc == 0
        for q in range(len(syss)):
            if st[q] <= 0:
                assert np.all(xc[q] == 7.0)
    finally:
        grp.dispose()
        for h in ha + hb:
            L.kkt_destroy(h)


def test_group_of_kkt_group_max_handles(gpu):
    # a full group (KKT_GROUP_MAX = 128 handles, the reductions' launches then
    # hold 256 segments, the tables' limit): every handle's solution and count
    # bitwise its own iterative solve's
    _say("group iterative solve with 128 handles")
    L = lib()
    dev = torch.device("cuda:0")
    kinds = ("beam", "tet", "rigid")
    hs, dv, db, dx, single, xs = [], [], [], [], [], []
    try:
        for q in range(128):
            ro, ci, va, b = G.serving_instance(kinds[q % 3], 150 + 7 * q, 5000 + q)
            ro, ci, va, b = (np.asarray(ro, np.int32), np.asarray(ci, np.int32), np.asarray(va, np.float64),
                             np.asarray(b, np.float64))
            v1 = va * (1.0 + 2e-3 * np.random.default_rng(q).uniform(-1, 1, va.shape))
            hs.append(_handle(L, ro, ci, va, KKT_PREC_FP32_REFINE))
            dv.append(torch.from_numpy(v1).to(dev))
            db.append(torch.from_numpy(b).to(dev))
            dx.append(torch.zeros(len(b), dtype=torch.float64, device=dev))
        torch.cuda.synchronize()
        for q in range(128):
            single.append(L.kkt_iterative_solve_device(hs[q], dv[q].data_ptr(), db[q].data_ptr(), dx[q].data_ptr(),
                                                       1e-10, 30))
        torch.cuda.synchronize()
        xs = [t.cpu().numpy() for t in dx]
        for t in dx:
            t.fill_(3.0)
        torch.cuda.synchronize()
        grp = IterativeGroup()
        try:
            rc, st = grp.solve(hs, [t.data_ptr() for t in dv], [t.data_ptr() for t in db], [t.data_ptr() for t in dx],
                               1e-10, 30)
            assert rc == 0, grp.last_error()
        finally:
            grp.dispose()
        torch.cuda.synchronize()
        assert list(st) == single and min(single) > 0
        for q in range(128):
            assert np.array_equal(dx[q].cpu().numpy(), xs[q]), q
    finally:
        for h in hs:
            L.kkt_destroy(h)
