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
        rc, _ = grp.solve(hs, *args, 1e-10, 30)
        assert rc == -14 and "precision" in grp.last_error()
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
