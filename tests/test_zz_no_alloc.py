"""GPU test of kkt_solve_device_noalloc, the refined solve that allocates nothing
(its working set lives in the library's static per-device scratch region),
through the C ABI against kkt_solve_device on the same factor: two calls in a
row with the second on a smaller system, then the first system again (the
scratch keeps the previous call's contents), a system too large for the
scratch (rejected with KKT_ERR_ARGUMENT), and a stalled refinement (reported,
no FP64 refactor)."""
import ctypes

import numpy as np
import pytest

from artisynth_core_amd import KKT_PREC_FP32_REFINE, KKT_PREC_FP64
from artisynth_core_amd import generators as G
from artisynth_core_amd._lib import KktInfo, lib
from artisynth_core_amd._lib import ptr as nptr
from kkt_helpers import rel_residual

pytestmark = pytest.mark.gpu


class _Handle:
    def __init__(self, ro, ci, va, prec, ir_steps=1, tol=1e-10):
        import torch
        self.L = lib()
        self.h = self.L.kkt_create()
        assert self.h
        L, h = self.L, self.h
        self.n = len(ro) - 1
        self.ro = np.ascontiguousarray(ro, np.int32)
        self.ci = np.ascontiguousarray(ci, np.int32)
        self.va = np.ascontiguousarray(va, np.float64)
        assert L.kkt_set_precision(h, prec) == 0
        assert L.kkt_set_ir_steps(h, ir_steps) == 0
        assert L.kkt_set_refine_tol(h, tol, 10) == 0
        assert L.kkt_set_pattern(h, self.n, len(self.va), nptr(self.ro, ctypes.c_int), nptr(self.ci, ctypes.c_int),
                                 1) == 0
        assert L.kkt_analyze_values(h, nptr(self.va, ctypes.c_double)) == 0, L.kkt_last_error(h)
        self.dv = torch.from_numpy(self.va).cuda()
        assert L.kkt_factor_device(h, self.dv.data_ptr()) == 0, L.kkt_last_error(h)

    def solve(self, b, noalloc):
        import torch
        db = torch.from_numpy(np.ascontiguousarray(b, np.float64)).cuda()
        dx = torch.full((self.n,), float("nan"), dtype=torch.float64, device="cuda")
        f = self.L.kkt_solve_device_noalloc if noalloc else self.L.kkt_solve_device
        rc = f(self.h, db.data_ptr(), dx.data_ptr())
        torch.cuda.synchronize()
        return rc, dx.cpu().numpy()

    def info(self):
        inf = KktInfo()
        self.L.kkt_get_info(self.h, ctypes.byref(inf))
        return inf.as_dict()

    def close(self):
        self.L.kkt_destroy(self.h)


def _compare(H, b):
    rc, x_ref = H.solve(b, noalloc=False)
    assert rc == 0, H.L.kkt_last_error(H.h)
    steps_ref = H.info()["refine_steps"]
    rc, x = H.solve(b, noalloc=True)
    assert rc == 0, H.L.kkt_last_error(H.h)
    assert H.info()["refine_steps"] == steps_ref
    # the same kernels in the same order (graph replay vs eager launches)
    assert np.linalg.norm(x - x_ref) <= 1e-12 * np.linalg.norm(x_ref)
    assert rel_residual(H.ro, H.ci, H.va, x, b) <= 1e-8
    return x


def test_noalloc_solve_matches_solve_device(gpu):
    print("[gpu-case] noalloc solve: beam 24x8x8 fp32, beam8 fp64, beam 24x8x8 again", flush=True)
    beam = G.HexBeam(24, 8, 8, 1.0, 0.2, material=("neohookean", 3.4e4, 1.7e6), h=0.005)
    rng = np.random.default_rng(5)
    u = rng.uniform(-1e-4, 1e-4, (beam.nn, 3))
    vals, b = beam.system(u, np.zeros_like(u), beam.gravity_loads())
    A = _Handle(beam.rowOffs, beam.colIdxs, vals, KKT_PREC_FP32_REFINE)
    small, svals, sb = G.small_beam(8)
    B = _Handle(small.rowOffs, small.colIdxs, svals, KKT_PREC_FP64, ir_steps=1)
    try:
        xa = _compare(A, b)
        assert A.info()["refine_steps"] >= 1
        xb = _compare(B, sb)  # second call, smaller system
        xa2 = _compare(A, b)  # first system again over the smaller call's leftovers
        assert np.array_equal(xa, xa2)
        assert rel_residual(small.rowOffs, small.colIdxs, svals, xb, sb) <= 1e-8
    finally:
        A.close()
        B.close()


def test_noalloc_solve_rejects_oversize_and_reports_stall(gpu):
    print("[gpu-case] noalloc solve: oversize diagonal n=1.3M, stalled refinement", flush=True)
    n = 1_300_000
    ro = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    va = 1.0 + np.arange(n) % 7
    D = _Handle(ro, ci, va, KKT_PREC_FP64, ir_steps=0)
    try:
        b = np.ones(n)
        rc, _ = D.solve(b, noalloc=True)
        assert rc == -14
        msg = D.L.kkt_last_error(D.h).decode()
        assert "scratch" in msg and "kkt_solve_device" in msg
        rc, x = D.solve(b, noalloc=False)
        assert rc == 0 and np.abs(x - 1.0 / va).max() <= 1e-15
    finally:
        D.close()
    beam, vals, b = G.small_beam(26)
    S = _Handle(beam.rowOffs, beam.colIdxs, vals, KKT_PREC_FP32_REFINE, tol=1e-30)
    try:
        rc, x = S.solve(b, noalloc=True)
        assert rc == -16
        assert "does not refactor" in S.L.kkt_last_error(S.h).decode()
        inf = S.info()
        assert inf["fp64_fallbacks"] == 0 and inf["precision"] == KKT_PREC_FP32_REFINE
        assert rel_residual(beam.rowOffs, beam.colIdxs, vals, x, b) <= 1e-8
    finally:
        S.close()
