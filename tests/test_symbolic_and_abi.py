"""CPU tests: host symbolic analysis (bit-exact against the oracle's independent
symbolic elimination + invariants), the C-ABI library surface, and the host logic
of the DirectSolver / KKTSolver mirror."""
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from artisynth_core_amd import generators as G
from artisynth_core_amd import _lib
from artisynth_core_amd.solvers import (KKTSolver, MI355XSolver, Matrix, NumericalException,
                                        kkt_upper_csr)
from kkt_helpers import P5_COLS, P5_ROWS, P5_VALS
from symbolic_helpers import product_symbolic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _cases():
    beam, vals, b = G.small_beam(8)
    yield "beam8", beam.rowOffs, beam.colIdxs, vals
    beam, vals, b = G.small_beam(26)
    yield "beam26", beam.rowOffs, beam.colIdxs, vals
    d = np.load(os.path.join(ROOT, "tests", "golden", "testMatrix_upper.npz"))
    yield "testMatrix", d["rowOffs"], d["colIdxs"], d["vals"]
    yield "pardiso5", P5_ROWS, P5_COLS, P5_VALS[0][0]
    beam = G.HexBeam(12, 5, 4, 1.0, 0.2)
    vals, b = beam.system(np.zeros((beam.nn, 3)), np.zeros((beam.nn, 3)), beam.gravity_loads())
    yield "beam12x5x4", beam.rowOffs, beam.colIdxs, vals


@pytest.mark.parametrize("case", list(_cases()), ids=lambda c: c[0])
@pytest.mark.parametrize("ordering", [1, 0])
def test_symbolic_matches_oracle(case, ordering):
    name, ro, ci, va = case
    n = len(ro) - 1
    S = product_symbolic(ro, ci, va, ordering=ordering)
    perm = S["perm"]
    # permutation is a bijection
    assert np.array_equal(np.sort(perm), np.arange(n))
    # elimination tree and column counts: bit-exact against symbolic elimination
    par, cc = O.symbolic(ro, ci, perm)
    assert np.array_equal(par, S["parent"])
    assert np.array_equal(cc, S["colcount"])
    assert S["info"]["nnzL_exact"] == int(cc.sum())
    assert np.all((S["parent"] == -1) | (S["parent"] > np.arange(n)))
    # supernode row structures: bit-exact against the oracle's union construction
    rp, rows = O.supernode_rows(ro, ci, perm, S["sup_ptr"])
    assert np.array_equal(rp, S["rows_ptr"])
    assert np.array_equal(rows, S["rows"])
    sup = S["sup_ptr"]
    ns = len(sup) - 1
    assert sup[0] == 0 and sup[-1] == n and np.all(np.diff(sup) > 0) and np.all(np.diff(sup) <= 2048)
    # nested structure: child rows outside the parent's columns lie in the parent's rows
    for s in range(ns):
        p = S["sparent"][s]
        r = rows[rp[s]:rp[s + 1]]
        if p < 0:
            assert len(r) == 0
            continue
        assert p > s and sup[p] <= r[0] < sup[p + 1]
        outside = r[r >= sup[p + 1]]
        assert np.all(np.isin(outside, rows[rp[p]:rp[p + 1]]))
        assert S["level"][p] > S["level"][s]


def test_delayed_rows_follow_their_neighbours():
    beam, vals, b = G.small_beam(26)
    S = product_symbolic(beam.rowOffs, beam.colIdxs, vals)
    assert S["info"]["ndelayed"] == beam.numG
    iperm = np.empty(beam.n, np.int64)
    iperm[S["perm"]] = np.arange(beam.n)
    A = sp.csr_matrix((vals, beam.colIdxs, beam.rowOffs), shape=(beam.n, beam.n))
    A = A + A.T
    for lam in range(beam.sizeM, beam.n):
        nbrs = A.indices[A.indptr[lam]:A.indptr[lam + 1]]
        nbrs = nbrs[nbrs != lam]
        assert np.all(iperm[nbrs] < iperm[lam])


def test_nested_dissection_reduces_fill():
    beam = G.HexBeam(20, 6, 6, 1.0, 0.2)
    vals, b = beam.system(np.zeros((beam.nn, 3)), np.zeros((beam.nn, 3)), beam.gravity_loads())
    nd = product_symbolic(beam.rowOffs, beam.colIdxs, vals, ordering=1)["info"]
    nat = product_symbolic(beam.rowOffs, beam.colIdxs, vals, ordering=0)["info"]
    assert nd["factor_flops"] < 0.5 * nat["factor_flops"]


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "kkt_mi355x.h")).read()
    declared = set(re.findall(r"\b(kkt_[a-z0-9_]+)\s*\(", hdr))
    L = _lib.lib()
    missing = [s for s in sorted(declared) if not hasattr(L, s)]
    assert not missing, missing
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert declared == bound
    assert L.kkt_version().decode().startswith("kkt_mi355x")


def test_abi_argument_and_state_errors_without_gpu():
    L = _lib.lib()
    h = L.kkt_create()
    assert h
    try:
        assert L.kkt_analyze(h) == -8  # STATE: analyze before setPattern
        assert b"setPattern" in L.kkt_last_error(h)
        ro = np.array([0, 1, 2], np.int32)
        ci = np.array([0, 3], np.int32)  # column out of range
        rc = L.kkt_set_pattern(h, 2, 2, _lib.ptr(ro, _lib.ctypes.c_int), _lib.ptr(ci, _lib.ctypes.c_int), 1)
        assert rc == -4
        ci = np.array([0, 1], np.int32)
        assert L.kkt_set_pattern(h, 2, 2, _lib.ptr(ro, _lib.ctypes.c_int),
                                 _lib.ptr(ci, _lib.ctypes.c_int), 0) == -9  # GENERAL rejected
        assert L.kkt_set_pattern(h, 2, 2, _lib.ptr(ro, _lib.ctypes.c_int),
                                 _lib.ptr(ci, _lib.ctypes.c_int), 1) == 0
        assert L.kkt_last_error(h) is None
        assert L.kkt_factor(h, _lib.ptr(np.ones(2), _lib.ctypes.c_double)) == -8
        assert L.kkt_set_precision(h, 7) == -14
        assert L.kkt_set_refine_tol(h, -1.0, 3) == -14
        assert L.kkt_set_ir_steps(h, -1) == -14
        forms = np.ones(2, np.int32)
        assert L.kkt_set_forward_forms(h, _lib.ptr(forms, _lib.ctypes.c_int)) == -8  # STATE: no analysis
        assert L.kkt_get_forward_forms(h, _lib.ptr(forms, _lib.ctypes.c_int)) == -8
    finally:
        L.kkt_destroy(h)
    L.kkt_destroy(None)


def test_batch_abi_argument_errors_without_gpu():
    L = _lib.lib()
    b = L.kkt_batch_create()
    assert b
    I = _lib.ctypes.c_int
    try:
        assert L.kkt_batch_analyze(b, None) == -8  # STATE: no systems
        sn = np.array([2, 1], np.int32)
        ro = np.array([0, 1, 2, 0, 1], np.int32)
        ci = np.array([0, 1, 0], np.int32)
        assert L.kkt_batch_set_systems(b, 0, _lib.ptr(sn, I), _lib.ptr(ro, I), _lib.ptr(ci, I), 1) == -14
        bad = np.array([0, 1, 2, 0, 1], np.int32)
        badc = np.array([0, 5, 0], np.int32)  # column outside system 0
        assert L.kkt_batch_set_systems(b, 2, _lib.ptr(sn, I), _lib.ptr(bad, I), _lib.ptr(badc, I), 1) == -4
        assert b"system 0" in L.kkt_batch_last_error(b)
        assert L.kkt_batch_set_systems(b, 2, _lib.ptr(sn, I), _lib.ptr(ro, I), _lib.ptr(ci, I), 1) == 0
        assert L.kkt_batch_last_error(b) is None
        assert L.kkt_batch_set_precision(b, 9) == -14
        si = _lib.KktBatchSysInfo()
        assert L.kkt_batch_system_info(b, 5, _lib.ctypes.byref(si)) == -14
    finally:
        L.kkt_batch_destroy(b)
    L.kkt_batch_destroy(None)


def test_direct_solver_mirror_state_machine():
    s = MI355XSolver()
    from artisynth_core_amd.solvers import ImproperStateException
    with pytest.raises(ImproperStateException):
        s.factor(np.ones(3))
    with pytest.raises(ImproperStateException):
        s.solve(np.zeros(3), np.ones(3))
    assert s.iterativeSolve(np.ones(3), np.zeros(3), np.ones(3), 10) == -1
    assert s.hasAutoIterativeSolving()
    s.dispose()
    s.dispose()
    with pytest.raises(ImproperStateException):
        s.setIterativeRefinementSteps(1)


def test_kkt_csr_layout_matches_kktsolver():
    # KKTSolver.java:171-454: rows < sizeM hold M's upper entries then G^T's entries
    # (column sizeM+j); row sizeM+i holds only -Rg[i]
    M = sp.csr_matrix(np.array([[4.0, 1, 0], [1, 3, 2], [0, 2, 5]]))
    GT = sp.csr_matrix(np.array([[1.0, 0], [0, 1], [1, 1]]))
    Rg = np.array([0.1, 0.2])
    ro, ci, va = kkt_upper_csr(M, GT, Rg)
    assert list(ro) == [0, 3, 6, 9, 10, 11]
    assert list(ci) == [0, 1, 3, 1, 2, 4, 2, 3, 4, 3, 4]
    assert list(va) == [4, 1, 1, 3, 2, 1, 5, 1, 1, -0.1, -0.2]
    ro, ci, va = kkt_upper_csr(M, GT, None)
    assert va[-2:].tolist() == [0.0, 0.0]


def test_estimate_optimal_count():
    # KKTSolver.java:889-922
    k = KKTSolver.__new__(KKTSolver)
    k.directTimeMsec, k.directCnt = 100.0, 1
    k.firstIterativeTimeMsec, k.iterativeTimeMsec, k.iterativeCnt = 10.0, 10.0, 1
    assert k.estimateOptimalCount() == 1 + 19  # dt <= 0
    k.iterativeTimeMsec, k.iterativeCnt = 10.0 + 14.0, 2  # it0=10, avg 12 -> dt = 2
    assert k.estimateOptimalCount() == int(np.ceil(np.sqrt(2 * (100 - 10 + 2) / 2)))
    k.iterativeTimeMsec, k.iterativeCnt = 300.0, 2
    assert k.estimateOptimalCount() == 3


def test_bench_policy_matches_kktsolver_mirror():
    # the bench sequence's step choice (bench.FactorAndSolvePolicy) follows
    # KKTSolver.factorAndSolve (KKTSolver.java:930-1014): fed the same step times
    # and outcomes, it chooses what the mirrored KKTSolver state would choose,
    # step by step, including a failed iterative attempt that forces a factor
    import bench
    rng = np.random.default_rng(5)
    pol = bench.FactorAndSolvePolicy()
    k = KKTSolver.__new__(KKTSolver)
    k.directTimeMsec, k.directCnt = 0.0, 0
    k.iterativeTimeMsec, k.firstIterativeTimeMsec, k.iterativeCnt = 0.0, 0.0, 0
    kinds = []
    for step in range(60):
        mirror = k.directCnt > 0 and (k.iterativeCnt == 0 or k.iterativeCnt + 1 < k.estimateOptimalCount())
        assert pol.try_iterative() == mirror, step
        converged = mirror and rng.uniform() > 0.1
        if converged:
            ms = 2.0 + 1.5 * k.iterativeCnt + rng.uniform(0, 0.5)  # cost grows after a factor
            pol.iterative_done(ms)
            k.iterativeTimeMsec += ms
            if k.iterativeCnt == 0:
                k.firstIterativeTimeMsec = k.iterativeTimeMsec
            k.iterativeCnt += 1
            kinds.append("i")
        else:
            ms = 16.0 + rng.uniform(0, 1.0)
            pol.direct_done(ms)
            k.directTimeMsec += ms
            k.directCnt += 1
            k.iterativeCnt, k.iterativeTimeMsec = 0, 0.0
            kinds.append("D")
    runs = "".join(kinds).split("D")[1:]
    assert kinds[0] == "D" and max(len(r) for r in runs) < 19  # refactors before 19 iterative steps
    assert sum(1 for r in runs if len(r) >= 3) >= 3  # and reuses factors over several steps


def test_mooney_rivlin_tangent_is_the_stress_derivative():
    # the precision workload's material: A = dP/dF by central differences, symmetric
    # in (iJ)<->(kL), and linear elasticity (mu = 2 c10, lambda = kappa - 2mu/3) at F = I
    rng = np.random.default_rng(3)
    F = np.eye(3) + 0.1 * rng.standard_normal((6, 3, 3))
    c10, kappa = 1e3, 1e5
    P, A = G.mooney_rivlin_PA(F, c10, kappa)
    eps = 1e-6
    for k in range(3):
        for L in range(3):
            d = np.zeros((3, 3))
            d[k, L] = eps
            Pp, _ = G.mooney_rivlin_PA(F + d, c10, kappa)
            Pm, _ = G.mooney_rivlin_PA(F - d, c10, kappa)
            fd = (Pp - Pm) / (2 * eps)
            assert np.abs(fd - A[:, :, :, k, L]).max() <= 1e-7 * np.abs(A).max()
    assert np.allclose(A, np.transpose(A, (0, 3, 4, 1, 2)), rtol=0, atol=1e-9 * np.abs(A).max())
    P0, A0 = G.mooney_rivlin_PA(np.eye(3)[None], c10, kappa)
    mu, lam = 2 * c10, kappa - 2 * (2 * c10) / 3
    I3 = np.eye(3)
    lin = (mu * (np.einsum("ik,JL->iJkL", I3, I3) + np.einsum("iL,Jk->iJkL", I3, I3)) +
           lam * np.einsum("iJ,kL->iJkL", I3, I3))
    assert np.abs(P0).max() < 1e-9 and np.allclose(A0[0], lin, rtol=0, atol=1e-9 * kappa)


def test_beam_dynamics_integrates_the_constrained_step():
    # the integrated ConstrainedBackwardEuler sequence of the benchmark, on a small
    # beam with the oracle as the FP64 solve: the attached end stays fixed (G v = 0),
    # the beam falls under gravity, each step's right-hand side is M v + h (fext -
    # fint) (b = M v + h f plus the h D v Jacobian term) at the state the previous
    # steps produced
    import oracle as Orc
    states = []

    def solve(beam, vals, b):
        S = product_symbolic(beam.rowOffs, beam.colIdxs, vals)
        x = Orc.Factor(beam.rowOffs, beam.colIdxs, vals, perm=S["perm"], sup_ptr=S["sup_ptr"]).solve(b)
        states.append(x)
        return x

    beam, V, B, tip = G.large_beam_dynamics(solve, nsteps=4, nx=8, ny=2, nz=2, width=0.2, tip_scale=0.0)
    assert len(V) == len(B) == len(tip) == 4
    for x in states:
        v = x[:beam.sizeM].reshape(-1, 3)
        assert np.abs(v[beam.attached]).max() <= 1e-12
    assert tip[0] < tip[1] < tip[2] < tip[3]
    # step 0 from rest: b = h * fext (gravity only), the lambda rows zero
    assert np.allclose(B[0][:beam.sizeM], beam.h * beam.gravity_loads().reshape(-1), rtol=0, atol=1e-15)
    assert not B[0][beam.sizeM:].any()


def test_serving_loop_mechanics_without_gpu():
    # the serving loop's host side: configuration checks, registration errors,
    # per-instance order and error reporting (without a GPU every admission fails
    # with KKT_ERR_DEVICE, which each request reports, and the loop keeps going)
    import ctypes
    L = _lib.lib()
    cfg = _lib.KktServeConfig()
    L.kkt_serve_default_config(ctypes.byref(cfg))
    assert (cfg.budget_bytes, cfg.nworkers, cfg.max_iter, cfg.tol_exp) == (200 << 30, 12, 30, 10)
    assert (cfg.group_runners, cfg.group_max) == (2, 64)
    for field, value in (("nworkers", 0), ("group_runners", -1), ("group_runners", 9), ("group_max", 0),
                         ("group_max", 129)):
        bad = _lib.KktServeConfig.from_buffer_copy(cfg)
        setattr(bad, field, value)
        assert not L.kkt_serve_create(ctypes.byref(bad)), field
    cfg.nworkers = 2
    s = L.kkt_serve_create(ctypes.byref(cfg))
    assert s
    try:
        ro = np.array([0, 2, 3], np.int32)
        ci = np.array([0, 1, 1], np.int32)
        va = np.array([2.0, 1.0, 3.0])
        I, D = ctypes.c_int, ctypes.c_double
        assert L.kkt_serve_register(s, 7, 2, 3, _lib.ptr(ro, I), _lib.ptr(ci, I), _lib.ptr(va, D), 1) == 0
        assert L.kkt_serve_register(s, 7, 2, 3, _lib.ptr(ro, I), _lib.ptr(ci, I), _lib.ptr(va, D), 1) == -14
        assert b"already registered" in L.kkt_serve_last_error(s)
        assert L.kkt_serve_register(s, 8, 2, 3, _lib.ptr(ro, I), _lib.ptr(ci, I), _lib.ptr(va, D), 0) == -9
        b = np.ones(2)
        x = np.zeros(2)
        assert L.kkt_serve_submit(s, 99, _lib.ptr(va, D), _lib.ptr(b, D), _lib.ptr(x, D)) == -14
        ids = [L.kkt_serve_submit(s, 7, _lib.ptr(va, D), _lib.ptr(b, D), _lib.ptr(x, D)) for _ in range(3)]
        assert all(i > 0 for i in ids) and ids == sorted(ids)
        for i in ids:
            res = _lib.KktServeResult()
            rc = L.kkt_serve_wait(s, i, ctypes.byref(res))
            assert rc == res.status and rc in (-15, -1, -2)  # no device: admission fails
            assert b"admission of instance 7 failed" in res.message
            assert res.submit_ms <= res.start_ms <= res.done_ms
        assert L.kkt_serve_wait(s, ids[0], None) == -14  # each request is waited for once
        st = _lib.KktServeStats()
        L.kkt_serve_get_stats(s, ctypes.byref(st))
        assert (st.submitted, st.completed, st.errors, st.admissions) == (3, 3, 3, 0)
        assert (st.group_calls, st.group_requests) == (0, 0)  # no device: no group object
    finally:
        L.kkt_serve_destroy(s)
    L.kkt_serve_destroy(None)
