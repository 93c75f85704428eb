"""GPU tests of the solver contract around the factor: the FP32-stall -> FP64
refactor fallback and its error code, the BackwardEuler Matrix-object path
(analyze(Matrix) + autoFactorAndSolve on an SPD system without constraints),
and the full 256-system batch of the batching workload against the oracle,
including systems of very different scale in one batch."""
import concurrent.futures as cf
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from artisynth_core_amd import KKT_PREC_FP32_REFINE, KKT_PREC_FP64
from artisynth_core_amd import generators as G
from artisynth_core_amd._lib import ERRORS
from artisynth_core_amd.solvers import MI355XSolver, Matrix, NumericalException
from kkt_helpers import full_from_upper, rel_err, rel_residual
from symbolic_helpers import product_symbolic

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


def _oracle_solve(ro, ci, va, b):
    S = product_symbolic(ro, ci, va)
    return O.Factor(ro, ci, va, perm=S["perm"], sup_ptr=S["sup_ptr"]).solve(b)


def test_fp32_stall_falls_back_to_fp64(gpu):
    # KKTSolver's refactor-on-failure contract (KKTSolver.java:981-1013) inside the
    # solve: FP32 refinement that cannot reach the tolerance in one step stalls,
    # the handle refactors in FP64 with the same values and refines to the
    # tolerance; a tolerance not even FP64 reaches returns KKT_ERR_REFINE_STALL
    _say("fp32 stall fallback")
    beam, vals, b = G.small_beam(26)
    s = MI355XSolver(precision=KKT_PREC_FP32_REFINE)
    try:
        s.analyzeArrays(vals, beam.colIdxs, beam.rowOffs, beam.n, Matrix.SYMMETRIC)
        s.setRefinement(1e-13, 1)
        s.factor(vals)
        assert s.info()["precision"] == KKT_PREC_FP32_REFINE
        x = np.zeros(beam.n)
        s.solve(x, b)
        inf = s.info()
        assert inf["fp64_fallbacks"] == 1 and inf["precision"] == KKT_PREC_FP64
        assert inf["last_rel_residual"] <= 1e-13
        assert rel_residual(beam.rowOffs, beam.colIdxs, vals, x, b) <= 1e-12
        # the FP64 factor stays (no second fallback) and still refines to the tolerance
        x2 = np.zeros(beam.n)
        s.solve(x2, 2.0 * b)
        assert s.info()["fp64_fallbacks"] == 1
        assert rel_residual(beam.rowOffs, beam.colIdxs, vals, x2, 2.0 * b) <= 1e-12
        # multi-RHS: a new FP32 factor, the block stalls, one FP64 refactor for all columns
        s.factor(vals)
        assert s.info()["precision"] == KKT_PREC_FP32_REFINE
        B = np.concatenate([b, -b, 3.0 * b])
        X = np.zeros(3 * beam.n)
        s.solve(X, B, 3)
        assert s.info()["fp64_fallbacks"] == 2
        for j in range(3):
            xj, bj = X[j * beam.n:(j + 1) * beam.n], B[j * beam.n:(j + 1) * beam.n]
            assert rel_residual(beam.rowOffs, beam.colIdxs, vals, xj, bj) <= 1e-12
        # unreachable tolerance: the FP64 refactor also stalls -> REFINE_STALL
        s.factor(vals)
        s.setRefinement(1e-30, 2)
        with pytest.raises(NumericalException) as ei:
            s.solve(np.zeros(beam.n), b)
        assert "status -16" in str(ei.value) and ERRORS[-16] == "REFINE_STALL"
        assert "refinement stalled" in s.getLastErrorMessage()
        assert s.info()["fp64_fallbacks"] == 3
        # the handle keeps working after the error
        s.setRefinement(1e-10, 10)
        s.factor(vals)
        x3 = np.zeros(beam.n)
        s.solve(x3, b)
        assert rel_residual(beam.rowOffs, beam.colIdxs, vals, x3, b) <= 1e-10
        assert s.info()["fp64_fallbacks"] == 3
    finally:
        s.dispose()


@pytest.mark.parametrize("prec", [KKT_PREC_FP64, KKT_PREC_FP32_REFINE])
def test_backward_euler_matrix_path(gpu, prec):
    # MechSystemSolver.doDirectSolve (MechSystemSolver.java:2113-2121) on the
    # DirectSolver Matrix-object entry points (CuDssSolver.java:310-369, 866-913):
    # an SPD S = M + h D + h^2 K without constraints, analyze(M, size, SPD), then
    # autoFactorAndSolve with the stale factor (BiCGStab branch), with tolExp = 0
    # (factor branch), and after the matrix values change again
    _say("backward euler Matrix path prec=%d" % prec)
    beam = G.HexBeam(24, 8, 8, 1.0, 0.2, material=("neohookean", 3.4e4, 1.7e6), h=0.005)
    rng = np.random.default_rng(21)
    u = rng.uniform(-1e-4, 1e-4, (beam.nn, 3))
    vals, bk = beam.system(u, np.zeros_like(u), beam.gravity_loads())
    m = beam.sizeM
    A = full_from_upper(beam.rowOffs, beam.colIdxs, vals)
    S = sp.csr_matrix(A[:m, :m])
    S.sort_indices()
    b = bk[:m].copy()
    s = MI355XSolver(precision=prec)
    if prec == KKT_PREC_FP64:
        s.setIterativeRefinementSteps(1)
    try:
        s.analyze(S, m, Matrix.SPD)
        s.factor()
        x = np.zeros(m)
        s.solve(x, b)
        U = sp.triu(S, format="csr")
        U.sort_indices()
        ro, ci, va = U.indptr.astype(np.int32), U.indices.astype(np.int32), U.data
        xo = _oracle_solve(ro, ci, va, b)
        assert rel_err(x, xo) <= 1e-8
        assert np.linalg.norm(S @ x - b) <= 1e-8 * np.linalg.norm(b)
        assert s.info()["nneg"] == 0 and s.info()["npos"] == m
        # values change in place in the caller's matrix object (same pattern, a
        # perturbation symmetric in (i, j))
        rows = np.repeat(np.arange(m), np.diff(S.indptr))
        S.data *= 1.0 + 2e-3 * np.sin(0.37 * (rows + S.indices))
        x2 = np.zeros(m)
        s.autoFactorAndSolve(x2, b, 10)
        assert s.lastAutoIterations > 0  # BiCGStab with the stale factor
        assert np.linalg.norm(S @ x2 - b) <= 1e-9 * np.linalg.norm(b)
        x3 = np.zeros(m)
        s.autoFactorAndSolve(x3, b, 0)  # tolExp <= 0: factor + solve
        assert s.lastAutoIterations == 0
        assert np.linalg.norm(S @ x3 - b) <= 1e-8 * np.linalg.norm(b)
        U = sp.triu(S, format="csr")
        U.sort_indices()
        xo3 = _oracle_solve(U.indptr.astype(np.int32), U.indices.astype(np.int32), U.data, b)
        assert rel_err(x3, xo3) <= 1e-8
        assert rel_err(x2, xo3) <= 1e-8
    finally:
        s.dispose()


_BATCH = {}


def _batch_mix_256():
    # generated once per test process (about 15 s on the host), with the
    # oracle's solution of every system alone (threads: the oracle's C calls
    # release the interpreter lock)
    if not _BATCH:
        systems = G.batch_mix(256)
        with cf.ThreadPoolExecutor(8) as ex:
            xos = list(ex.map(lambda sd: _oracle_solve(*sd), systems))
        _BATCH["systems"], _BATCH["xo"] = systems, xos
    return _BATCH["systems"], _BATCH["xo"]


@pytest.mark.parametrize("prec", [KKT_PREC_FP64, KKT_PREC_FP32_REFINE])
def test_full_batch_mix_vs_oracle(gpu, prec):
    # the batching workload at its full size: 256 uneven systems (DOF U[300, 4000],
    # rigid bodies + tet blocks + 10-200 contact equalities) in one batched factor
    # and solve, every system against the oracle alone; FP32 fronts refine each
    # system to 1e-10 on its own residual
    _say("batch mix 256 systems prec=%d" % prec)
    systems, xos = _batch_mix_256()
    from artisynth_core_amd.batch import BatchedKKT
    B = BatchedKKT(systems, precision=prec, refine_tol=(1e-10, 10))
    try:
        B.factor()
        x = B.solve()
        res = B.per_system_residuals(x, systems)
        errs = [rel_err(B.split(x)[q], xos[q]) for q in range(len(systems))]
        assert max(res) <= (1e-8 if prec == KKT_PREC_FP64 else 1e-10), max(res)
        assert max(errs) <= 1e-8, max(errs)
        if prec == KKT_PREC_FP32_REFINE:
            steps = [B.system_info(q)["refine_steps"] for q in range(len(systems))]
            assert min(steps) >= 1 and B.info()["fp64_fallbacks"] == 0
    finally:
        B.dispose()


@pytest.mark.parametrize("prec", [KKT_PREC_FP64, KKT_PREC_FP32_REFINE])
def test_batch_of_mixed_scales_matches_alone(gpu, prec):
    # systems scaled by 1e-6 and 1e3 in one batch: the pivot perturbation threshold
    # is taken per system (connected component) and the refinement stops per
    # system, so each system's inertia, refinement steps and solution are what it
    # gets alone (FP64: one refinement step each; FP32: refined to 1e-10 each)
    _say("batch of mixed scales prec=%d" % prec)
    base = G.batch_mix(12, seed=91, dof_lo=300, dof_hi=1200)
    systems = []
    for q, (ro, ci, va, b) in enumerate(base):
        f = (1e-6, 1.0, 1e3)[q % 3]
        systems.append((ro, ci, va * f, b * f))
    from artisynth_core_amd.batch import BatchedKKT
    B = BatchedKKT(systems, precision=prec, refine_tol=(1e-10, 10))
    try:
        B.factor()
        xb = B.split(B.solve())
        for q, (ro, ci, va, b) in enumerate(systems):
            s = MI355XSolver(precision=prec)
            s.setIterativeRefinementSteps(1)
            s.setRefinement(1e-10, 10)
            try:
                s.analyzeArrays(va, ci, ro, len(ro) - 1, Matrix.SYMMETRIC)
                s.factor(va)
                xa = np.zeros(len(ro) - 1)
                s.solve(xa, b)
                ia = s.info()
            finally:
                s.dispose()
            si = B.system_info(q)
            assert si["n"] == len(ro) - 1
            assert (si["npos"], si["nneg"], si["nperturbed"]) == (ia["npos"], ia["nneg"], ia["nperturbed"])
            assert rel_residual(ro, ci, va, xb[q], b) <= (1e-8 if prec == KKT_PREC_FP64 else 1e-10)
            if prec == KKT_PREC_FP64:
                assert si["refine_steps"] == 1
                assert rel_err(xb[q], xa) <= 1e-12, q
            else:
                assert si["refine_steps"] == ia["refine_steps"], q
                assert si["last_rel_residual"] <= 1e-10
                assert rel_err(xb[q], xa) <= 1e-9, q
        assert B.info()["fp64_fallbacks"] == 0
    finally:
        B.dispose()


@pytest.mark.parametrize("prec", [KKT_PREC_FP64, KKT_PREC_FP32_REFINE])
def test_max_supernode_setting(gpu, prec):
    # the analysis honours the supernode width limit (more, narrower supernodes
    # and more tree levels at 64 columns than at the default) and the solve on
    # that partition matches the oracle factored on the same partition; widths
    # outside 64..2048 are rejected
    _say("max supernode 64 vs default, prec %d" % prec)
    beam, vals, b = G.small_beam(26)
    ro, ci = beam.rowOffs, beam.colIdxs
    shapes = {}
    for width in (None, 64):
        s = MI355XSolver(precision=prec)
        try:
            if width:
                s.setMaxSupernode(width)
            s.analyzeArrays(vals, ci, ro, beam.n, Matrix.SYMMETRIC)
            s.factor(vals)
            x = np.zeros(beam.n)
            s.solve(x, b)
            sym = s.symbolic()
            k = np.diff(sym["sup_ptr"])
            shapes[width] = (len(k), int(k.max()), s.info()["nlevels"])
            if width:
                assert k.max() <= width
            xo = O.Factor(ro, ci, vals, perm=sym["perm"], sup_ptr=sym["sup_ptr"]).solve(b)
            assert rel_err(x, xo) <= (1e-8 if prec == KKT_PREC_FP64 else 1e-7)
            assert rel_residual(ro, ci, vals, x, b) <= 1e-8
        finally:
            s.dispose()
    assert shapes[64][0] > shapes[None][0] and shapes[64][2] >= shapes[None][2], shapes
    s = MI355XSolver()
    try:
        for bad in (32, 4096):
            with pytest.raises(NumericalException):
                s.setMaxSupernode(bad)
    finally:
        s.dispose()
