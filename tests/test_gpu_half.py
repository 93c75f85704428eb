"""GPU tests of the half-precision copy of an FP32 factor's triangular-solve
arrays, which the stale-factor iterative solve's preconditioner reads:

- one preconditioner application with the copy against the FP32 arrays' (a
  direct solve without refinement): close to fp16 rounding, not equal (the copy
  is what runs), the FP32 solve itself pinned to the oracle;
- the iterative solve with and without the copy: the same FP64 residual target,
  solutions within the tolerance, iteration counts close;
- a factor with entries beyond the half range keeps to the FP32 arrays;
- graph replays of the iterative solve equal eager launches bitwise."""
import ctypes
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from artisynth_core_amd import KKT_PREC_FP32_REFINE
from artisynth_core_amd import generators as G
from artisynth_core_amd._lib import ptr
from artisynth_core_amd.solvers import MI355XSolver, Matrix
from kkt_helpers import rel_err, rel_residual

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


def _precond(s, b):
    """x0 = M^-1 b of the iterative solve (a tolerance it meets at once)."""
    x = np.zeros(len(b))
    bb = np.ascontiguousarray(b)
    va = s._vals_for_test
    rc = s._L.kkt_iterative_solve(s._h, ptr(va, ctypes.c_double), ptr(bb, ctypes.c_double),
                                  ptr(x, ctypes.c_double), 1e30, 5)
    assert rc == 1, rc
    return x


def _beam_solver(nx, half):
    beam, vals, b = G.small_beam(nx)
    s = MI355XSolver(precision=KKT_PREC_FP32_REFINE)
    s.setIterativeHalf(half)
    s.analyzeArrays(vals, beam.colIdxs, beam.rowOffs, beam.n, Matrix.SYMMETRIC)
    s.factor(vals)
    s._vals_for_test = np.ascontiguousarray(vals)
    return s, beam, vals, b


def test_half_preconditioner_application(gpu):
    _say("half preconditioner application, beam 26")
    s, beam, vals, b = _beam_solver(26, True)
    try:
        assert s.info()["precond_half"] == 1
        xh = _precond(s, b)
        # the FP32 arrays: a direct solve whose refinement target is met at once
        s.setRefinement(1e30, 1)
        x32 = np.zeros(beam.n)
        s.solve(x32, b)
        sym = s.symbolic()
        f = O.Factor(beam.rowOffs, beam.colIdxs, vals, perm=sym["perm"], sup_ptr=sym["sup_ptr"])
        xo = f.solve(b)
        assert rel_err(x32, xo) <= 1e-3  # FP32 factor, no refinement
        d = rel_err(xh, x32)
        assert 0.0 < d <= 2e-2, d  # the copy runs, and only its rounding separates them
    finally:
        s.dispose()


@pytest.mark.parametrize("nx", [12, 26])
def test_iterative_solve_with_and_without_half(gpu, nx):
    _say("iterative half vs fp32 arrays, beam %d" % nx)
    beam, vals, b = G.small_beam(nx)
    v2 = vals * np.random.default_rng(31).uniform(0.95, 1.05, vals.shape)
    out = {}
    for half in (1, 0):
        s = MI355XSolver(precision=KKT_PREC_FP32_REFINE)
        try:
            s.setIterativeHalf(half)
            s.analyzeArrays(vals, beam.colIdxs, beam.rowOffs, beam.n, Matrix.SYMMETRIC)
            s.factor(vals)
            assert s.info()["precond_half"] == half
            x = np.zeros(beam.n)
            it = s.iterativeSolve(v2, x, b, 10)
            assert it > 0
            assert rel_residual(beam.rowOffs, beam.colIdxs, v2, x, b) <= 1e-10
            out[half] = (it, x)
        finally:
            s.dispose()
    assert abs(out[1][0] - out[0][0]) <= 2, (out[1][0], out[0][0])
    assert rel_err(out[1][1], out[0][1]) <= 1e-8


def _overflow_system(n1=300):
    # [B C; C^T D] in natural order: B tridiagonal (diagonal 10, off-diagonal -3),
    # C = 1e6 I, D = -1e6 I (indefinite, condition number 2.6).  B's pivots (9 to
    # 10) stay above the static perturbation threshold (1e-6 max|a_ij| = 1) and
    # pass the in-block pivot test, so L has entries 1e6 / 9 in D's rows and the
    # solve form's contribution rows reach 1.25e5, beyond the half range.
    B = sp.diags([np.full(n1 - 1, -3.0), np.full(n1, 10.0), np.full(n1 - 1, -3.0)], [-1, 0, 1])
    C = sp.identity(n1) * 1e6
    D = -sp.identity(n1) * 1e6
    A = sp.bmat([[B, C], [C.T, D]]).tocsr()
    U = sp.csr_matrix(sp.triu(A))
    U.sort_indices()
    return U.indptr.astype(np.int32), U.indices.astype(np.int32), U.data.astype(np.float64), A


def test_out_of_range_factor_keeps_fp32_arrays(gpu):
    _say("half copy out of range")
    ro, ci, va, A = _overflow_system()
    n = len(ro) - 1
    b = np.cos(np.arange(n) * 0.37)
    s = MI355XSolver(precision=KKT_PREC_FP32_REFINE, ordering=0)
    s.setIterativeHalf(True)
    try:
        s.analyzeArrays(va, ci, ro, n, Matrix.SYMMETRIC)
        s.factor(va)
        assert s.info()["precond_half"] == 0
        x = np.zeros(n)
        it = s.iterativeSolve(va * 1.01, x, b, 10)
        assert it > 0
        assert rel_residual(ro, ci, va * 1.01, x, b) <= 1e-10
        # and a well-scaled factor on the same handle takes the copy again
        beam, vals, bb = G.small_beam(8)
        s.analyzeArrays(vals, beam.colIdxs, beam.rowOffs, beam.n, Matrix.SYMMETRIC)
        s.factor(vals)
        assert s.info()["precond_half"] == 1
        x = np.zeros(beam.n)
        assert s.iterativeSolve(vals * 1.02, x, bb, 10) > 0
        assert rel_residual(beam.rowOffs, beam.colIdxs, vals * 1.02, x, bb) <= 1e-10
    finally:
        s.dispose()


def test_half_iterative_graph_replays_equal_eager(gpu):
    _say("half iterative graphs vs eager")
    beam, vals, b = G.small_beam(26)
    v2 = vals * np.random.default_rng(33).uniform(0.97, 1.03, vals.shape)
    res = []
    for graphs in (1, 0):
        s = MI355XSolver(precision=KKT_PREC_FP32_REFINE)
        try:
            s._L.kkt_set_graphs(s._h, graphs)
            s.analyzeArrays(vals, beam.colIdxs, beam.rowOffs, beam.n, Matrix.SYMMETRIC)
            s.factor(vals)
            xs = []
            for rep in range(3):
                x = np.zeros(beam.n)
                it = s.iterativeSolve(v2, x, b, 10)
                xs.append((it, x))
            for it, x in xs[1:]:
                assert it == xs[0][0] and np.array_equal(x, xs[0][1])
            res.append(xs[0])
        finally:
            s.dispose()
    assert res[0][0] == res[1][0]
    assert np.array_equal(res[0][1], res[1][1])
