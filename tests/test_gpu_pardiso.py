"""GPU parity against MKL PARDISO run with the reference's settings
(tests/golden/pardiso_fixtures.npz): the HIP path through the C ABI, both
factor precisions, reaches PARDISO's solution to 1e-8 (relative), the same
inertia, no perturbed pivots, and the stale-factor iterative solve reaches
PARDISO's CGS result.  nnz(L) of the two orderings is printed side by side."""
import sys

import numpy as np
import pytest

from artisynth_core_amd import KKT_PREC_FP32_REFINE, KKT_PREC_FP64
from artisynth_core_amd.solvers import MI355XSolver, Matrix
from kkt_helpers import rel_residual
from pardiso_cases import ALL, case, info, sampled

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ALL)
@pytest.mark.parametrize("prec", [KKT_PREC_FP64, KKT_PREC_FP32_REFINE])
def test_gpu_matches_pardiso(gpu, name, prec):
    print("[gpu-case] pardiso fixture %s prec=%d" % (name, prec), flush=True)
    ro, ci, va, b, hyb, fx = case(name)
    n = len(ro) - 1
    s = MI355XSolver(precision=prec)
    try:
        s.analyzeArrays(va, ci, ro, n, Matrix.SYMMETRIC)
        s.factor(va)
        x = np.zeros(n)
        s.solve(x, b)
        pi = info(fx)
        err = sampled(fx, x)
        inf = s.info()
        sys.stdout.write("[pardiso] %s prec=%d err=%.2e nnzL stored %d exact %d PARDISO %d\n" % (
            name, prec, err, inf["nnzL"], inf["nnzL_exact"], pi["nnzL"]))
        assert rel_residual(ro, ci, va, x, b) <= 1e-8
        assert err <= 1e-8, (name, err)
        assert (inf["npos"], inf["nneg"]) == (pi["npos"], pi["nneg"])
        if prec == KKT_PREC_FP64:
            assert inf["nperturbed"] == pi["nperturbed"] == 0
        if hyb is not None:
            va2, b2 = hyb
            x2 = np.zeros(n)
            it = s.iterativeSolve(np.ascontiguousarray(va2), x2, np.ascontiguousarray(b2), 10)
            assert it > 0 and fx["cgs_iters"] > 0
            assert rel_residual(ro, ci, va2, x2, b2) <= 1e-9
            assert sampled(fx, x2, "x_hybrid") <= 1e-8
    finally:
        s.dispose()
