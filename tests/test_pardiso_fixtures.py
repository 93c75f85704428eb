"""CPU tests: the oracle against MKL PARDISO run with the reference's settings
(tests/golden/pardiso_fixtures.npz, written by tests/golden/make_pardiso.py).

PARDISO orders with METIS and pivots in its own supernodes, the oracle uses the
product's nested dissection and panel-restricted Bunch-Kaufman, so the factors
differ; what must agree is what the matrix fixes: the solution (to the
conditioning of each case), the inertia (Sylvester's law) and the absence of
perturbed pivots on these well-posed systems.  nnz(L) is compared as a fill
quality figure of the ordering, not for equality."""
import numpy as np
import pytest

import oracle as O
from pardiso_cases import ALL, case, info, sampled
from kkt_helpers import rel_residual
from symbolic_helpers import product_symbolic


def _oracle(ro, ci, va):
    S = product_symbolic(ro, ci, va)
    return O.Factor(ro, ci, va, perm=S["perm"], sup_ptr=S["sup_ptr"]), S


@pytest.mark.parametrize("name", ALL)
def test_oracle_matches_pardiso(name):
    ro, ci, va, b, hyb, fx = case(name)
    f, S = _oracle(ro, ci, va)
    x = f.solve(b)
    pi = info(fx)
    # the solution: PARDISO's x (one refinement step) and the oracle's agree
    assert sampled(fx, x) <= 1e-10, (name, sampled(fx, x))
    assert abs(np.linalg.norm(x) - fx["x_norm"]) <= 1e-10 * fx["x_norm"]
    oi = f.info()
    assert (oi["npos"], oi["nneg"]) == (pi["npos"], pi["nneg"])
    assert oi["nperturbed"] == pi["nperturbed"] == 0
    # fill of the product's ordering against METIS inside PARDISO (exact counts;
    # PARDISO's iparm[17] of a tiny matrix counts its dense supernode storage)
    if len(ro) > 200:
        ratio = S["info"]["nnzL_exact"] / pi["nnzL"]
        assert 0.5 < ratio < 1.5, (name, ratio)
    if hyb is not None:
        # stale-factor iterative solve: PARDISO's CGS and the oracle's BiCGStab
        # (the bridge's algorithm) reach the same solution of the new system
        va2, b2 = hyb
        it, x2 = f.bicgstab(b2, 1e-10, 32, vals=va2)
        assert it > 0 and fx["cgs_iters"] > 0
        assert rel_residual(ro, ci, va2, x2, b2) <= 1e-9
        assert sampled(fx, x2, "x_hybrid") <= 1e-8, (name, sampled(fx, x2, "x_hybrid"))


def test_planted_wrong_sign_is_caught_by_the_pardiso_fixture():
    # a sign error in the constraint block gives a solution far from PARDISO's
    ro, ci, va, b, hyb, fx = case("beam8")
    from artisynth_core_amd import generators as G
    beam, _, _ = G.small_beam(8)
    bad = va.copy()
    bad[beam._gpos] *= -1.0
    f, _ = _oracle(ro, ci, bad)
    assert sampled(fx, f.solve(b)) > 1e-3
