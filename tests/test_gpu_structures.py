"""GPU tests on structures the mesh problems do not produce, each against the
oracle factored on the same ordering and supernodes, in both precisions:

- a star (one vertex coupled to every other): its centre's supernode has
  hundreds of children, so the forward's update-slot gather runs many passes;
- a random KKT [S G^T; G 0] with an irregular S and zero-diagonal constraint rows;
- many tiny independent blocks (1x1, 2x2 indefinite, 3x3), one assembly tree each;
- the 1x1 system."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from artisynth_core_amd import KKT_PREC_FP32_REFINE, KKT_PREC_FP64
from artisynth_core_amd.solvers import MI355XSolver, Matrix, kkt_upper_csr
from kkt_helpers import rel_err, rel_residual

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


def _upper(A):
    U = sp.csr_matrix(sp.triu(sp.csr_matrix(A)))
    U.sort_indices()
    return U.indptr.astype(np.int32), U.indices.astype(np.int32), U.data.astype(np.float64)


def _star(n, rng):
    # SPD: centre row 0 coupled to all others, diagonally dominant
    A = sp.lil_matrix((n, n))
    for i in range(1, n):
        A[0, i] = A[i, 0] = rng.uniform(-1, 1)
        A[i, i] = 2.0 + rng.uniform(0, 1)
    A[0, 0] = float(n)
    return _upper(A.tocsr())


def _random_kkt(nm, ng, rng):
    R = sp.random(nm, nm, density=6.0 / nm, random_state=rng.integers(1 << 30))
    S = (R @ R.T + sp.identity(nm) * 1.0).tocsr()
    GT = sp.lil_matrix((nm, ng))
    cols = rng.permutation(nm)[: 2 * ng]
    for j in range(ng):  # each constraint couples two distinct DOFs: full row rank
        GT[cols[2 * j], j] = 1.0
        GT[cols[2 * j + 1], j] = -rng.uniform(0.5, 1.5)
    return kkt_upper_csr(S, GT.tocsr(), None)


def _tiny_blocks(nb, rng):
    blocks = []
    for q in range(nb):
        kind = q % 3
        if kind == 0:
            blocks.append(sp.csr_matrix(np.array([[rng.uniform(1, 2) * (1 if q % 2 else -1)]])))
        elif kind == 1:  # indefinite 2x2 with a zero diagonal
            blocks.append(sp.csr_matrix(np.array([[0.0, 1.0 + q * 0.01], [1.0 + q * 0.01, 0.0]])))
        else:
            M = rng.standard_normal((3, 3))
            blocks.append(sp.csr_matrix(M @ M.T + 3 * np.eye(3)))
    return _upper(sp.block_diag(blocks).tocsr())


def _check(name, ro, ci, va, prec):
    n = len(ro) - 1
    b = np.cos(np.arange(n) * 0.7) + 0.1
    s = MI355XSolver(precision=prec)
    try:
        s.analyzeArrays(va, ci, ro, n, Matrix.SYMMETRIC)
        s.factor(va)
        x = np.zeros(n)
        s.solve(x, b)
        sym = s.symbolic()
        f = O.Factor(ro, ci, va, perm=sym["perm"], sup_ptr=sym["sup_ptr"])
        xo = f.solve(b)
        assert rel_residual(ro, ci, va, x, b) <= 1e-8, name
        assert rel_err(x, xo) <= 1e-8, name
        inf = s.info()
        A = sp.csr_matrix((va, ci, ro), shape=(n, n))
        A = (A + sp.triu(A, 1).T).toarray()
        ev = np.linalg.eigvalsh(A)
        assert (inf["npos"], inf["nneg"]) == (int((ev > 0).sum()), int((ev < 0).sum())), name
        return sym
    finally:
        s.dispose()


@pytest.mark.parametrize("prec", [KKT_PREC_FP64, KKT_PREC_FP32_REFINE])
def test_star_many_children(gpu, prec):
    _say("star n=600 prec=%d" % prec)
    rng = np.random.default_rng(11)
    ro, ci, va = _star(600, rng)
    sym = _check("star", ro, ci, va, prec)
    sp_ = sym["sparent"]
    nch = np.bincount(sp_[sp_ >= 0], minlength=len(sp_))
    assert nch.max() > 8  # the slot gather takes several passes of eight children


@pytest.mark.parametrize("prec", [KKT_PREC_FP64, KKT_PREC_FP32_REFINE])
def test_random_kkt(gpu, prec):
    _say("random kkt prec=%d" % prec)
    rng = np.random.default_rng(12)
    ro, ci, va = _random_kkt(700, 120, rng)
    _check("random kkt", ro, ci, va, prec)


@pytest.mark.parametrize("prec", [KKT_PREC_FP64, KKT_PREC_FP32_REFINE])
def test_many_tiny_blocks(gpu, prec):
    _say("tiny blocks prec=%d" % prec)
    rng = np.random.default_rng(13)
    ro, ci, va = _tiny_blocks(150, rng)
    _check("tiny blocks", ro, ci, va, prec)


@pytest.mark.parametrize("prec", [KKT_PREC_FP64, KKT_PREC_FP32_REFINE])
def test_one_by_one(gpu, prec):
    _say("1x1 prec=%d" % prec)
    ro = np.array([0, 1], np.int32)
    ci = np.array([0], np.int32)
    va = np.array([-4.0])
    _check("1x1", ro, ci, va, prec)
