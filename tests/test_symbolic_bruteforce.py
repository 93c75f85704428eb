"""CPU tests: the oracle's symbolic analysis (orc_symbolic: elimination tree and
column counts; orc_supernode_rows: supernode row structures) pinned by brute
force -- dense boolean Gaussian elimination of the permuted pattern -- on
patterns of at most a few hundred rows, independently of both the oracle's
column-merge algorithm and the product's row-subtree counting.

Brute force: B = the symmetric boolean pattern (diagonal set), permuted; for
k = 0..n-1 every pair of rows i, j > k with B[i,k] and B[j,k] becomes nonzero
(the fill of eliminating k).  Then struct L(:,k) = {i > k: B[i,k]},
colcount[k] = |struct| + 1 and parent[k] = min struct (or -1).  Supernodes are
eliminated as dense blocks: the rows of supernode s are the rows beyond it that
are nonzero in any of its columns, and their pairs all become nonzero."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from artisynth_core_amd import generators as G
from kkt_helpers import KKT5_COLS, KKT5_ROWS, P5_COLS, P5_ROWS, SPD5_COLS, SPD5_ROWS
from symbolic_helpers import product_symbolic

HERE = os.path.dirname(os.path.abspath(__file__))


def dense_pattern(ro, ci, perm):
    n = len(ro) - 1
    B = np.zeros((n, n), bool)
    r = np.repeat(np.arange(n), np.diff(ro))
    B[r, ci] = True
    B |= B.T
    B[np.arange(n), np.arange(n)] = True
    return B[np.ix_(perm, perm)]


def brute_etree_colcounts(ro, ci, perm):
    B = dense_pattern(ro, ci, perm)
    n = len(B)
    for k in range(n):
        below = np.nonzero(B[k + 1:, k])[0] + k + 1
        if len(below):
            B[np.ix_(below, below)] = True
    L = np.tril(B, -1)
    cc = L.sum(0).astype(np.int32) + 1
    parent = np.full(n, -1, np.int32)
    for k in range(n):
        nz = np.nonzero(L[:, k])[0]
        if len(nz):
            parent[k] = nz[0]
    return parent, cc


def brute_supernode_rows(ro, ci, perm, sup_ptr):
    B = dense_pattern(ro, ci, perm)
    rows_ptr, rows = [0], []
    for s in range(len(sup_ptr) - 1):
        f, l = sup_ptr[s], sup_ptr[s + 1]
        R = np.nonzero(B[l:, f:l].any(1))[0] + l
        if len(R):
            B[np.ix_(R, R)] = True
        rows.extend(R.tolist())
        rows_ptr.append(len(rows))
    return np.array(rows_ptr, np.int64), np.array(rows, np.int32)


def _random_pattern(n, density, seed):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=density, random_state=np.random.RandomState(seed), format="csr")
    A = sp.triu(A + A.T + sp.eye(n), format="csr")
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), rng


def _cases():
    yield "pardiso5", P5_ROWS, P5_COLS
    yield "spd5", SPD5_ROWS, SPD5_COLS
    yield "kkt5", KKT5_ROWS, KKT5_COLS
    beam, vals, b = G.small_beam(8)
    yield "beam8", beam.rowOffs, beam.colIdxs
    for q, (n, d) in enumerate([(60, 0.05), (120, 0.02), (200, 0.01), (150, 0.004)]):
        ro, ci, _ = _random_pattern(n, d, 100 + q)
        yield "random%d" % n, ro, ci


CASES = list(_cases())


@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_oracle_symbolic_matches_brute_force(case):
    name, ro, ci = case
    n = len(ro) - 1
    rng = np.random.default_rng(n)
    for perm in (np.arange(n, dtype=np.int32), rng.permutation(n).astype(np.int32),
                 product_symbolic(ro, ci)["perm"]):
        par, cc = O.symbolic(ro, ci, perm)
        bpar, bcc = brute_etree_colcounts(ro, ci, perm)
        assert np.array_equal(par, bpar), name
        assert np.array_equal(cc, bcc), name


@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_oracle_supernode_rows_match_brute_force(case):
    name, ro, ci = case
    n = len(ro) - 1
    S = product_symbolic(ro, ci)
    perm = S["perm"]
    rng = np.random.default_rng(n + 1)
    # the product's (relaxed) partition, single columns, and random contiguous cuts
    cuts = np.sort(rng.choice(np.arange(1, n), size=min(n - 1, max(1, n // 7)), replace=False))
    for sup in (S["sup_ptr"], np.arange(n + 1, dtype=np.int32),
                np.concatenate([[0], cuts, [n]]).astype(np.int32)):
        rp, rows = O.supernode_rows(ro, ci, perm, sup)
        brp, brows = brute_supernode_rows(ro, ci, perm, sup)
        assert np.array_equal(rp, brp), name
        assert np.array_equal(rows, brows), name


def test_planted_off_by_one_is_caught():
    # the pins have teeth: an oracle fed a permutation shifted by one position, or
    # counting columns without their diagonal, disagrees with the brute force
    beam, vals, b = G.small_beam(8)
    ro, ci = beam.rowOffs, beam.colIdxs
    perm = product_symbolic(ro, ci)["perm"]
    bpar, bcc = brute_etree_colcounts(ro, ci, perm)
    par, cc = O.symbolic(ro, ci, np.roll(perm, 1))
    assert not (np.array_equal(par, bpar) and np.array_equal(cc, bcc))
    par, cc = O.symbolic(ro, ci, perm)
    assert not np.array_equal(cc - 1, bcc)
    assert not np.array_equal(np.where(par >= 0, par - 1, -1), bpar)
    # supernode rows off by one column at every boundary
    sup = product_symbolic(ro, ci)["sup_ptr"]
    bad = sup.copy()
    bad[1:-1] = np.maximum(bad[1:-1] - 1, bad[:-2] + 1)
    rp, rows = O.supernode_rows(ro, ci, perm, bad)
    brp, brows = brute_supernode_rows(ro, ci, perm, sup)
    assert not (np.array_equal(rp, brp) and np.array_equal(rows, brows))
