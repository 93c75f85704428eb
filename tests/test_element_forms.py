"""Element-batched device assembly forms (CuDssSolver.add*ElementDeviceValues,
cudssBridgeKernels.cu:336-739).

CPU: the oracle's restatement of each form is pinned by closed forms: the
isotropic tangent, the generators' independent tensor-form element stiffness of
linear tets and trilinear hexes, the rigid-body null space of a frame-attached
element, and the volumetric term of a one-pressure element.  GPU: each device
form against the oracle on random descriptors, random (duplicate and negative)
slots, mixed free / frame-attached nodes and repeated calls."""
import sys

import numpy as np
import pytest

import oracle as O
from artisynth_core_amd import generators as G

TET_NATURAL = np.array([[-1.0, -1.0, -1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])


def _dense_slots(nn, rd=3):
    """Pair slots of all (i, j) node pairs into a dense (rd nn) x (rd nn) array."""
    pairs, slots = [], []
    N = rd * nn
    for i in range(nn):
        for j in range(nn):
            pairs += [i, j]
            blk = np.full(36, -1, np.int64)
            for r in range(rd):
                for c in range(rd):
                    blk[r * rd + c] = (rd * i + r) * N + rd * j + c
            slots.append(blk)
    return np.array(pairs, np.int32), np.concatenate(slots)


def _eye_T(nn):
    T = np.zeros((nn, 18))
    T[:, :9] = np.eye(3).ravel()
    return T.ravel()


def _hex_natural_grads():
    """dN_a/dxi at the 2x2x2 Gauss points (weights 1) of the trilinear hex, in the
    node order of the generators' hex grid."""
    XI = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                   [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)
    g = 1.0 / np.sqrt(3.0)
    out = np.zeros((8, 8, 3))
    for q, (a, b, c) in enumerate(XI * g):
        for n, (xa, ya, za) in enumerate(XI):
            out[q, n] = [xa * (1 + ya * b) * (1 + za * c) / 8, ya * (1 + xa * a) * (1 + za * c) / 8,
                         za * (1 + xa * a) * (1 + ya * b) / 8]
    return out


def _assemble(nn, S, C, rd=3):
    K = np.zeros((rd * nn) ** 2)
    O.scatter_add(K, S, C, 1.0)
    return K.reshape(rd * nn, rd * nn)


def _random_tet(rng):
    while True:
        X = rng.standard_normal((4, 3))
        if abs(np.linalg.det(X[1:] - X[0])) > 0.3:
            return X


def test_linear_elastic_D_is_the_isotropic_tangent():
    for E, nu in [(1.0, 0.25), (1e5, 0.33), (3e4, 0.49)]:
        lam, mu = G.lame(E, nu)
        assert np.allclose(O.linear_elastic_D(E, nu), O.isotropic_D(lam, mu), rtol=1e-14)


def test_linear_elastic_tet_matches_the_tensor_form():
    # gradients of a linear tet in closed form; K = V (lam g g^T + mu ...) from the
    # generators' independent einsum restatement
    rng = np.random.default_rng(3)
    E, nu = 2.3e3, 0.31
    lam, mu = G.lame(E, nu)
    X = _random_tet(rng)
    Ke, vol = G.tet_linear_stiffness(X, np.array([[0, 1, 2, 3]]), lam, mu)
    g = np.linalg.solve(np.c_[np.ones(4), X], np.eye(4))[1:].T  # (4, 3) grad N_a
    pn, sl36 = _dense_slots(4)
    sl9 = sl36.reshape(-1, 36)[:, :9].ravel()
    S, C = O.linear_elastic_stiffness3_element(
        [4], [0, 16], [0, 1], [0, 4], pn, sl9, [E, nu], g.ravel(), [vol[0]], 1)
    assert np.allclose(_assemble(4, S, C), Ke[0], rtol=1e-12, atol=1e-12 * np.abs(Ke).max())
    # the geometry form from natural gradients and the one-point weight 1/6
    S, C = O.linear_elastic_stiffness3_element_geometry(
        [4], [0, 4], [0, 16], [0, 1], [0, 4], pn, sl36, [3] * 4, _eye_T(4), [E, nu], X.ravel(),
        TET_NATURAL.ravel(), [1.0 / 6.0], 1)
    # det J is signed: a left-handed node order gives a negative volume
    sgn = np.sign(np.linalg.det(X[1:] - X[0]))
    assert np.allclose(sgn * _assemble(4, S, C), Ke[0], rtol=1e-12, atol=1e-12 * np.abs(Ke).max())


def test_geometry_form_hex_matches_the_beam_element():
    # trilinear hex, 2x2x2 Gauss: the generators' linear element tangent (4th-order
    # tensor form with dN/dX of the axis-aligned hex)
    E, nu = 1e5, 0.33
    beam = G.HexBeam(1, 1, 1, 0.7, 0.3, height=0.4, material=("linear", E, nu))
    Ke, _ = beam.element_tangent(np.zeros((beam.nodes.shape[0], 3)))
    X = beam.nodes[beam.elems[0]]
    NG = _hex_natural_grads()
    pn, sl36 = _dense_slots(8)
    S, C = O.linear_elastic_stiffness3_element_geometry(
        [8], [0, 8], [0, 64], [0, 8], [0, 64], pn, sl36, [3] * 8, _eye_T(8), [E, nu], X.ravel(),
        NG.ravel(), np.ones(8), 1)
    assert np.allclose(_assemble(8, S, C), Ke[0], rtol=1e-11, atol=1e-11 * np.abs(Ke).max())


def test_frame_attached_element_reduces_to_zero_stiffness():
    # every node attached to one frame, T_q = [I; [r_q]x]: T K T^T summed over the
    # pairs is P^T K P with P the rigid-motion map, which a small-strain element
    # stiffness annihilates; with T = I the form is the plain one
    rng = np.random.default_rng(5)
    X = rng.standard_normal((8, 3)) * 0.1 + np.array(
        [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    org = np.array([0.3, -0.2, 0.5])
    T = np.zeros((8, 18))
    for q in range(8):
        r = X[q] - org
        rx = np.array([[0, -r[2], r[1]], [r[2], 0, -r[0]], [-r[1], r[0], 0]])
        T[q] = np.vstack([np.eye(3), rx]).ravel()
    pn, _ = _dense_slots(8)
    sl = np.tile(np.arange(36), 64)  # every pair onto the one 6x6 frame block
    NG = _hex_natural_grads()
    S, C = O.linear_elastic_stiffness3_element_geometry(
        [8], [0, 8], [0, 64], [0, 8], [0, 64], pn, sl, [6] * 8, T.ravel(), [1e3, 0.3], X.ravel(),
        NG.ravel(), np.ones(8), 1)
    Kf = np.zeros(36)
    O.scatter_add(Kf, S, C, 1.0)
    _, sl3 = _dense_slots(8)
    S3, C3 = O.linear_elastic_stiffness3_element_geometry(
        [8], [0, 8], [0, 64], [0, 8], [0, 64], pn, sl3, [3] * 8, _eye_T(8), [1e3, 0.3], X.ravel(),
        NG.ravel(), np.ones(8), 1)
    K3 = _assemble(8, S3, C3)
    assert np.abs(Kf).max() < 1e-10 * np.abs(K3).max()
    # the 6 rigid modes are the element's null space
    w = np.linalg.eigvalsh(K3)
    assert (np.abs(w) < 1e-9 * w.max()).sum() == 6


def test_dilational_form_is_the_volumetric_term():
    # one pressure: c_i = V g_i, Rinv = kappa / V gives kappa V g_i g_j^T, the lambda
    # term of the tensor form with mu = 0
    rng = np.random.default_rng(7)
    X = _random_tet(rng)
    kappa = 4.5e3
    Ke, vol = G.tet_linear_stiffness(X, np.array([[0, 1, 2, 3]]), kappa, 0.0)
    g = np.linalg.solve(np.c_[np.ones(4), X], np.eye(4))[1:].T
    pn, sl36 = _dense_slots(4)
    cons = (vol[0] * g).ravel()
    S, C = O.dilational_stiffness3_element([4], [0, 4], [1], [0, 16], [0, 4], [0, 1], pn, sl36,
                                           [3] * 4, _eye_T(4), cons, [kappa / vol[0]], 1)
    assert np.allclose(_assemble(4, S, C), Ke[0], rtol=1e-12, atol=1e-12 * np.abs(Ke).max())
    # two pressures: K = Cmat Rinv Cmat^T with Cmat (3 nn x np) stacked node by node
    npr = 2
    cv = rng.standard_normal((4, npr, 3))
    R = rng.standard_normal((npr, npr))
    S, C = O.dilational_stiffness3_element([4], [0, 4], [npr], [0, 16], [0, 4 * npr], [0, npr * npr],
                                           pn, sl36, [3] * 4, _eye_T(4), cv.ravel(), R.ravel(), 1)
    Cmat = cv.transpose(0, 2, 1).reshape(12, npr)
    assert np.allclose(_assemble(4, S, C), Cmat @ R @ Cmat.T, rtol=1e-12, atol=1e-12)
    # no pressure: nothing
    S, C = O.dilational_stiffness3_element([4], [0, 4], [0], [0, 16], [0, 0], [0, 0], pn, sl36,
                                           [3] * 4, _eye_T(4), np.zeros(0), np.zeros(0), 1)
    assert S.size == 0


# ---------------------------------------------------------------- GPU

def _say(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


def _random_elements(rng, nelems, nnz, nips_max=4):
    """Random compact descriptors: node counts 4 / 8 / 10, a random subset of the
    node pairs, 1..nips_max integration points, nodes free (dim 3, T = I) or
    frame-attached (dim 6, random T), slots in [-1, nnz) with duplicates."""
    nc = rng.choice([4, 8, 10], nelems).astype(np.int32)
    no = np.r_[0, np.cumsum(nc)].astype(np.int32)
    pairs = []
    po = [0]
    for e in range(nelems):
        allp = [(i, j) for i in range(nc[e]) for j in range(nc[e])]
        pick = rng.choice(len(allp), size=int(rng.integers(1, len(allp) + 1)), replace=False)
        for q in sorted(pick):
            pairs += list(allp[q])
        po.append(po[-1] + len(pick))
    po = np.array(po, np.int32)
    pn = np.array(pairs, np.int32)
    ni = rng.integers(1, nips_max + 1, nelems)
    io = np.r_[0, np.cumsum(ni)].astype(np.int32)
    go = np.r_[0, np.cumsum(ni * nc)].astype(np.int32)
    nd = np.where(rng.random(no[-1]) < 0.3, 6, 3).astype(np.int32)
    T = np.zeros((no[-1], 18))
    for q in range(no[-1]):
        T[q, :9] = np.eye(3).ravel() if nd[q] == 3 else rng.standard_normal(9)
        if nd[q] == 6:
            T[q, 9:] = rng.standard_normal(9)
    sl = rng.integers(-1, nnz, 36 * po[-1]).astype(np.int32)
    return dict(nc=nc, no=no, po=po, pn=pn, io=io, go=go, nd=nd, T=T.ravel(), sl=sl)


def _solver_with_pattern():
    from artisynth_core_amd.solvers import Matrix, MI355XSolver
    beam, vals, b = G.small_beam(8)
    s = MI355XSolver()
    s.analyzeArrays(vals, beam.colIdxs, beam.rowOffs, beam.n, Matrix.SYMMETRIC)
    return s, beam.nnz


def _check(s, ref):
    got = s.getDeviceValues(np.zeros(ref.size))
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12 * max(np.abs(ref).max(), 1.0))


@pytest.mark.gpu
def test_device_element_forms_match_the_oracle(gpu):
    _say("element-batched device forms")
    rng = np.random.default_rng(31)
    s, nnz = _solver_with_pattern()
    try:
        ne = 120
        d = _random_elements(rng, ne, nnz)
        ngr = int(d["go"][-1])
        nip = int(d["io"][-1])
        grads = rng.standard_normal(3 * ngr)
        Ds = rng.standard_normal(36 * nip)
        sig = rng.standard_normal(6 * nip)
        dvs = rng.uniform(0.1, 1.0, nip)
        s.clearDeviceValues()
        ref = np.zeros(nnz)
        for rep in range(2):  # the second call reuses the cached plan
            s.addMaterialStiffness3ElementDeviceValues(d["nc"], d["no"], d["po"], d["io"], d["go"], d["pn"],
                                                       d["sl"], d["nd"], d["T"], grads, Ds, sig, dvs, ne, 0.7)
            S, C = O.material_stiffness3_element(d["nc"], d["no"], d["po"], d["io"], d["go"], d["pn"], d["sl"],
                                                 d["nd"], d["T"], grads, Ds, sig, dvs, ne)
            O.scatter_add(ref, S, C, 0.7)
            _check(s, ref)
        # linear elastic, 9 slots per pair, no reduction
        sl9 = rng.integers(-1, nnz, 9 * int(d["po"][-1])).astype(np.int32)
        params = np.c_[rng.uniform(1e2, 1e4, ne), rng.uniform(0.1, 0.45, ne)].ravel()
        s.addLinearElasticStiffness3ElementDeviceValues(d["nc"], d["po"], d["io"], d["go"], d["pn"], sl9,
                                                        params, grads, dvs, ne, -1.3)
        S, C = O.linear_elastic_stiffness3_element(d["nc"], d["po"], d["io"], d["go"], d["pn"], sl9, params,
                                                   grads, dvs, ne)
        O.scatter_add(ref, S, C, -1.3)
        _check(s, ref)
        # geometry form: perturbed unit-cube / tet node positions keep det J > 0
        pos = np.zeros((int(d["no"][-1]), 3))
        ngrad = np.zeros((ngr, 3))
        w = rng.uniform(0.1, 1.0, nip)
        HEX = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]],
                       float)
        NGH = _hex_natural_grads()
        for e in range(ne):
            n0, nn = d["no"][e], d["nc"][e]
            base = HEX if nn == 8 else np.r_[[[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]],
                                             rng.random((nn - 4, 3)) * 0.2 + 0.2]
            pos[n0:n0 + nn] = base + rng.standard_normal((nn, 3)) * 0.03
            for ip in range(d["io"][e + 1] - d["io"][e]):
                g0 = d["go"][e] + ip * nn
                if nn == 8:
                    ngrad[g0:g0 + 8] = NGH[ip % 8]
                else:
                    ngrad[g0:g0 + 4] = TET_NATURAL
                    ngrad[g0 + 4:g0 + nn] = rng.standard_normal((nn - 4, 3)) * 0.01
        s.addLinearElasticStiffness3ElementGeometryDeviceValues(
            d["nc"], d["no"], d["po"], d["io"], d["go"], d["pn"], d["sl"], d["nd"], d["T"], params,
            pos.ravel(), ngrad.ravel(), w, ne, 0.4)
        S, C = O.linear_elastic_stiffness3_element_geometry(
            d["nc"], d["no"], d["po"], d["io"], d["go"], d["pn"], d["sl"], d["nd"], d["T"], params,
            pos.ravel(), ngrad.ravel(), w, ne)
        O.scatter_add(ref, S, C, 0.4)
        _check(s, ref)
        # dilational: 0..3 pressures per element (0 adds nothing)
        pc = rng.integers(0, 4, ne).astype(np.int32)
        co = np.r_[0, np.cumsum(pc * d["nc"])].astype(np.int32)
        ro = np.r_[0, np.cumsum(pc * pc)].astype(np.int32)
        cons = rng.standard_normal(3 * int(co[-1]))
        rinv = rng.standard_normal(int(ro[-1]))
        s.addDilationalStiffness3ElementDeviceValues(d["nc"], d["no"], pc, d["po"], co, ro, d["pn"], d["sl"],
                                                     d["nd"], d["T"], cons, rinv, ne, 2.0)
        S, C = O.dilational_stiffness3_element(d["nc"], d["no"], pc, d["po"], co, ro, d["pn"], d["sl"], d["nd"],
                                               d["T"], cons, rinv, ne)
        O.scatter_add(ref, S, C, 2.0)
        _check(s, ref)
    finally:
        s.dispose()


@pytest.mark.gpu
def test_device_element_forms_reject_bad_descriptors(gpu):
    # argument checks of the reference (CuDssSolver.java:600-620) and the native
    # consistency checks: nothing is added when a call is rejected
    from artisynth_core_amd.solvers import NumericalException
    _say("element-batched device forms: bad descriptors")
    rng = np.random.default_rng(32)
    s, nnz = _solver_with_pattern()
    try:
        d = _random_elements(rng, 10, nnz)
        nip, ngr = int(d["io"][-1]), int(d["go"][-1])
        args = [d["nc"], d["no"], d["po"], d["io"], d["go"], d["pn"], d["sl"], d["nd"], d["T"],
                np.zeros(3 * ngr), np.zeros(36 * nip), np.zeros(6 * nip), np.ones(nip)]
        s.clearDeviceValues()
        with pytest.raises(ValueError):
            s.addMaterialStiffness3ElementDeviceValues(*args, 11, 1.0)  # nelems past the arrays
        bad = list(args)
        bad[5] = d["pn"].copy()
        bad[5][0] = 50  # node index outside its element
        with pytest.raises(NumericalException):
            s.addMaterialStiffness3ElementDeviceValues(*bad, 10, 1.0)
        bad = list(args)
        bad[7] = d["nd"].copy()
        bad[7][0] = 7  # target dimension beyond 6
        with pytest.raises(NumericalException):
            s.addMaterialStiffness3ElementDeviceValues(*bad, 10, 1.0)
        assert not s.getDeviceValues(np.ones(nnz)).any()
    finally:
        s.dispose()
