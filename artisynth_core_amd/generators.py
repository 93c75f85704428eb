"""Synthetic KKT systems of the benchmark configurations (host numpy/scipy).

These stand in for ArtiSynth's model assembly (the caller of the solver):
trilinear hex elements with 2x2x2 Gauss integration, lumped mass, Rayleigh
stiffness damping, point-attachment constraints G^T = I3 blocks, and the
ConstrainedBackwardEuler system S = M + h D + h^2 K (MechSystemSolver.java:
2686-2692, 2900-3019) assembled into the KKT upper CSR of KKTSolver
(KKTSolver.java:171-454), right-hand side b = M v + h f (:4851-4862).

Materials restated from ArtiSynth:
  LinearMaterial(E, nu)                         (LinearMaterial.java:65-69)
  NeoHookeanMaterial: W = mu(I1-3)/2 - mu lnJ + lam lnJ^2/2
                                                (NeoHookeanMaterial.java:113-156)
"""
import numpy as np
import scipy.sparse as sp

from .solvers import kkt_upper_csr

GRAVITY = np.array([0.0, 0.0, -9.8])  # MechModel DEFAULT_GRAVITY


def hex_grid(nx, ny, nz, lx, ly, lz):
    """Nodes (N,3) and hex connectivity (E,8) of a box grid centred at the origin
    (FemFactory.createHexGrid layout: x fastest)."""
    xs = np.linspace(-lx / 2, lx / 2, nx + 1)
    ys = np.linspace(-ly / 2, ly / 2, ny + 1)
    zs = np.linspace(-lz / 2, lz / 2, nz + 1)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    nodes = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)

    def nid(i, j, k):
        return (k * (ny + 1) + j) * (nx + 1) + i

    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    elems = np.stack([nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                      nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1),
                      nid(i, j + 1, k + 1)], 1)
    return nodes, elems


_HEX_XI = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                    [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)


def hex_gauss_grads(hx, hy, hz):
    """dN_a/dX at the 8 Gauss points of an axis-aligned hx*hy*hz hex: (8,8,3),
    and the quadrature weight * det J per point."""
    g = 1.0 / np.sqrt(3.0)
    gps = _HEX_XI * g
    dN = np.zeros((8, 8, 3))
    for q, (a, b, c) in enumerate(gps):
        for n, (xa, ya, za) in enumerate(_HEX_XI):
            dN[q, n, 0] = xa * (1 + ya * b) * (1 + za * c) / 8 * (2 / hx)
            dN[q, n, 1] = ya * (1 + xa * a) * (1 + za * c) / 8 * (2 / hy)
            dN[q, n, 2] = za * (1 + xa * a) * (1 + ya * b) / 8 * (2 / hz)
    return dN, hx * hy * hz / 8.0


def lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def neo_hookean_PA(F, lam, mu):
    """First Piola stress P and tangent A = dP/dF (A[...,i,J,k,L]) of
    W = mu(I1-3)/2 - mu lnJ + lam lnJ^2/2, for F of shape (...,3,3)."""
    J = np.linalg.det(F)
    Fi = np.linalg.inv(F)  # Fi[J,i] = F^-1
    FiT = np.swapaxes(Fi, -1, -2)
    lnJ = np.log(J)
    P = mu * (F - FiT) + (lam * lnJ)[..., None, None] * FiT
    I3 = np.eye(3)
    A = mu * np.einsum("ik,JL->iJkL", I3, I3)[None] * np.ones(F.shape[:-2] + (1, 1, 1, 1))
    c1 = (mu - lam * lnJ)[..., None, None, None, None]
    A = A + c1 * np.einsum("...Li,...Jk->...iJkL", Fi, Fi)
    A = A + lam * np.einsum("...Ji,...Lk->...iJkL", Fi, Fi)
    return P, A


class HexBeam:
    """Cantilever hex beam with its left face attached by point constraints."""

    def __init__(self, nx, ny, nz, length=1.0, width=0.2, height=None, density=1000.0,
                 material=("linear", 1e5, 0.33), stiffness_damping=0.002, h=0.01):
        height = width if height is None else height
        self.nodes, self.elems = hex_grid(nx, ny, nz, length, width, height)
        self.nn = len(self.nodes)
        self.sizeM = 3 * self.nn
        self.h = h
        self.beta = stiffness_damping
        self.material = material
        hx, hy, hz = length / nx, width / ny, height / nz
        self.dN, self.wdet = hex_gauss_grads(hx, hy, hz)
        # B_g[(i,J), (a,k)] = delta_ik dN_a/dX_J at Gauss point g
        self._Bmat = np.zeros((8, 9, 24))
        for g in range(8):
            for i in range(3):
                for J in range(3):
                    self._Bmat[g, 3 * i + J, i::3] = self.dN[g, :, J]
        # lumped mass
        m_node = np.zeros(self.nn)
        np.add.at(m_node, self.elems.ravel(), density * hx * hy * hz / 8.0)
        self.mass = np.repeat(m_node, 3)
        # attachments: nodes on the x = -length/2 face
        xmin = self.nodes[:, 0].min()
        self.attached = np.nonzero(np.isclose(self.nodes[:, 0], xmin))[0]
        self.tip = np.nonzero(np.isclose(self.nodes[:, 0], self.nodes[:, 0].max()))[0]
        nG = 3 * len(self.attached)
        rows = (3 * self.attached[:, None] + np.arange(3)[None]).ravel()
        self.GT = sp.csr_matrix((np.ones(nG), (rows, np.arange(nG))), shape=(self.sizeM, nG))
        self.numG = nG
        self._build_pattern()

    def _build_pattern(self):
        E, npe = self.elems.shape
        nd = 3 * npe
        dofs = (3 * self.elems[:, :, None] + np.arange(3)[None, None]).reshape(E, nd)
        r = np.repeat(dofs, nd, axis=1).ravel()
        c = np.tile(dofs, (1, nd)).ravel()
        n = self.sizeM
        keep = r <= c
        self._ekeep = keep
        key = r[keep].astype(np.int64) * n + c[keep]
        ukeys, inv = np.unique(key, return_inverse=True)
        self._einv = inv
        ur, uc = ukeys // n, ukeys % n
        # make sure the diagonal is present
        S = sp.csr_matrix((np.ones(len(ur)), (ur, uc)), shape=(n, n))
        self.Spat = S
        self._ukeys = ukeys
        ro, ci, _ = kkt_upper_csr(S, self.GT, None)
        self.rowOffs, self.colIdxs = ro, ci
        # map from S-upper entry (ukeys order) to KKT value position
        Sr = np.repeat(np.arange(n), np.diff(S.indptr))
        skey = Sr.astype(np.int64) * n + S.indices
        assert np.array_equal(skey, ukeys)
        pos = np.empty(len(ukeys), np.int64)
        kr = np.repeat(np.arange(len(ro) - 1), np.diff(ro))
        kkey = kr.astype(np.int64) * (n + self.numG) + ci
        tgt = Sr.astype(np.int64) * (n + self.numG) + S.indices
        pos = np.searchsorted(kkey, tgt)
        assert np.array_equal(kkey[pos], tgt)
        self._kpos = pos
        # positions of G^T entries and of the lambda diagonal
        isG = (kr < n) & (ci >= n)
        self._gpos = np.nonzero(isG)[0]
        self._lampos = np.nonzero(kr >= n)[0]
        self.nnz = len(ci)
        self.n = n + self.numG
        self._diagpos = None

    def element_tangent(self, u):
        """Per-element 24x24 tangent stiffness and 24 internal forces at displacement u."""
        E = len(self.elems)
        mat = self.material
        if mat[0] == "linear":
            lam, mu = lame(mat[1], mat[2])
            F = np.broadcast_to(np.eye(3), (1, 8, 3, 3))
            I3 = np.eye(3)
            A = (mu * (np.einsum("ik,JL->iJkL", I3, I3) + np.einsum("iL,Jk->iJkL", I3, I3)) +
                 lam * np.einsum("iJ,kL->iJkL", I3, I3))
            Ke = np.einsum("gaJ,iJkL,gbL->aibk", self.dN, A, self.dN) * self.wdet
            Ke = np.broadcast_to(Ke.reshape(1, 24, 24), (E, 24, 24))
            ue = u[self.elems]  # (E,8,3)
            fint = np.einsum("eaibk,ebk->eai", Ke.reshape(E, 8, 3, 8, 3), ue).reshape(E, 24)
            return Ke, fint
        if mat[0] == "neohookean":
            mu, kappa = mat[1], mat[2]
            lam = kappa - 2.0 * mu / 3.0
            ue = u[self.elems]  # (E,8,3)
            F = np.eye(3)[None, None] + np.einsum("eai,gaJ->egiJ", ue, self.dN)
            P, A = neo_hookean_PA(F, lam, mu)
            A9 = np.ascontiguousarray(A.reshape(E, 8, 9, 9))
            Ke = np.zeros((E, 24, 24))
            for g in range(8):
                Bg = self._Bmat[g]  # (9, 24): grad u (i,J) from nodal u (a,k)
                C = (A9[:, g].reshape(E * 9, 9) @ Bg).reshape(E, 9, 24)
                Ke += (Bg.T @ C.transpose(1, 0, 2).reshape(9, E * 24)).reshape(24, E, 24).transpose(1, 0, 2)
            Ke *= self.wdet
            fint = np.einsum("egiJ,gaJ->eai", P, self.dN).reshape(E, 24) * self.wdet
            return Ke, fint
        raise ValueError("unknown material %r" % (mat,))

    def system(self, u, v, fext):
        """KKT values (upper CSR) and right-hand side for one ConstrainedBackwardEuler
        step at displacement u, velocity v (both (nn,3)), external nodal forces fext."""
        E = len(self.elems)
        Ke, fint_e = self.element_tangent(u)
        n = self.sizeM
        # K upper entries
        kv = np.bincount(self._einv, weights=np.ascontiguousarray(Ke).reshape(-1)[self._ekeep.ravel()],
                         minlength=len(self._ukeys))
        fint = np.zeros(n)
        nd = 3 * self.elems.shape[1]
        dofs = (3 * self.elems[:, :, None] + np.arange(3)[None, None]).reshape(E, nd)
        np.add.at(fint, dofs.ravel(), fint_e.ravel())
        h, beta = self.h, self.beta
        a = h * beta + h * h
        sv = a * kv
        diag = (self._ukeys // n) == (self._ukeys % n)
        sv[diag] += self.mass
        vals = np.zeros(self.nnz)
        vals[self._kpos] = sv
        vals[self._gpos] = 1.0
        vals[self._lampos] = 0.0
        # f = fext - fint - D v, D = beta K; b = M v + h f (constrainedBackwardEuler,
        # MechSystemSolver.java:4851-4862) plus the right-hand-side Jacobian term
        # (a2 df/dv + a3 df/dx) v with a2 = -h, a3 = 0, i.e. + h D v
        # (KKTFactorAndSolve, MechSystemSolver.java:2686-2692, 2900-2950)
        vv = v.reshape(-1)
        Kv = self._kmul(kv, vv)
        f = fext.reshape(-1) - fint - beta * Kv
        bm = self.mass * vv + h * f + h * beta * Kv
        bg = np.zeros(self.numG)
        return vals, np.concatenate([bm, bg])

    def _kmul(self, kv, x):
        n = self.sizeM
        r = self._ukeys // n
        c = self._ukeys % n
        y = np.bincount(r, weights=kv * x[c], minlength=n)
        off = r != c
        y += np.bincount(c[off], weights=kv[off] * x[r[off]], minlength=n)
        return y

    def gravity_loads(self):
        m_node = self.mass[0::3]
        return m_node[:, None] * GRAVITY[None]


def small_beam(nx=8, seed=0x1234):
    """Config S: hex beam nx x 2 x 2, E=1e5, nu=0.33, rho=1000, left face attached,
    one step h=0.01 from rest; loads = gravity + seeded uniform[-1,1] N per node."""
    beam = HexBeam(nx, 2, 2, 1.0, 0.2, density=1000.0, material=("linear", 1e5, 0.33), h=0.01)
    rng = np.random.default_rng(seed)
    fext = beam.gravity_loads() + rng.uniform(-1, 1, size=(beam.nn, 3))
    u = np.zeros((beam.nn, 3))
    vals, b = beam.system(u, u, fext)
    return beam, vals, b


def large_beam_sequence(nsteps=20, seed=0x5EED, nx=100, ny=16, nz=16, length=1.0, width=0.16):
    """Config L: hex beam nx x ny x nz, neo-Hookean mu=3.4e4, kappa=1.7e6, rho=1000,
    h=0.005; the 289 left-face nodes attached (867 constraint rows).  The 20 step
    states follow a prescribed, growing cantilever bending under seeded random tip
    loads (quasi-static shape u_z ~ x^2 (3L - x)); each step's system is assembled at
    its own state, so values change from step to step while the pattern is fixed.
    Returns (beam, [vals_k], [b_k])."""
    beam = HexBeam(nx, ny, nz, length, width, density=1000.0,
                   material=("neohookean", 3.4e4, 1.7e6), h=0.005)
    rng = np.random.default_rng(seed)
    X = beam.nodes
    s = (X[:, 0] - X[:, 0].min()) / length
    shape = s * s * (3.0 - s) / 2.0
    tipload = np.zeros(3)
    grav = beam.gravity_loads()
    u_prev = np.zeros((beam.nn, 3))
    out_v, out_b = [], []
    for k in range(nsteps):
        tipload = tipload + rng.uniform(-1.0, 1.0, size=3) * np.array([0.2, 1.0, 1.0])
        defl = 2e-4 * (k + 1) + 1e-4 * tipload  # tip deflection (m)
        u = np.zeros((beam.nn, 3))
        u[:, 1] = defl[1] * shape
        u[:, 2] = -abs(defl[2]) * shape - 2e-4 * (k + 1) * shape
        u[:, 0] = defl[0] * 0.01 * s
        v = (u - u_prev) / beam.h
        fext = grav.copy()
        fext[beam.tip] += tipload[None, :] / len(beam.tip)
        vals, b = beam.system(u, v, fext)
        out_v.append(vals)
        out_b.append(b)
        u_prev = u
    return beam, out_v, out_b


def large_beam_dynamics(solve, nsteps=20, seed=0x5EED, nx=100, ny=16, nz=16, length=1.0, width=0.16,
                        tip_scale=20.0):
    """Config L as an integrated ConstrainedBackwardEuler run: the hex beam of
    large_beam_sequence released from rest under gravity, its free end pushed by a
    seeded random-walk load (steps of up to tip_scale N), h = 0.005.  Step k's KKT
    system is assembled at the state (u_k, v_k) the previous steps produced and
    handed to solve(beam, vals, b) -- the caller's FP64 direct solve -- whose first sizeM
    entries are v_{k+1}; then u_{k+1} = u_k + h v_{k+1} (MechSystemSolver.java:
    4804-4939: velocity solve, then the position update).  Returns (beam,
    [vals_k], [b_k], tip deflection (m) after each step)."""
    beam = HexBeam(nx, ny, nz, length, width, density=1000.0,
                   material=("neohookean", 3.4e4, 1.7e6), h=0.005)
    rng = np.random.default_rng(seed)
    grav = beam.gravity_loads()
    u = np.zeros((beam.nn, 3))
    v = np.zeros((beam.nn, 3))
    tipload = np.zeros(3)
    out_v, out_b, tip = [], [], []
    for k in range(nsteps):
        tipload = tipload + tip_scale * rng.uniform(-1.0, 1.0, size=3) * np.array([0.2, 1.0, 1.0])
        fext = grav.copy()
        fext[beam.tip] += tipload[None, :] / len(beam.tip)
        vals, b = beam.system(u, v, fext)
        out_v.append(vals)
        out_b.append(b)
        x = np.asarray(solve(beam, vals, b))
        v = x[:beam.sizeM].reshape(-1, 3).copy()
        u = u + beam.h * v
        tip.append(float(np.linalg.norm(u[beam.tip], axis=1).mean()))
    return beam, out_v, out_b, tip


# ---------------------------------------------------------------------------
# Linear tetrahedral FEM blocks, rigid bodies and contact constraints
# ---------------------------------------------------------------------------
def tet_grid(nx, ny, nz, lx, ly, lz, origin=(0.0, 0.0, 0.0)):
    """Structured tet mesh: each cell of an nx*ny*nz box grid split into 6 tets
    along its main diagonal (conforming across cells)."""
    xs = np.linspace(0, lx, nx + 1) + origin[0]
    ys = np.linspace(0, ly, ny + 1) + origin[1]
    zs = np.linspace(0, lz, nz + 1) + origin[2]
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    nodes = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)

    def nid(i, j, k):
        return (k * (ny + 1) + j) * (nx + 1) + i

    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    e = [(1, 0, 0), (0, 1, 0), (0, 0, 1)]
    tets = []
    for p in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
        a = (0, 0, 0)
        b = tuple(a[q] + e[p[0]][q] for q in range(3))
        c = tuple(b[q] + e[p[1]][q] for q in range(3))
        tets.append(np.stack([nid(i, j, k), nid(i + b[0], j + b[1], k + b[2]),
                              nid(i + c[0], j + c[1], k + c[2]), nid(i + 1, j + 1, k + 1)], 1))
    return nodes, np.concatenate(tets, 0)


def delaunay_box(npts_x, ny, nz, lx, ly, lz, seed, jitter=0.25):
    """Seeded Delaunay tet mesh of a jittered lattice filling an lx*ly*lz box
    (interior points jittered, boundary points on the box faces)."""
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    xs = np.linspace(0, lx, npts_x)
    ys = np.linspace(0, ly, ny)
    zs = np.linspace(0, lz, nz)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    h = np.array([lx / (npts_x - 1), ly / (ny - 1), lz / (nz - 1)])
    inner = np.all((P > 1e-12) & (P < np.array([lx, ly, lz]) - 1e-12), axis=1)
    P[inner] += rng.uniform(-jitter, jitter, size=(inner.sum(), 3)) * h
    tri = Delaunay(P)
    T = tri.simplices.astype(np.int64)
    D = np.stack([P[T[:, 1]] - P[T[:, 0]], P[T[:, 2]] - P[T[:, 0]], P[T[:, 3]] - P[T[:, 0]]], 2)
    vol = np.linalg.det(D) / 6.0
    keep = np.abs(vol) > 1e-6 * np.abs(vol).mean()
    T = T[keep]
    flip = vol[keep] < 0
    T[flip] = T[flip][:, [0, 2, 1, 3]]
    return P, T


def tet_linear_stiffness(nodes, tets, lam, mu):
    """Element stiffness (E,12,12) and volumes of linear tets, isotropic material."""
    P0 = nodes[tets[:, 0]]
    Dm = np.stack([nodes[tets[:, 1]] - P0, nodes[tets[:, 2]] - P0, nodes[tets[:, 3]] - P0], 2)
    vol = np.abs(np.linalg.det(Dm)) / 6.0
    Dinv = np.linalg.inv(Dm)  # rows: gradients of N1..N3
    g = np.zeros((len(tets), 4, 3))
    g[:, 1:] = Dinv
    g[:, 0] = -Dinv.sum(1)
    I3 = np.eye(3)
    # Ke[a,i,b,k] = V (lam g_a[i] g_b[k] + mu (g_a.g_b) delta_ik + mu g_a[k] g_b[i])
    Ke = (lam * np.einsum("eai,ebk->eaibk", g, g) +
          mu * np.einsum("eaj,ebj,ik->eaibk", g, g, I3) +
          mu * np.einsum("eak,ebi->eaibk", g, g))
    Ke = Ke * vol[:, None, None, None, None]
    return Ke.reshape(len(tets), 12, 12), vol


def assemble_tet(nodes, tets, lam, mu, density):
    """Stiffness (scipy CSR, full) and lumped mass vector of a linear tet body."""
    Ke, vol = tet_linear_stiffness(nodes, tets, lam, mu)
    E = len(tets)
    dofs = (3 * tets[:, :, None] + np.arange(3)[None, None]).reshape(E, 12)
    r = np.repeat(dofs, 12, axis=1).ravel()
    c = np.tile(dofs, (1, 12)).ravel()
    n = 3 * len(nodes)
    K = sp.csr_matrix((Ke.ravel(), (r, c)), shape=(n, n))
    m = np.zeros(len(nodes))
    np.add.at(m, tets.ravel(), np.repeat(density * vol / 4.0, 4))
    return K, np.repeat(m, 3)


def mooney_rivlin_PA(F, c10, kappa):
    """First Piola stress P and tangent A = dP/dF (A[...,i,J,k,L]) of the compressible
    Mooney-Rivlin material with c01 = 0: W = c10 (J^-2/3 I1 - 3) + kappa/2 (J - 1)^2.
    With H = F^-T and Jm = J^-2/3:
      P = 2 c10 Jm (F - I1/3 H) + kappa (J - 1) J H
      A = 2 c10 Jm [d_ik d_JL - 2/3 (F_iJ H_kL + H_iJ F_kL) + 2/9 I1 H_iJ H_kL
                    + I1/3 H_iL H_kJ] + kappa [(2J - 1) J H_iJ H_kL - (J - 1) J H_iL H_kJ]
    At F = I this is linear elasticity with mu = 2 c10, lambda = kappa - 2 mu / 3."""
    J = np.linalg.det(F)
    H = np.swapaxes(np.linalg.inv(F), -1, -2)
    I1 = np.einsum("...iJ,...iJ->...", F, F)
    Jm = J ** (-2.0 / 3.0)
    a = (2.0 * c10 * Jm)[..., None, None]
    P = a * (F - (I1 / 3.0)[..., None, None] * H) + (kappa * (J - 1.0) * J)[..., None, None] * H
    I3 = np.eye(3)
    a4 = a[..., None, None]
    A = a4 * (np.einsum("ik,JL->iJkL", I3, I3)
              - (2.0 / 3.0) * (np.einsum("...iJ,...kL->...iJkL", F, H) + np.einsum("...iJ,...kL->...iJkL", H, F))
              + (2.0 / 9.0) * I1[..., None, None, None, None] * np.einsum("...iJ,...kL->...iJkL", H, H)
              + (I1 / 3.0)[..., None, None, None, None] * np.einsum("...iL,...kJ->...iJkL", H, H))
    A = A + (kappa * (2.0 * J - 1.0) * J)[..., None, None, None, None] * np.einsum("...iJ,...kL->...iJkL", H, H)
    A = A - (kappa * (J - 1.0) * J)[..., None, None, None, None] * np.einsum("...iL,...kJ->...iJkL", H, H)
    return P, A


class TetBody(HexBeam):
    """Linear-tet body (4 nodes, constant strain) with its x = min face attached by
    point constraints, compressible Mooney-Rivlin material (c10, c01 = 0, kappa);
    the KKT assembly (pattern, slot map, ConstrainedBackwardEuler values) is the
    hex beam's."""

    def __init__(self, nodes, tets, c10, kappa, density=1000.0, stiffness_damping=0.002, h=0.01):
        self.nodes = np.asarray(nodes, np.float64)
        self.elems = np.asarray(tets, np.int64)
        self.nn = len(self.nodes)
        self.sizeM = 3 * self.nn
        self.h = h
        self.beta = stiffness_damping
        self.material = ("mooney-rivlin", c10, kappa)
        P0 = self.nodes[self.elems[:, 0]]
        Dm = np.stack([self.nodes[self.elems[:, q]] - P0 for q in (1, 2, 3)], 2)
        self.vol = np.abs(np.linalg.det(Dm)) / 6.0
        Dinv = np.linalg.inv(Dm)
        g = np.zeros((len(self.elems), 4, 3))
        g[:, 1:] = Dinv
        g[:, 0] = -Dinv.sum(1)
        self.grads = g  # reference shape-function gradients, (E, 4, 3)
        m_node = np.zeros(self.nn)
        np.add.at(m_node, self.elems.ravel(), np.repeat(density * self.vol / 4.0, 4))
        self.mass = np.repeat(m_node, 3)
        xmin = self.nodes[:, 0].min()
        self.attached = np.nonzero(np.isclose(self.nodes[:, 0], xmin))[0]
        self.tip = np.nonzero(np.isclose(self.nodes[:, 0], self.nodes[:, 0].max()))[0]
        nG = 3 * len(self.attached)
        rows = (3 * self.attached[:, None] + np.arange(3)[None]).ravel()
        self.GT = sp.csr_matrix((np.ones(nG), (rows, np.arange(nG))), shape=(self.sizeM, nG))
        self.numG = nG
        self._build_pattern()

    def element_tangent(self, u):
        c10, kappa = self.material[1], self.material[2]
        g = self.grads
        F = np.eye(3)[None] + np.einsum("eai,eaJ->eiJ", u[self.elems], g)
        P, A = mooney_rivlin_PA(F, c10, kappa)
        E = len(self.elems)
        C = np.einsum("eiJkL,ebL->eiJbk", A, g)
        Ke = np.einsum("eaJ,eiJbk->eaibk", g, C).reshape(E, 12, 12) * self.vol[:, None, None]
        fint = np.einsum("eiJ,eaJ->eai", P, g).reshape(E, 12) * self.vol[:, None]
        return Ke, fint


def tet_precision_sequence(nsteps=4, dims=(121, 24, 24), seed=0x7E7, c10=1e3, kappa=1e5,
                           size=(0.1, 0.02, 0.02), h=0.01):
    """Precision workload: seeded Delaunay tet box (jittered lattice of dims points,
    size 0.1 x 0.02 x 0.02 m; (121, 24, 24) gives 69,696 nodes = 209,088 DOF),
    Mooney-Rivlin c10 = 1e3, c01 = 0, kappa = 1e5 (nu ~ 0.49), rho = 1000, its x = 0
    face attached.  The nsteps states follow a prescribed, growing bending of the box
    (tip deflection up to a few percent of its length) under seeded tip loads; each
    step's KKT system is assembled at its state.  Returns (body, [vals_k], [b_k])."""
    nodes, tets = delaunay_box(dims[0], dims[1], dims[2], size[0], size[1], size[2], seed)
    body = TetBody(nodes, tets, c10, kappa, density=1000.0, h=h)
    rng = np.random.default_rng(seed + 1)
    X = body.nodes
    s = (X[:, 0] - X[:, 0].min()) / size[0]
    shape = s * s * (3.0 - s) / 2.0
    grav = body.gravity_loads()
    u_prev = np.zeros((body.nn, 3))
    out_v, out_b = [], []
    for k in range(nsteps):
        defl = size[0] * 0.01 * (k + 1) * (1.0 + 0.2 * rng.uniform(-1, 1, 3))
        u = np.zeros((body.nn, 3))
        u[:, 1] = 0.3 * defl[1] * shape
        u[:, 2] = -defl[2] * shape
        v = (u - u_prev) / h
        fext = grav.copy()
        fext[body.tip] += rng.uniform(-1e-2, 1e-2, 3)[None, :] / len(body.tip)
        vals, b = body.system(u, v, fext)
        out_v.append(vals)
        out_b.append(b)
        u_prev = u
    return body, out_v, out_b


def rigid_contact_system(ndof, seed, h=0.01):
    """One system of the batching mix: 2-20 rigid bodies (6 DOF, random SPD spatial
    inertia) plus a small linear tet FEM block, with 10-200 contact rows treated as
    equalities (random unit normals; each row couples one rigid body to a distinct
    FEM node, so G has full row rank).  Returns (M_or_S, GT, rhs_m, rhs_g)."""
    rng = np.random.default_rng(seed)
    nb = int(rng.integers(2, 21))
    n_rigid = 6 * nb
    n_fem_nodes = max(8, (ndof - n_rigid) // 3)
    # tet box with about n_fem_nodes nodes
    a = max(1, int(round((n_fem_nodes / 2.0) ** (1.0 / 3.0))))
    dims = [2 * a - 1, a - 1, a - 1]
    while (dims[0] + 1) * (dims[1] + 1) * (dims[2] + 1) < n_fem_nodes:
        dims[int(np.argmin(dims))] += 1
    dims = [max(1, d) for d in dims]
    nodes, tets = tet_grid(dims[0], dims[1], dims[2], 0.02 * dims[0], 0.02 * dims[1], 0.02 * dims[2])
    lam, mu = lame(rng.uniform(5e4, 5e5), 0.3)
    K, mfem = assemble_tet(nodes, tets, lam, mu, 1000.0)
    nfem = K.shape[0]
    # rigid bodies: mass m I3 and a random SPD inertia tensor
    blocks = []
    for b in range(nb):
        m = rng.uniform(0.1, 2.0)
        Q = np.linalg.qr(rng.standard_normal((3, 3)))[0]
        J = Q @ np.diag(rng.uniform(1e-3, 1e-2, 3)) @ Q.T
        Mb = np.zeros((6, 6))
        Mb[:3, :3] = m * np.eye(3)
        Mb[3:, 3:] = J
        blocks.append(sp.csr_matrix(Mb))
    beta = 0.01
    Sfem = sp.diags(mfem) + (h * beta + h * h) * K
    S = sp.block_diag(blocks + [Sfem], format="csr")
    sizeM = S.shape[0]
    ncon = int(rng.integers(10, 201))
    ncon = min(ncon, len(nodes))
    fem_nodes = rng.choice(len(nodes), size=ncon, replace=False)
    rows, cols, vals = [], [], []
    centers = rng.uniform(-0.1, 0.1, size=(nb, 3))
    for q in range(ncon):
        nrm = rng.standard_normal(3)
        nrm /= np.linalg.norm(nrm)
        b = int(rng.integers(0, nb))
        pnt = centers[b] + rng.uniform(-0.05, 0.05, 3)
        ang = np.cross(pnt - centers[b], nrm)
        for d in range(3):
            rows.append(6 * b + d); cols.append(q); vals.append(nrm[d])
            rows.append(6 * b + 3 + d); cols.append(q); vals.append(ang[d])
            rows.append(n_rigid + 3 * fem_nodes[q] + d); cols.append(q); vals.append(-nrm[d])
    GT = sp.csr_matrix((vals, (rows, cols)), shape=(sizeM, ncon))
    bm = rng.uniform(-1, 1, sizeM)
    bg = np.zeros(ncon)
    return S, GT, bm, bg


def batch_mix(nsys=256, seed=0xBA7C, dof_lo=300, dof_hi=4000):
    """The batching workload: nsys independent KKT systems with DOF ~ U[dof_lo, dof_hi]."""
    rng = np.random.default_rng(seed)
    systems = []
    for s in range(nsys):
        ndof = int(rng.integers(dof_lo, dof_hi + 1))
        S, GT, bm, bg = rigid_contact_system(ndof, seed * 1000 + s)
        ro, ci, va = kkt_upper_csr(S, GT, None)
        systems.append((ro, ci, va, np.concatenate([bm, bg])))
    return systems


def block_diag_csr(systems):
    """Concatenates upper-CSR systems (rowOffs, colIdxs, vals, b) into one block-
    diagonal upper CSR; returns (rowOffs, colIdxs, vals, b, row_offsets_per_system)."""
    ros, cis, vas, bs, offs = [], [], [], [], [0]
    nnz_off = 0
    for ro, ci, va, b in systems:
        n = len(ro) - 1
        ros.append(ro[:-1] + nnz_off)
        cis.append(ci + offs[-1])
        vas.append(va)
        bs.append(b)
        nnz_off += int(ro[-1])
        offs.append(offs[-1] + n)
    ro = np.concatenate(ros + [np.array([nnz_off])]).astype(np.int32)
    return (ro, np.concatenate(cis).astype(np.int32), np.concatenate(vas), np.concatenate(bs),
            np.array(offs, np.int64))


def _beam_dims(ndof, aspect=5):
    """Hex grid (nx, ny, nz) whose node count * 3 is close to ndof, aspect nx ~ 5 ny."""
    nodes = max(8, ndof // 3)
    a = max(1, int(round((nodes / aspect) ** (1.0 / 3.0))) - 1)
    dims = [max(1, aspect * (a + 1) - 1), a, a]
    while (dims[0] + 1) * (dims[1] + 1) * (dims[2] + 1) < nodes:
        dims[0] += 1
    return dims


def serving_instance(kind, ndof, seed):
    """One simulation instance of the serving mix: upper-CSR pattern, base values
    and a right-hand side.  kind: 'beam' (linear hex beam, attached end), 'tet'
    (linear tet box, attached end) or 'rigid' (rigid bodies + tet block + contacts)."""
    rng = np.random.default_rng(seed)
    if kind == "rigid":
        S, GT, bm, bg = rigid_contact_system(min(ndof, 8000), seed)
        ro, ci, va = kkt_upper_csr(S, GT, None)
        return ro, ci, va, np.concatenate([bm, bg])
    if kind == "beam":
        nx, ny, nz = _beam_dims(ndof)
        beam = HexBeam(nx, ny, nz, 1.0, 0.2 * ny / max(nx, 1) * 5, material=("linear", 1e5, 0.33),
                       h=0.01)
        fext = beam.gravity_loads() + rng.uniform(-1, 1, (beam.nn, 3))
        u = np.zeros((beam.nn, 3))
        vals, b = beam.system(u, u, fext)
        return beam.rowOffs, beam.colIdxs, vals, b
    # tet box with its x = 0 face attached by point constraints
    nx, ny, nz = _beam_dims(ndof)
    nodes, tets = tet_grid(nx, ny, nz, 0.01 * nx, 0.01 * ny, 0.01 * nz)
    lam, mu = lame(1e5, 0.3)
    K, m = assemble_tet(nodes, tets, lam, mu, 1000.0)
    h = 0.01
    S = sp.diags(m) + (h * 0.002 + h * h) * K
    att = np.nonzero(np.isclose(nodes[:, 0], 0.0))[0]
    nG = 3 * len(att)
    rows = (3 * att[:, None] + np.arange(3)[None]).ravel()
    GT = sp.csr_matrix((np.ones(nG), (rows, np.arange(nG))), shape=(S.shape[0], nG))
    ro, ci, va = kkt_upper_csr(S, GT, None)
    b = np.concatenate([rng.uniform(-1, 1, S.shape[0]) * 1e-2, np.zeros(nG)])
    return ro, ci, va, b


def serving_mix(ninst=64, seed=0x5E7E, dof_lo=1000, dof_hi=90000):
    """The serving workload: ninst instances, DOF log-uniform in [dof_lo, dof_hi], kinds cycling
    beam / tet / rigid."""
    rng = np.random.default_rng(seed)
    out = []
    kinds = ("beam", "tet", "rigid")
    for i in range(ninst):
        ndof = int(np.exp(rng.uniform(np.log(dof_lo), np.log(dof_hi))))
        kind = kinds[i % 3]
        out.append((kind, ndof) + serving_instance(kind, ndof, seed * 100 + i))
    return out


def perturbed_values(vals, step, seed, eps=1e-3):
    """Values of a later step of an instance: seeded relative perturbation of every
    entry (zero entries stay zero, the pattern and symmetry are unchanged)."""
    rng = np.random.default_rng(seed * 7919 + step)
    return vals * (1.0 + eps * rng.uniform(-1, 1, len(vals)))
