"""Writes tests/golden/pardiso_fixtures.npz: MKL PARDISO results, with the
reference's PARDISO settings, for the systems the parity tests use.

Run in the build container (MKL 2021.4 under /opt/conda/lib; the reference pins
2021.1.1, src/maspack/solvers/lib/Makefile:43-48):

    python tests/golden/make_pardiso.py

PARDISO is driven as the reference's Pardiso4 drives it (oracle/pardiso.py:
mtype=-2, parallel METIS, Bunch-Kaufman, perturbation 1e-8, scaling and matching,
one refinement step; phases 11, 22, 33 and the stale-factor CGS phase 33 with
iparm[3]=102), with one MKL thread so that METIS, and with it nnz(L), is
deterministic.  Per case the fixture holds:
  x            phase-33 solution (full, or a seeded sample of 2,000 entries plus
               ||x|| for the large beam)
  info         [npos iparm[21], nneg iparm[22], nperturbed iparm[13], nnzL iparm[17],
               refinement steps iparm[6]]
  cgs_iters    iparm[19] of a stale-factor iterative solve (cases with a hybrid step)
  x_hybrid     that solve's result (sampled like x)
  vals_sha256  SHA-256 of the float64 value bytes the case was built from, so a
               test that regenerates the inputs can check they are the same
Inputs come from the repo's seeded generators (and tests/golden/testMatrix_upper.npz),
so the tests rebuild them rather than reading them here.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

SAMPLE = 2000


def sha(a):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a, np.float64).tobytes()).digest(), np.uint8)


def cases():
    """(name, rowOffs, colIdxs, vals, b, hybrid (vals2, b2) or None, sampled)."""
    from artisynth_core_amd import generators as G
    from kkt_helpers import KKT5_B, KKT5_COLS, KKT5_ROWS, KKT5_VALS, P5_B, P5_COLS, P5_ROWS, P5_VALS
    for q, (v, _) in enumerate(P5_VALS):
        yield "p5_%d" % q, P5_ROWS, P5_COLS, v, P5_B, None, False
    yield "kkt5", KKT5_ROWS, KKT5_COLS, KKT5_VALS, KKT5_B, None, False
    for nx in (8, 26):
        beam, vals, b = G.small_beam(nx)
        yield "beam%d" % nx, beam.rowOffs, beam.colIdxs, vals, b, (vals * 1.05, b), False
    d = np.load(os.path.join(HERE, "testMatrix_upper.npz"))
    ro, ci, va = d["rowOffs"], d["colIdxs"], d["vals"]
    from kkt_helpers import full_from_upper
    A = full_from_upper(ro, ci, va)
    norm = np.sqrt((A.data ** 2).sum() / A.nnz)
    rng = np.random.default_rng(0x1234)
    va2 = va + (rng.random(len(va)) - 0.5) * (norm / 1e4)
    n = len(ro) - 1
    yield "testMatrix", ro, ci, va, np.ones(n), (va2, np.ones(n)), False
    beam = G.HexBeam(24, 8, 8, 1.0, 0.2, material=("neohookean", 3.4e4, 1.7e6), h=0.005)
    rng = np.random.default_rng(5)
    u = rng.uniform(-1e-4, 1e-4, (beam.nn, 3))
    vals, b = beam.system(u, np.zeros_like(u), beam.gravity_loads())
    yield "beam24x8x8", beam.rowOffs, beam.colIdxs, vals, b, None, False
    body, V, B = G.tet_precision_sequence(nsteps=2, dims=(21, 6, 6))
    yield "tet21x6x6", body.rowOffs, body.colIdxs, V[1], B[1], (V[0], B[0]), False
    beam, V, B = G.large_beam_sequence(nsteps=2)
    yield "beamL", beam.rowOffs, beam.colIdxs, V[0], B[0], (V[1], B[1]), True


def main():
    import pardiso as P
    ok, what = P.available()
    if not ok:
        raise SystemExit(what)
    P.set_threads(1)
    out = {}
    meta = {"mkl": what, "threads": 1,
            "iparm": {"1": 3, "7": 1, "9": 8, "10": 1, "12": 1, "20": 1, "26": 0}, "cases": []}
    for name, ro, ci, va, b, hyb, sampled in cases():
        p = P.Pardiso(ro, ci, va)
        p.factor()
        x = p.solve(b)
        inf = p.info()
        n = len(ro) - 1
        idx = np.sort(np.random.default_rng(0xF1C5).choice(n, SAMPLE, replace=False)) if sampled else None
        out[name + "/vals_sha256"] = sha(va)
        out[name + "/b_sha256"] = sha(b)
        out[name + "/info"] = np.array([inf["npos"], inf["nneg"], inf["nperturbed"], inf["nnzL"],
                                        inf["refine_steps"]], np.int64)
        out[name + "/x"] = x[idx] if sampled else x
        out[name + "/x_norm"] = np.float64(np.linalg.norm(x))
        if sampled:
            out[name + "/x_idx"] = idx.astype(np.int32)
        line = "%s: n=%d %s" % (name, n, inf)
        if hyb is not None:
            va2, b2 = hyb
            out[name + "/vals2_sha256"] = sha(va2)
            it, x2 = p.iterative_solve(va2, b2, 10)
            out[name + "/cgs_iters"] = np.int64(it)
            out[name + "/x_hybrid"] = x2[idx] if sampled else x2
            out[name + "/x_hybrid_norm"] = np.float64(np.linalg.norm(x2))
            line += " cgs_iters=%d" % it
        p.release()
        meta["cases"].append(name)
        print(line, flush=True)
    out["meta_json"] = np.frombuffer(json.dumps(meta).encode(), np.uint8)
    np.savez_compressed(os.path.join(HERE, "pardiso_fixtures.npz"), **out)


if __name__ == "__main__":
    main()
