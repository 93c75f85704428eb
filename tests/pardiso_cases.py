"""The MKL PARDISO fixtures (tests/golden/pardiso_fixtures.npz, written by
tests/golden/make_pardiso.py): case inputs rebuilt from the seeded generators and
checked against the fixture's input hashes, and the fixture values per case."""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

_fix = None


def fixtures():
    global _fix
    if _fix is None:
        _fix = dict(np.load(os.path.join(HERE, "golden", "pardiso_fixtures.npz")))
    return _fix


def _sha(a):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a, np.float64).tobytes()).digest(), np.uint8)


SMALL = ["p5_0", "p5_1", "p5_2", "kkt5", "beam8", "beam26", "testMatrix", "beam24x8x8", "tet21x6x6"]
ALL = SMALL + ["beamL"]


def case(name):
    """(rowOffs, colIdxs, vals, b, hybrid or None, fixture dict for the case)."""
    import make_pardiso
    for nm, ro, ci, va, b, hyb, sampled in make_pardiso.cases():
        if nm != name:
            continue
        F = fixtures()
        fx = {k.split("/", 1)[1]: v for k, v in F.items() if k.startswith(name + "/")}
        assert np.array_equal(_sha(va), fx["vals_sha256"]), "inputs of %s changed since the fixture" % name
        assert np.array_equal(_sha(b), fx["b_sha256"]), "right-hand side of %s changed" % name
        if hyb is not None:
            assert np.array_equal(_sha(hyb[0]), fx["vals2_sha256"])
        return (np.ascontiguousarray(ro, np.int32), np.ascontiguousarray(ci, np.int32),
                np.ascontiguousarray(va, np.float64), np.ascontiguousarray(b, np.float64), hyb, fx)
    raise KeyError(name)


def sampled(fx, x, key="x"):
    """x restricted to the fixture's sample (if the case is sampled) and the
    relative error against the fixture's values there."""
    ref = fx[key]
    if "x_idx" in fx:
        x = np.asarray(x)[fx["x_idx"]]
    return float(np.linalg.norm(x - ref) / max(np.linalg.norm(ref), 1e-300))


def info(fx):
    i = fx["info"]
    return {"npos": int(i[0]), "nneg": int(i[1]), "nperturbed": int(i[2]), "nnzL": int(i[3]),
            "refine_steps": int(i[4])}
