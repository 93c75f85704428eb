"""KKT problems on disk in the format of KKTSolver.printLinearProblem
(KKTSolver.java:1746-1788) / pardisoMklTest.cc:215-300."""
import numpy as np

import oracle as O
from artisynth_core_amd import generators as G
from artisynth_core_amd.kktio import read_linear_problem, write_linear_problem
from kkt_helpers import P5_B, P5_COLS, P5_ROWS, P5_VALS, rel_residual


def test_round_trip_is_exact(tmp_path):
    beam, vals, b = G.small_beam(8)
    p = tmp_path / "beam.txt"
    write_linear_problem(p, beam.rowOffs, beam.colIdxs, vals, b[:beam.sizeM], b[beam.sizeM:],
                         comment="small beam")
    r = read_linear_problem(p)
    assert r["symmetric"] and r["rhs_present"] and r["n"] == beam.n
    assert np.array_equal(r["rowOffs"], beam.rowOffs)
    assert np.array_equal(r["colIdxs"], beam.colIdxs)
    assert np.array_equal(r["vals"], vals) and np.array_equal(r["rhs"], b)
    # the problem read back solves like the original (oracle, FP64)
    x = O.Factor(r["rowOffs"], r["colIdxs"], r["vals"]).solve(r["rhs"])
    assert rel_residual(beam.rowOffs, beam.colIdxs, vals, x, b) <= 1e-10


def test_reads_the_reference_layout(tmp_path):
    # hand-written in the reference's layout: comment line, SYMMETRIC, one-based
    # offsets and indices, "%g" values, no right-hand side (the reader uses ones)
    ro = np.asarray(P5_ROWS)
    ci = np.asarray(P5_COLS)
    v5, x5 = P5_VALS[0]
    txt = ["# 5x5 indefinite (PardisoSolverTest)", "SYMMETRIC 5",
           " ".join(str(v + 1) for v in ro), " ".join(str(v + 1) for v in ci),
           " ".join("%g" % v for v in v5)]
    p = tmp_path / "p5.txt"
    p.write_text("\n".join(txt) + "\n")
    r = read_linear_problem(p)
    assert r["symmetric"] and not r["rhs_present"]
    assert np.array_equal(r["rhs"], np.ones(5))
    assert np.array_equal(r["rowOffs"], ro) and np.array_equal(r["colIdxs"], ci)
    assert np.array_equal(r["vals"], v5)
    # with the test's right-hand side the known answer comes back (oracle, FP64)
    x = O.Factor(r["rowOffs"], r["colIdxs"], r["vals"]).solve(P5_B)
    assert np.allclose(x, x5, rtol=0, atol=1e-12)


def test_omit_lower_right_diagonal(tmp_path):
    # the constraint rows' diagonal entries (the last numG values) are left out and
    # the offsets clamped to nnz + 1; the reader can put explicit zeros back
    beam, vals, b = G.small_beam(8)
    p = tmp_path / "omit.txt"
    write_linear_problem(p, beam.rowOffs, beam.colIdxs, vals, b[:beam.sizeM], b[beam.sizeM:],
                         omit_lower_right_diagonal=True, num_g=beam.numG)
    r = read_linear_problem(p)
    assert int(r["rowOffs"][-1]) == beam.nnz - beam.numG
    assert np.all(np.diff(r["rowOffs"])[beam.sizeM:] == 0)
    r2 = read_linear_problem(p, add_missing_diagonal=True)
    assert np.array_equal(r2["rowOffs"], beam.rowOffs)
    assert np.array_equal(r2["colIdxs"], beam.colIdxs)
    assert np.array_equal(r2["vals"], vals)  # the omitted diagonal values are 0
