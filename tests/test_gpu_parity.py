"""GPU parity tests: the HIP path through the C ABI against the CPU oracle and the
reference's known answers.  Marked gpu; run with `pytest -m gpu`."""
import sys

import numpy as np
import pytest

import oracle as O
from artisynth_core_amd import KKT_PREC_FP32_REFINE, KKT_PREC_FP64
from artisynth_core_amd import generators as G
from artisynth_core_amd.solvers import ImproperStateException, MI355XSolver, Matrix, NumericalException
from kkt_helpers import (IND3_COLS, IND3_ROWS, IND3_VALS, KKT5_B, KKT5_COLS, KKT5_ROWS, KKT5_VALS,
                         P5_B, P5_COLS, P5_ROWS, P5_VALS, SPD5_COLS, SPD5_RHS, SPD5_ROWS, SPD5_VALS,
                         SPD5_X, full_from_upper, rel_err, rel_residual)

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[gpu-case] " + msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec", [KKT_PREC_FP64, KKT_PREC_FP32_REFINE])
def test_pardiso_5x5_known_answers(gpu, prec):
    # PardisoSolverTest.java:407-543: inertia 2-/3+, 0 perturbed pivots, 15-digit x
    _say("pardiso5x5 prec=%d" % prec)
    s = MI355XSolver(precision=prec)
    try:
        s.analyzeArrays(P5_VALS[0][0], P5_COLS, P5_ROWS, 5, Matrix.SYMMETRIC)
        for vals, xchk in P5_VALS:
            s.factor(vals)
            x = np.zeros(5)
            s.solve(x, P5_B)
            assert np.abs(x - xchk).max() < 1e-12, (x, xchk)
        s.factor(P5_VALS[0][0])
        assert s.getNumNegEigenvalues() == 2
        assert s.getNumPosEigenvalues() == 3
        assert s.getNumPerturbedPivots() == 0
        # multiple right-hand sides, column-major (PardisoSolverTest.java:479-497)
        B = np.concatenate([P5_B, 2 * P5_B, 4 * P5_B])
        X = np.zeros(15)
        s.solve(X, B, 3)
        xc = P5_VALS[0][1]
        assert np.abs(X - np.concatenate([xc, 2 * xc, 4 * xc])).max() < 1e-12
    finally:
        s.dispose()


def test_cudss_fixtures(gpu):
    _say("cudss fixtures")
    s = MI355XSolver()
    try:
        s.analyzeArrays(SPD5_VALS, SPD5_COLS, SPD5_ROWS, 5, Matrix.SPD)
        s.factor(SPD5_VALS)
        x = np.zeros(5)
        s.solve(x, SPD5_RHS)
        assert np.abs(x - SPD5_X).max() < 1e-12
        s.factor(SPD5_VALS * 2.5)  # refactor with scaled values (CuDssSolverTest.java:190-217)
        s.solve(x, SPD5_RHS)
        assert np.abs(x - SPD5_X / 2.5).max() < 1e-12
        s.analyzeArrays(IND3_VALS, IND3_COLS, IND3_ROWS, 3, Matrix.SYMMETRIC)
        s.factor(IND3_VALS)
        b = np.array([1.0, 2.0, 3.0])
        s.solve(x, b)
        A = full_from_upper(IND3_ROWS, IND3_COLS, IND3_VALS).toarray()
        assert np.abs(x[:3] - np.linalg.solve(A, b)).max() < 1e-13
        assert s.getNumNegEigenvalues() == 1 and s.getNumPosEigenvalues() == 2
        s.analyzeArrays(KKT5_VALS, KKT5_COLS, KKT5_ROWS, 5, Matrix.SYMMETRIC)
        s.factor(KKT5_VALS)
        x = np.zeros(5)
        s.solve(x, KKT5_B)
        assert rel_residual(KKT5_ROWS, KKT5_COLS, KKT5_VALS, x, KKT5_B) <= 1e-9
    finally:
        s.dispose()


def test_state_errors(gpu):
    _say("state errors")
    s = MI355XSolver()
    with pytest.raises(ImproperStateException):
        s.solve(np.zeros(5), SPD5_RHS)
    s.analyzeArrays(SPD5_VALS, SPD5_COLS, SPD5_ROWS, 5, Matrix.SPD)
    with pytest.raises(ImproperStateException):
        s.solve(np.zeros(5), SPD5_RHS)
    assert s.iterativeSolve(SPD5_VALS, np.zeros(5), SPD5_RHS, 10) <= 0
    s.dispose()
    s.dispose()  # idempotent (CuDssSolverTest.java:333-338)


@pytest.mark.parametrize("nx", [8, 26])
@pytest.mark.parametrize("prec", [KKT_PREC_FP64, KKT_PREC_FP32_REFINE])
def test_small_beam_vs_oracle(gpu, nx, prec):
    _say("small beam nx=%d prec=%d" % (nx, prec))
    beam, vals, b = G.small_beam(nx)
    s = MI355XSolver(precision=prec)
    try:
        s.analyzeArrays(vals, beam.colIdxs, beam.rowOffs, beam.n, Matrix.SYMMETRIC)
        assert s.verifyDeviceStructure()
        s.factor(vals)
        x = np.zeros(beam.n)
        s.solve(x, b)
        sym = s.symbolic()
        ref = O.Factor(beam.rowOffs, beam.colIdxs, vals, perm=sym["perm"], sup_ptr=sym["sup_ptr"])
        xo = ref.solve(b)
        assert rel_residual(beam.rowOffs, beam.colIdxs, vals, x, b) <= 1e-8
        assert rel_err(x, xo) <= 1e-8
        inf = s.info()
        oi = ref.info()
        assert (inf["npos"], inf["nneg"]) == (oi["npos"], oi["nneg"])
        assert inf["nnzL"] == oi["nnzL"]
        if prec == KKT_PREC_FP64:
            # same panel-restricted Bunch-Kaufman decisions as the oracle
            assert (inf["n2x2"], inf["nperturbed"]) == (oi["n2x2"], oi["nperturbed"])
    finally:
        s.dispose()


@pytest.mark.parametrize("prec", [KKT_PREC_FP64, KKT_PREC_FP32_REFINE])
def test_delaunay_tet_kkt_vs_oracle(gpu, prec):
    # unstructured mesh (the precision workload's generator at a small lattice):
    # Mooney-Rivlin tet box in a bent state, attached end
    _say("delaunay tet 21x6x6 prec=%d" % prec)
    body, V, B = G.tet_precision_sequence(nsteps=2, dims=(21, 6, 6))
    vals, b = V[1], B[1]
    s = MI355XSolver(precision=prec)
    try:
        s.analyzeArrays(vals, body.colIdxs, body.rowOffs, body.n, Matrix.SYMMETRIC)
        assert s.verifyDeviceStructure()
        s.factor(vals)
        x = np.zeros(body.n)
        s.solve(x, b)
        sym = s.symbolic()
        ref = O.Factor(body.rowOffs, body.colIdxs, vals, perm=sym["perm"], sup_ptr=sym["sup_ptr"])
        xo = ref.solve(b)
        assert rel_residual(body.rowOffs, body.colIdxs, vals, x, b) <= 1e-8
        assert rel_err(x, xo) <= 1e-8
        inf, oi = s.info(), ref.info()
        assert (inf["npos"], inf["nneg"]) == (oi["npos"], oi["nneg"]) == (body.sizeM, body.numG)
    finally:
        s.dispose()


@pytest.mark.parametrize("prec", [KKT_PREC_FP64, KKT_PREC_FP32_REFINE])
def test_forward_forms_are_bitwise_equal(gpu, prec):
    # the forward sweep of a level runs as one launch (every row block of a
    # large front gathers the front's right-hand side itself) or as two (a gather
    # launch, then the rows); both add the same terms in the same order, so the
    # solve (triangular solves and refinement) is bitwise the same for any mix of
    # forms, on graph replay and eager launches alike, and matches the oracle
    _say("forward forms prec=%d" % prec)
    body, V, B = G.tet_precision_sequence(nsteps=2, dims=(21, 6, 6))
    vals, b = V[1], B[1]
    s = MI355XSolver(precision=prec)
    try:
        s.analyzeArrays(vals, body.colIdxs, body.rowOffs, body.n, Matrix.SYMMETRIC)
        s.factor(vals)
        nl = s.info()["nlevels"]
        assert list(s.getForwardForms()) == [1] * nl  # the default
        xs = []
        rng = np.random.default_rng(5)
        plans = [[1] * nl, [2] * nl, list(rng.integers(1, 3, nl)), [1] * nl]
        groups = [[1] * nl, [8] * nl, list(2 ** rng.integers(0, 4, nl)), [2] * nl]
        for graphs in (1, 0):
            s._L.kkt_set_graphs(s._h, graphs)
            for forms, rg in zip(plans, groups):
                s.setForwardForms(forms)
                s.setForwardRowGroups(rg)
                assert list(s.getForwardForms()) == [int(f) for f in forms]
                assert list(s.getForwardRowGroups()) == [int(g) for g in rg]
                x = np.zeros(body.n)
                s.solve(x, b)
                xs.append(x)
        for x in xs[1:]:
            assert np.array_equal(x, xs[0])
        sym = s.symbolic()
        ref = O.Factor(body.rowOffs, body.colIdxs, vals, perm=sym["perm"], sup_ptr=sym["sup_ptr"])
        assert rel_residual(body.rowOffs, body.colIdxs, vals, xs[0], b) <= 1e-8
        assert rel_err(xs[0], ref.solve(b)) <= 1e-8
        with pytest.raises(Exception):
            s.setForwardForms([3] * nl)
        with pytest.raises(Exception):
            s.setForwardRowGroups([3] * nl)
    finally:
        s.dispose()


def test_iterative_solve_stale_factor(gpu):
    # CuDssSolverTest.java:993-1094 / PardisoSolverTest.java:668-688: values perturbed
    # after the factor; BiCGStab with the stale factor reaches 1e-9
    _say("iterative stale factor")
    beam, vals, b = G.small_beam(26)
    s = MI355XSolver()
    try:
        s.analyzeArrays(vals, beam.colIdxs, beam.rowOffs, beam.n, Matrix.SYMMETRIC)
        s.factor(vals)
        v2 = vals * 1.05
        x = np.zeros(beam.n)
        it = s.iterativeSolve(v2, x, b, 10)
        assert it > 0
        assert rel_residual(beam.rowOffs, beam.colIdxs, v2, x, b) <= 1e-9
        sym = s.symbolic()
        ref = O.Factor(beam.rowOffs, beam.colIdxs, vals, perm=sym["perm"], sup_ptr=sym["sup_ptr"])
        it_o, xo = ref.bicgstab(b, 1e-10, 32, vals=v2)
        assert it_o > 0
        assert abs(it - it_o) <= 1
        assert rel_err(x, xo) <= 1e-8
    finally:
        s.dispose()


@pytest.mark.parametrize("prec", [KKT_PREC_FP64, KKT_PREC_FP32_REFINE])
def test_multipanel_beam_vs_oracle(gpu, prec):
    # fronts wider than one 64-column panel and several tree levels
    _say("multipanel beam 24x8x8 prec=%d" % prec)
    beam = G.HexBeam(24, 8, 8, 1.0, 0.2, material=("neohookean", 3.4e4, 1.7e6), h=0.005)
    rng = np.random.default_rng(5)
    u = rng.uniform(-1e-4, 1e-4, (beam.nn, 3))
    vals, b = beam.system(u, np.zeros_like(u), beam.gravity_loads())
    s = MI355XSolver(precision=prec)
    try:
        s.analyzeArrays(vals, beam.colIdxs, beam.rowOffs, beam.n, Matrix.SYMMETRIC)
        sym = s.symbolic()
        assert np.diff(sym["sup_ptr"]).max() > 64
        assert s.verifyDeviceStructure()
        s.factor(vals)
        x = np.zeros(beam.n)
        s.solve(x, b)
        ref = O.Factor(beam.rowOffs, beam.colIdxs, vals, perm=sym["perm"], sup_ptr=sym["sup_ptr"])
        xo = ref.solve(b)
        assert rel_residual(beam.rowOffs, beam.colIdxs, vals, x, b) <= 1e-8
        assert rel_err(x, xo) <= 1e-7
        inf, oi = s.info(), ref.info()
        assert (inf["npos"], inf["nneg"]) == (oi["npos"], oi["nneg"]) == (beam.sizeM, beam.numG)
        if prec == KKT_PREC_FP64:
            assert (inf["n2x2"], inf["nperturbed"]) == (oi["n2x2"], oi["nperturbed"])
        # bitwise reproducible: no order-dependent atomics
        x2 = np.zeros(beam.n)
        s.factor(vals)
        s.solve(x2, b)
        assert np.array_equal(x, x2)
    finally:
        s.dispose()


def test_test_matrix_on_gpu(gpu):
    # PardisoSolverTest.java:608-688 on the GPU path
    _say("testMatrix")
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "testMatrix_upper.npz"))
    ro, ci, va = d["rowOffs"], d["colIdxs"], d["vals"]
    n = len(ro) - 1
    s = MI355XSolver()
    try:
        s.analyzeArrays(va, ci, ro, n, Matrix.SYMMETRIC)
        s.factor(va)
        b = np.ones(n)
        x = np.zeros(n)
        s.solve(x, b)
        A = full_from_upper(ro, ci, va)
        assert np.abs(A @ x - b).max() <= 1e-10
        norm = np.sqrt((A.data ** 2).sum() / A.nnz)
        rng = np.random.default_rng(0x1234)
        va2 = va + (rng.random(len(va)) - 0.5) * (norm / 1e4)
        x2 = np.zeros(n)
        it = s.iterativeSolve(va2, x2, b, 10)
        assert it > 0
        assert np.abs(full_from_upper(ro, ci, va2) @ x2 - b).max() <= 1e-9
    finally:
        s.dispose()


def test_batched_systems_vs_oracle(gpu):
    # several uneven rigid-body/tet/contact systems factored as one batch: each
    # system's slice matches the oracle on that system alone
    _say("batched mix 12 systems")
    from artisynth_core_amd.batch import BatchedKKT
    systems = G.batch_mix(12, seed=77, dof_lo=300, dof_hi=1500)
    B = BatchedKKT(systems)
    try:
        B.factor()
        x = B.solve()
        res = B.per_system_residuals(x, systems)
        assert max(res) <= 1e-8
        for q, (ro, ci, va, b) in enumerate(systems):
            xo = O.Factor(ro, ci, va).solve(b)
            assert rel_err(B.split(x)[q], xo) <= 1e-8
    finally:
        B.dispose()


def test_batched_autotuned_shapes(gpu):
    # the autotuner on a batch: the tuned solve still matches the oracle per
    # system, and the tuned shapes are within the kernels' limits
    _say("batched mix 12 systems, autotuned solve shapes")
    from artisynth_core_amd.batch import BatchedKKT
    systems = G.batch_mix(12, seed=78, dof_lo=300, dof_hi=1500)
    B = BatchedKKT(systems)
    try:
        B.factor()
        x0 = B.solve()
        B.autotune_solve(2)
        x = B.solve()
        assert max(B.per_system_residuals(x, systems)) <= 1e-8
        assert rel_err(x, x0) <= 1e-9
        for q, (ro, ci, va, b) in enumerate(systems):
            xo = O.Factor(ro, ci, va).solve(b)
            assert rel_err(B.split(x)[q], xo) <= 1e-8
    finally:
        B.dispose()


def test_serving_alone_equals_mix(gpu):
    # each instance's results are the same alone and in a mix (no eviction: the
    # same step kinds in the same order on each instance's own handle)
    _say("serving alone vs mix")
    from artisynth_core_amd.serving import SolveServer
    mix = [("beam", 900), ("tet", 1200), ("rigid", 1500), ("beam", 2500)]
    inst = [G.serving_instance(k, nd, 900 + i) for i, (k, nd) in enumerate(mix)]
    steps = 3
    seq = {i: [(G.perturbed_values(inst[i][2], s_, i), inst[i][3]) for s_ in range(steps)]
           for i in range(len(inst))}
    alone = {}
    for i in range(len(inst)):
        srv = SolveServer(nworkers=1)
        srv.register(i, inst[i][0], inst[i][1], inst[i][2])
        rs = [srv.submit(i, v, b) for v, b in seq[i]]
        for r in rs:
            r.wait()
            assert r.error is None, r.error
        alone[i] = [(r.x, r.kind) for r in rs]
        srv.shutdown()
    srv = SolveServer(nworkers=3)
    for i in range(len(inst)):
        srv.register(i, inst[i][0], inst[i][1], inst[i][2])
    rs = []
    for s_ in range(steps):
        for i in range(len(inst)):
            rs.append((i, s_, srv.submit(i, *seq[i][s_])))
    for i, s_, r in rs:
        r.wait()
        assert r.error is None, r.error
        assert r.kind == alone[i][s_][1]
        assert rel_err(r.x, alone[i][s_][0]) <= 1e-12
        assert rel_residual(inst[i][0], inst[i][1], seq[i][s_][0], r.x, seq[i][s_][1]) <= 1e-8
    st = srv.stats()
    assert st["evictions"] == 0 and st["admissions"] == len(inst) and st["completed"] == len(rs)
    srv.shutdown()


@pytest.mark.parametrize("prec", [KKT_PREC_FP32_REFINE, KKT_PREC_FP64])
def test_serving_device_inputs_equal_host_inputs(gpu, prec):
    # kkt_serve_submit_device (values, right-hand side and solution in HBM, read
    # in place) gives bitwise the results of kkt_serve_submit (host arrays copied
    # at submission) for the same instance and step sequence, in both precisions
    _say("serving device vs host inputs prec=%d" % prec)
    import torch
    from artisynth_core_amd.serving import SolveServer
    dev = torch.device("cuda:0")
    inst = [G.serving_instance("tet", 1500, 71), G.serving_instance("rigid", 1200, 72)]
    steps = 3
    seq = {i: [(G.perturbed_values(inst[i][2], s_, 50 + i, eps=3e-2), inst[i][3]) for s_ in range(steps)]
           for i in range(len(inst))}
    host, device = {}, {}
    srv = SolveServer(nworkers=2, precision=prec)
    try:
        for i in range(len(inst)):
            srv.register(i, inst[i][0], inst[i][1], inst[i][2])
        for i in range(len(inst)):
            rs = [srv.submit(i, v, b) for v, b in seq[i]]
            host[i] = [(r.wait().x, r.kind, r.error) for r in rs]
    finally:
        srv.shutdown()
    srv = SolveServer(nworkers=2, precision=prec)
    try:
        for i in range(len(inst)):
            srv.register(i, inst[i][0], inst[i][1], inst[i][2])
        for i in range(len(inst)):
            keep = []
            rs = []
            for v, b in seq[i]:
                dv = torch.from_numpy(np.ascontiguousarray(v)).to(dev)
                db = torch.from_numpy(np.ascontiguousarray(b)).to(dev)
                dx = torch.zeros(len(b), dtype=torch.float64, device=dev)
                keep.append((dv, db, dx))
                rs.append(srv.submit_device(i, dv.data_ptr(), db.data_ptr(), dx.data_ptr(), keep=(dv, db, dx)))
            device[i] = []
            for r, (dv, db, dx) in zip(rs, keep):
                r.wait()
                device[i].append((dx.cpu().numpy(), r.kind, r.error))
    finally:
        srv.shutdown()
    for i in range(len(inst)):
        kinds = [k for _, k, _ in host[i]]
        assert "direct" in kinds  # the first step of an instance factors
        for (xh, kh, eh), (xd, kd, ed) in zip(host[i], device[i]):
            assert eh is None and ed is None, (eh, ed)
            assert kh == kd
            assert np.array_equal(xh, xd)
        for s_, (xh, _, _) in enumerate(host[i]):
            assert rel_residual(inst[i][0], inst[i][1], seq[i][s_][0], xh, seq[i][s_][1]) <= 1e-8


def test_serving_eviction_and_readmission(gpu):
    # 18 instances under a budget that holds about four of them factored: requests
    # free idle instances' memory (least recently used first: factors released,
    # analyses kept, then analysis-only instances destroyed) and bring them back;
    # every result meets the tolerance and matches the instance's solve alone (an
    # instance that lost its factor refactors on its next step, so it matches to
    # the refinement tolerance rather than bitwise); an instance larger than the
    # whole budget fails its requests with KKT_ERR_ALLOC while the rest go on
    _say("serving eviction")
    import ctypes
    from artisynth_core_amd._lib import KktInfo, lib
    from artisynth_core_amd._lib import ptr as nptr
    from artisynth_core_amd.serving import SolveServer
    kinds = ("beam", "tet", "rigid")
    inst = [G.serving_instance(kinds[i % 3], 700 + 150 * i, 1700 + i) for i in range(18)]
    L = lib()
    need = []
    for ro, ci, va, b in inst:
        h = L.kkt_create()
        L.kkt_set_precision(h, KKT_PREC_FP32_REFINE)
        ro32, ci32 = np.ascontiguousarray(ro, np.int32), np.ascontiguousarray(ci, np.int32)
        va64 = np.ascontiguousarray(va)
        assert L.kkt_set_pattern(h, len(ro) - 1, int(ro[-1]), nptr(ro32, ctypes.c_int), nptr(ci32, ctypes.c_int), 1) == 0
        assert L.kkt_analyze_values(h, nptr(va64, ctypes.c_double)) == 0
        inf = KktInfo()
        L.kkt_get_info(h, ctypes.byref(inf))
        need.append(inf.factored_bytes)
        L.kkt_destroy(h)
    budget = int(sum(sorted(need)[-4:]) * 1.05)
    steps = 3
    seq = {i: [(G.perturbed_values(inst[i][2], s_, i, eps=2e-3), inst[i][3]) for s_ in range(steps)]
           for i in range(len(inst))}
    alone = {}
    srv = SolveServer(nworkers=1)
    for i in range(len(inst)):
        srv.register(i, inst[i][0], inst[i][1], inst[i][2])
    for i in range(len(inst)):
        alone[i] = [srv.submit(i, v, b).wait() for v, b in seq[i]]
    srv.shutdown()
    big = G.serving_instance("beam", 30000, 77)
    srv = SolveServer(budget_bytes=budget, nworkers=4)
    for i in range(len(inst)):
        srv.register(i, inst[i][0], inst[i][1], inst[i][2])
    srv.register(99, big[0], big[1], big[2])
    rs = []
    for s_ in range(steps):
        for i in range(len(inst)):
            rs.append((i, s_, srv.submit(i, *seq[i][s_])))
        if s_ == 1:
            rbig = srv.submit(99, big[2], big[3])
    worst = 0.0
    for i, s_, r in rs:
        r.wait()
        assert r.error is None, r.error
        ro, ci = inst[i][0], inst[i][1]
        va, b = seq[i][s_]
        assert rel_residual(ro, ci, va, r.x, b) <= 1e-9
        a = alone[i][s_]
        assert rel_residual(ro, ci, va, a.x, b) <= 1e-9
        worst = max(worst, rel_err(r.x, a.x))
        if r.kind == a.kind == "direct" and s_ == 0:
            assert rel_err(r.x, a.x) <= 1e-12  # first step: the same factor and solve
    assert worst <= 1e-6, worst
    rbig.wait()
    assert rbig.status == -2 and "more than the whole budget" in rbig.error
    st = srv.stats()
    print("[serve] budget %.1f MB, releases %d, evictions %d, admissions %d, peak %.1f MB, worst diff %.2e" % (
        budget / 2 ** 20, st["releases"], st["evictions"], st["admissions"], st["bytes_peak"] / 2 ** 20, worst),
          flush=True)
    assert st["releases"] > 0  # factors of idle instances released under the budget
    assert st["bytes_peak"] <= budget
    assert st["completed"] == len(rs) + 1 and st["errors"] == 1
    # one more step of every instance: the ones whose factor was released come
    # back through a refactor (the budget holds about four factors, so at least
    # the first of them), with their results to tolerance
    more = [(i, srv.submit(i, *seq[i][steps - 1])) for i in range(len(inst))]
    for i, r in more:
        r.wait()
        assert r.error is None, r.error
        assert rel_residual(inst[i][0], inst[i][1], seq[i][steps - 1][0], r.x, seq[i][steps - 1][1]) <= 1e-9
    st2 = srv.stats()
    assert st2["direct"] - st["direct"] >= len(inst) - 5, (st["direct"], st2["direct"])
    assert sum(1 for i, r in more if r.kind == "direct") >= len(inst) - 5
    assert st2["bytes_peak"] <= budget
    srv.shutdown()


@pytest.mark.parametrize("prec", [KKT_PREC_FP64, KKT_PREC_FP32_REFINE])
def test_multi_rhs_solve_equals_single_solves(gpu, prec):
    # KKTSolver block solves (KKTSolver.java:1537-1588) through kkt_solve_multi:
    # the right-hand sides go through the batched triangular solve eight at a time
    # (partial last group included), with per-column refinement; every column must
    # equal its single-RHS solve bitwise (same launches, same accumulation order)
    _say("multi-RHS vs single prec=%d" % prec)
    import ctypes
    import torch
    from artisynth_core_amd._lib import ptr
    beam = G.HexBeam(24, 8, 8, 1.0, 0.2, material=("neohookean", 3.4e4, 1.7e6), h=0.005)
    u = np.random.default_rng(31).uniform(-1e-4, 1e-4, (beam.nn, 3))
    vals, b = beam.system(u, np.zeros_like(u), beam.gravity_loads())
    rng = np.random.default_rng(32)
    n = beam.n
    s = MI355XSolver(precision=prec)
    try:
        if prec == KKT_PREC_FP64:
            s.setIterativeRefinementSteps(1)
        s.analyzeArrays(vals, beam.colIdxs, beam.rowOffs, n, Matrix.SYMMETRIC)
        s.factor(vals)
        for nrhs in (1, 3, 8, 9, 17, 70):
            B = rng.uniform(-1, 1, n * nrhs)
            B[:n] = b
            X = np.zeros(n * nrhs)
            s.solve(X, B, nrhs)
            for j in range(nrhs):
                xs = np.zeros(n)
                s.solve(xs, B[j * n:(j + 1) * n])
                assert np.array_equal(X[j * n:(j + 1) * n], xs), (nrhs, j)
                assert rel_residual(beam.rowOffs, beam.colIdxs, vals, xs, B[j * n:(j + 1) * n]) <= 1e-8
        # device arrays
        nrhs = 6
        B = rng.uniform(-1, 1, n * nrhs)
        dB = torch.tensor(B, device="cuda")
        dX = torch.zeros(n * nrhs, dtype=torch.float64, device="cuda")
        rc = s._L.kkt_solve_multi_device(s._h, nrhs, ctypes.c_void_p(dB.data_ptr()), ctypes.c_void_p(dX.data_ptr()))
        assert rc == 0
        X = dX.cpu().numpy()
        for j in range(nrhs):
            xs = np.zeros(n)
            s.solve(xs, B[j * n:(j + 1) * n])
            assert np.array_equal(X[j * n:(j + 1) * n], xs), j
        # overlapping input and output are rejected
        assert s._L.kkt_solve_multi_device(s._h, nrhs, ctypes.c_void_p(dB.data_ptr()),
                                           ctypes.c_void_p(dB.data_ptr() + 8 * n)) < 0
    finally:
        s.dispose()


@pytest.mark.parametrize("graphs", [1, 0])
def test_iterative_solve_iteration_limit(gpu, graphs):
    # kkt_iterative_solve with maxIter at, and one below, the iterations the solve
    # needs: converged with that count, then not converged (-13); the queued
    # half-iterations past the limit or past convergence must not change x
    _say("iterative iteration limit graphs=%d" % graphs)
    import ctypes
    from artisynth_core_amd._lib import ptr
    beam, vals, b = G.small_beam(26)
    s = MI355XSolver()
    try:
        s._L.kkt_set_graphs(s._h, graphs)
        s.analyzeArrays(vals, beam.colIdxs, beam.rowOffs, beam.n, Matrix.SYMMETRIC)
        s.factor(vals)
        # entry-wise perturbation (a uniform scale would leave the factor exact)
        v2 = vals * np.random.default_rng(21).uniform(0.9, 1.1, vals.shape)
        vv, bb = np.ascontiguousarray(v2), np.ascontiguousarray(b)

        def run(max_iter):
            x = np.zeros(beam.n)
            rc = s._L.kkt_iterative_solve(s._h, ptr(vv, ctypes.c_double), ptr(bb, ctypes.c_double),
                                          ptr(x, ctypes.c_double), 1e-10, max_iter)
            return rc, x

        need, x_need = run(32)
        assert need >= 2
        assert rel_residual(beam.rowOffs, beam.colIdxs, v2, x_need, b) <= 1e-10
        rc, x = run(need)
        assert rc == need
        assert np.array_equal(x, x_need)
        rc, _ = run(need - 1)
        assert rc == -13
        # the handle still solves afterwards
        rc, x = run(32)
        assert rc == need and np.array_equal(x, x_need)
    finally:
        s.dispose()


@pytest.mark.parametrize("prec", [KKT_PREC_FP64, KKT_PREC_FP32_REFINE])
def test_graph_replays_match_eager_launches(gpu, prec):
    # the factor, solve and BiCGStab launch sequences are captured once into HIP
    # graphs and replayed; every replay (with eager work of another handle in
    # between) must give the bitwise result of the eagerly launched kernels
    _say("graph replays vs eager prec=%d" % prec)
    beam = G.HexBeam(24, 8, 8, 1.0, 0.2, material=("neohookean", 3.4e4, 1.7e6), h=0.005)
    u = np.random.default_rng(9).uniform(-1e-4, 1e-4, (beam.nn, 3))
    vals, b = beam.system(u, np.zeros_like(u), beam.gravity_loads())
    rng = np.random.default_rng(10)
    rhs = [b] + [rng.uniform(-1, 1, beam.n) for _ in range(3)]
    solvers = []
    for graphs in (0, 1):
        s = MI355XSolver(precision=prec)
        s._L.kkt_set_graphs(s._h, graphs)
        s.setIterativeRefinementSteps(0)
        s.setRefinement(1e-10, 10)
        s.analyzeArrays(vals, beam.colIdxs, beam.rowOffs, beam.n, Matrix.SYMMETRIC)
        solvers.append(s)
    try:
        for rep in range(3):
            v = vals * (1.0 + 0.01 * rep)
            xs = []
            for s in solvers:
                s.factor(v)
                out = []
                for r in rhs:
                    x = np.zeros(beam.n)
                    s.solve(x, r)
                    out.append(x)
                xs.append(out)
            for xe, xg in zip(*xs):
                assert np.array_equal(xe, xg)
            assert rel_residual(beam.rowOffs, beam.colIdxs, v, xs[1][0], b) <= 1e-8
            # stale-factor BiCGStab: captured half-iterations vs eager launches
            v2 = v * (1.0 + 0.002 * (rep + 1))
            its, xi = [], []
            for s in solvers:
                x = np.zeros(beam.n)
                its.append(s.iterativeSolve(v2, x, b, 10))
                xi.append(x)
            assert its[0] == its[1] > 0
            assert np.array_equal(xi[0], xi[1])
            assert rel_residual(beam.rowOffs, beam.colIdxs, v2, xi[1], b) <= 1e-9
    finally:
        for s in solvers:
            s.dispose()


def _csr_slots(ro, ci, rows, cols):
    """Positions of (row, col) entries in an upper CSR (-1 below the diagonal)."""
    n = len(ro) - 1
    kr = np.repeat(np.arange(n), np.diff(ro)).astype(np.int64)
    keys = kr * n + np.asarray(ci, np.int64)
    rows = np.asarray(rows, np.int64)
    cols = np.asarray(cols, np.int64)
    q = rows * n + cols
    pos = np.searchsorted(keys, q)
    pos = np.minimum(pos, len(keys) - 1)
    ok = (keys[pos] == q) & (rows <= cols)
    return np.where(ok, pos, -1).astype(np.int32)


def test_device_value_assembly_ops(gpu):
    # CuDssSolver.*DeviceValues (CuDssSolver.java:421-788) against the oracle's
    # restatement: duplicate and negative slots, 3-DOF diagonals, scaled 3x3 blocks,
    # material stiffness blocks, and multiply with the assembled buffer
    _say("device value assembly ops")
    beam, vals, b = G.small_beam(8)
    rng = np.random.default_rng(21)
    s = MI355XSolver()
    try:
        s.analyzeArrays(vals, beam.colIdxs, beam.rowOffs, beam.n, Matrix.SYMMETRIC)
        nnz = beam.nnz
        s.clearDeviceValues()
        got = s.getDeviceValues(np.ones(nnz))
        assert not got.any()
        ref = np.zeros(nnz)
        slots = rng.integers(-3, nnz, 5000).astype(np.int32)
        contrib = rng.standard_normal(5000)
        s.addDeviceValues(slots, contrib, 5000, 0.5)
        O.scatter_add(ref, slots, contrib, 0.5)
        s.addDeviceValues(slots, contrib, 4000, -2.0)  # prefix: a second plan
        O.scatter_add(ref, slots[:4000], contrib[:4000], -2.0)
        ds = rng.integers(0, nnz, 3 * 300).astype(np.int32)
        ds[::7] = -1
        ms = rng.uniform(1, 2, 300)
        s.addScaledDiagonal3DeviceValues(ds, ms, 300, 1.5)
        O.scatter_add(ref, ds, np.repeat(ms, 3), 1.5)
        bsl = rng.integers(-1, nnz, 9 * 200).astype(np.int32)
        bv = rng.standard_normal(9 * 200)
        bs = rng.uniform(0.5, 2, 200)
        s.addScaledBlock3DeviceValues(bsl, bv, bs, 200, 0.25)
        O.scatter_add(ref, bsl, bv * np.repeat(bs, 9), 0.25)
        nb = 150
        msl = rng.integers(-1, nnz, 9 * nb).astype(np.int32)
        gis, gjs = rng.standard_normal((nb, 3)), rng.standard_normal((nb, 3))
        Ds = rng.standard_normal((nb, 6, 6))
        sig, dvs = rng.standard_normal((nb, 6)), rng.uniform(0.1, 1, nb)
        s.addMaterialStiffness3DeviceValues(msl, gis, gjs, Ds, sig, dvs, nb, 0.75)
        K = np.stack([O.material_stiffness3(gis[q], gjs[q], Ds[q], sig[q], dvs[q]) for q in range(nb)])
        O.scatter_add(ref, msl, K.reshape(-1), 0.75)
        got = s.getDeviceValues(np.zeros(nnz))
        assert np.allclose(got, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
        # the same calls again reuse the cached plans and add the same amounts
        s.addDeviceValues(slots, contrib, 5000, 0.5)
        O.scatter_add(ref, slots, contrib, 0.5)
        got = s.getDeviceValues(np.zeros(nnz))
        assert np.allclose(got, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
        # multiply uses the device buffer
        x = rng.standard_normal(beam.n)
        y = s.multiply(x, np.zeros(beam.n))
        assert np.allclose(y, full_from_upper(beam.rowOffs, beam.colIdxs, ref) @ x, rtol=1e-12,
                           atol=1e-12 * np.abs(y).max())
        with pytest.raises(NumericalException):
            s.addDeviceValues(np.array([nnz], np.int32), np.ones(1), 1, 1.0)
    finally:
        s.dispose()


@pytest.mark.parametrize("prec", [KKT_PREC_FP64, KKT_PREC_FP32_REFINE])
def test_device_assembled_tet_kkt_solves(gpu, prec):
    # the KKT of a tet body assembled on the device (lumped masses, material
    # stiffness blocks from shape-function gradients, constraint entries), then
    # factored from the device buffer: equals the host-assembled system
    _say("device-assembled tet KKT prec=%d" % prec)
    body, V, B = G.tet_precision_sequence(nsteps=1, dims=(21, 6, 6))
    u0 = np.zeros((body.nn, 3))
    fext = body.gravity_loads() + np.random.default_rng(22).uniform(-1e-2, 1e-2, (body.nn, 3))
    vals, b = body.system(u0, u0, fext)  # undeformed: linear isotropic tangent
    n, ro, ci = body.n, body.rowOffs, body.colIdxs
    c10, kappa = body.material[1], body.material[2]
    mu = 2 * c10
    D = O.isotropic_D(kappa - 2 * mu / 3, mu)
    E = len(body.elems)
    g = body.grads
    a_idx = np.repeat(np.arange(4), 4)
    b_idx = np.tile(np.arange(4), 4)
    na = body.elems[:, a_idx]  # (E, 16)
    nb_ = body.elems[:, b_idx]
    rows = (3 * na[:, :, None, None] + np.arange(3)[None, None, :, None]) + 0 * np.arange(3)[None, None, None, :]
    cols = (3 * nb_[:, :, None, None] + np.arange(3)[None, None, None, :]) + 0 * np.arange(3)[None, None, :, None]
    bslots = _csr_slots(ro, ci, rows.ravel(), cols.ravel())
    gis = g[:, a_idx].reshape(-1, 3)
    gjs = g[:, b_idx].reshape(-1, 3)
    nblk = E * 16
    Ds = np.broadcast_to(D, (nblk, 6, 6))
    sig = np.zeros((nblk, 6))
    dvs = np.repeat(body.vol, 16)
    nodes = np.arange(body.nn)
    dslots = _csr_slots(ro, ci, (3 * nodes[:, None] + np.arange(3)).ravel(),
                        (3 * nodes[:, None] + np.arange(3)).ravel())
    s = MI355XSolver(precision=prec)
    try:
        s.analyzeArrays(vals, ci, ro, n, Matrix.SYMMETRIC)
        s.clearDeviceValues()
        s.addScaledDiagonal3DeviceValues(dslots, body.mass[0::3], body.nn, 1.0)
        s.addMaterialStiffness3DeviceValues(bslots, gis, gjs, Ds, sig, dvs, nblk,
                                            body.h * body.beta + body.h * body.h)
        s.addDeviceValues(body._gpos.astype(np.int32), np.ones(len(body._gpos)), len(body._gpos), 1.0)
        got = s.getDeviceValues(np.zeros(body.nnz))
        assert np.allclose(got, vals, rtol=1e-12, atol=1e-12 * np.abs(vals).max())
        s.factorDeviceValues()
        x = np.zeros(n)
        s.solve(x, b)
        assert rel_residual(ro, ci, vals, x, b) <= 1e-8
        inf = s.info()
        assert (inf["npos"], inf["nneg"]) == (body.sizeM, body.numG)
    finally:
        s.dispose()


@pytest.mark.parametrize("prec", [KKT_PREC_FP64, KKT_PREC_FP32_REFINE])
def test_solve_shape_autotuner(gpu, prec):
    # kkt_autotune_solve re-plans the solve kernels' lane shapes per tree level:
    # every shape is a candidate, solutions move only at rounding level, a
    # re-captured graph repeats bitwise, and refinement still reaches its tolerance
    _say("solve-shape autotuner prec=%d" % prec)
    beam = G.HexBeam(24, 8, 8, 1.0, 0.2, material=("neohookean", 3.4e4, 1.7e6), h=0.005)
    u = np.random.default_rng(11).uniform(-1e-4, 1e-4, (beam.nn, 3))
    vals, b = beam.system(u, np.zeros_like(u), beam.gravity_loads())
    solvers = []
    for graphs in (0, 1):
        s = MI355XSolver(precision=prec)
        s._L.kkt_set_graphs(s._h, graphs)
        s.setIterativeRefinementSteps(0)
        s.setRefinement(1e-10, 10)
        s.analyzeArrays(vals, beam.colIdxs, beam.rowOffs, beam.n, Matrix.SYMMETRIC)
        solvers.append(s)
    try:
        x0 = []
        for s in solvers:
            s.factor(vals)
            f, bw = s.getSolveShapes()
            assert len(f) == s.info()["nlevels"] and set(f) == {16} and set(bw) == {16}
            x = np.zeros(beam.n)
            s.solve(x, b)
            x0.append(x)
        shapes = [s.autotuneSolve(3) for s in solvers]
        for f, bw in shapes:
            assert set(f) <= {4, 8, 16, 32} and set(bw) <= {4, 8, 16, 32}
        # per handle (eager, graphs): the first solve after tuning runs the new
        # items (the graph handle captures again), the second repeats it bitwise
        xs = []
        for s, xa in zip(solvers, x0):
            x = np.zeros(beam.n)
            s.solve(x, b)
            assert rel_residual(beam.rowOffs, beam.colIdxs, vals, x, b) <= 1e-10
            assert np.abs(x - xa).max() <= 1e-9 * np.abs(xa).max()
            x2 = np.zeros(beam.n)
            s.solve(x2, b)
            assert np.array_equal(x, x2)
            xs.append(x)
    finally:
        for s in solvers:
            s.dispose()
