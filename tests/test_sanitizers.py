"""CPU test: the host symbolic analysis (artisynth_core_amd/csrc/symbolic.cpp)
and the C oracle (oracle/kkt_oracle.c) built with AddressSanitizer and
UndefinedBehaviorSanitizer and run on the test matrices (tests/sanitize/
san_driver.cpp cross-checks the two; a sanitizer report fails the run), with
the driver's orderings and supernode partitions compared against the product
library's own host analysis.  The device plan builder uploads while it builds,
so it runs only with a GPU and is not covered here."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from artisynth_core_amd import generators as G
from kkt_helpers import P5_COLS, P5_ROWS, P5_VALS
from symbolic_helpers import product_symbolic, zero_diag_mask

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if not shutil.which("gcc") or not shutil.which("g++"):
        pytest.skip("gcc/g++ not available")
    d = tmp_path_factory.mktemp("san")
    obj = str(d / "kkt_oracle.o")
    exe = str(d / "san_driver")
    inc = ["-I" + os.path.join(ROOT, "oracle"), "-I" + os.path.join(ROOT, "artisynth_core_amd", "csrc")]
    subprocess.run(["gcc", "-std=c11"] + SAN + inc + ["-c", os.path.join(ROOT, "oracle", "kkt_oracle.c"), "-o", obj],
                   check=True, capture_output=True)
    subprocess.run(["g++", "-std=c++17"] + SAN + inc +
                   [os.path.join(ROOT, "tests", "sanitize", "san_driver.cpp"),
                    os.path.join(ROOT, "artisynth_core_amd", "csrc", "symbolic.cpp"), obj, "-o", exe],
                   check=True, capture_output=True)
    # the driver carries the sanitizer runtimes (an uninstrumented build would
    # pass every check below without testing anything)
    libs = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
    assert "libasan" in libs and "libubsan" in libs, libs
    return exe, d


def _cases():
    beam, vals, b = G.small_beam(8)
    yield "beam8", beam.rowOffs, beam.colIdxs, vals, b
    beam, vals, b = G.small_beam(26)
    yield "beam26", beam.rowOffs, beam.colIdxs, vals, b
    d = np.load(os.path.join(ROOT, "tests", "golden", "testMatrix_upper.npz"))
    n = len(d["rowOffs"]) - 1
    yield "testMatrix", d["rowOffs"], d["colIdxs"], d["vals"], np.linspace(1.0, 2.0, n)
    yield "pardiso5", P5_ROWS, P5_COLS, P5_VALS[0][0], np.arange(1.0, 6.0)
    ro, ci, va, b = G.serving_instance("rigid", 600, 11)
    yield "rigid_contact", ro, ci, va, b
    systems = G.batch_mix(nsys=3, seed=5, dof_lo=200, dof_hi=700)
    ro, ci, va = G.block_diag_csr(systems)[:3]
    yield "batch3", ro, ci, va, np.cos(np.arange(len(ro) - 1))


def _write(path, ro, ci, va, mask, b):
    ro = np.ascontiguousarray(ro, np.int32)
    n = len(ro) - 1
    with open(path, "wb") as f:
        f.write(np.array([n, int(ro[n])], np.int32).tobytes())
        f.write(ro.tobytes())
        f.write(np.ascontiguousarray(ci, np.int32).tobytes())
        f.write(np.ascontiguousarray(va, np.float64).tobytes())
        f.write(np.ascontiguousarray(mask, np.uint8).tobytes())
        f.write(np.ascontiguousarray(b, np.float64).tobytes())


def _read(path, n):
    raw = open(path, "rb").read()
    out, p = [], 0
    for _ in range(2):
        ns = int(np.frombuffer(raw, np.int32, 1, p)[0])
        p += 4
        perm = np.frombuffer(raw, np.int32, n, p)
        p += 4 * n
        sup = np.frombuffer(raw, np.int32, ns + 1, p)
        p += 4 * (ns + 1)
        nnzl = int(np.frombuffer(raw, np.int64, 1, p)[0])
        p += 8
        res = float(np.frombuffer(raw, np.float64, 1, p)[0])
        p += 8
        out.append((perm, sup, nnzl, res))
    return out


def _run(cmd):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    return subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=300)


def test_malformed_patterns_under_sanitizers(driver):
    exe, _ = driver
    r = _run([exe, "--bad"])
    assert r.returncode == 0, r.stderr[-3000:]


@pytest.mark.parametrize("case", list(_cases()), ids=lambda c: c[0])
def test_symbolic_and_oracle_under_sanitizers(driver, case):
    exe, d = driver
    name, ro, ci, va, b = case
    n = len(ro) - 1
    mask = zero_diag_mask(ro, ci, va)
    inp, outp = str(d / (name + ".in")), str(d / (name + ".out"))
    _write(inp, ro, ci, va, mask, b)
    r = _run([exe, inp, outp])
    assert r.returncode == 0, r.stderr[-3000:]
    for ordering, (perm, sup, nnzl, res) in enumerate(_read(outp, n)):
        S = product_symbolic(ro, ci, va, ordering=ordering)
        assert np.array_equal(perm, S["perm"]), (name, ordering)
        assert np.array_equal(sup, S["sup_ptr"]), (name, ordering)
        assert nnzl == S["info"]["nnzL"], (name, ordering)
        assert 0.0 <= res <= 1e-10, (name, ordering, res)
