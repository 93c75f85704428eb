"""Build recipe of the native library (libkkt_mi355x.so) for gfx950.

Compiles the HIP kernels, the host symbolic analysis and the C ABI with hipcc
into artisynth_core_amd/libkkt_mi355x.so.  Run `python -m artisynth_core_amd.build`
or call build().
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libkkt_mi355x.so")
# the HBM streaming probe of bench.py --mode stream (measurement only)
PROBE_LIB = os.path.join(HERE, "libkkt_probe.so")

SOURCES = ["kernels.hip", "solver.hip", "symbolic.cpp", "serve.cpp"]
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c


def build(verbose=False, jobs=3):
    os.makedirs(OBJ, exist_ok=True)
    hipcc = _hipcc()
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [hipcc] + FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    for p in procs:
        _wait(p)
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    _wait(("link", subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    cmd = [hipcc] + FLAGS + ["-x", "hip", "-shared", os.path.join(CSRC, "probe.hip"), "-o", PROBE_LIB]
    if verbose:
        print(" ".join(cmd), flush=True)
    _wait(("probe.hip", subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    return LIB


def _wait(item):
    src, p = item
    out, _ = p.communicate()
    txt = out.decode(errors="replace")
    if p.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, txt))
    lines = [l for l in txt.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if lines:
        sys.stderr.write("\n".join(lines[-30:]) + "\n")


if __name__ == "__main__":
    print(build(verbose=True))
