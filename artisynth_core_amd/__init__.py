"""MI355X-native sparse KKT direct solver backend for ArtiSynth's DirectSolver contract.

Native library: libkkt_mi355x.so (C ABI in include/kkt_mi355x.h), built in-tree by
artisynth_core_amd.build.  Python layer: `solvers` (DirectSolver / KKTSolver mirror),
`generators` (synthetic KKT systems of the benchmark configurations).
"""
from ._lib import KKT_PREC_FP32_REFINE, KKT_PREC_FP64, lib  # noqa: F401

__all__ = ["lib", "KKT_PREC_FP64", "KKT_PREC_FP32_REFINE"]
