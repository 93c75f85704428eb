"""Host-only symbolic analysis through the C ABI (kkt_symbolic_*), for CPU tests."""
import ctypes

import numpy as np

from artisynth_core_amd._lib import KktInfo, lib, ptr


def zero_diag_mask(rowOffs, colIdxs, vals):
    n = len(rowOffs) - 1
    mask = np.ones(n, np.uint8)
    for i in range(n):
        for p in range(rowOffs[i], rowOffs[i + 1]):
            if colIdxs[p] == i and vals[p] != 0.0:
                mask[i] = 0
    return mask


def product_symbolic(rowOffs, colIdxs, vals=None, ordering=1):
    L = lib()
    ro = np.ascontiguousarray(rowOffs, np.int32)
    ci = np.ascontiguousarray(colIdxs, np.int32)
    n = len(ro) - 1
    mask = zero_diag_mask(ro, ci, vals) if vals is not None else None
    h = L.kkt_symbolic_create(n, int(ro[n]), ptr(ro, ctypes.c_int), ptr(ci, ctypes.c_int),
                              ptr(mask, ctypes.c_ubyte) if mask is not None else None, ordering)
    assert h, "symbolic analysis failed"
    try:
        inf = KktInfo()
        L.kkt_symbolic_info(h, ctypes.byref(inf))
        ns = inf.nsuper
        out = {"info": inf.as_dict(), "perm": np.zeros(n, np.int32), "parent": np.zeros(n, np.int32),
               "colcount": np.zeros(n, np.int32), "sup_ptr": np.zeros(ns + 1, np.int32),
               "sparent": np.zeros(ns, np.int32), "rows_ptr": np.zeros(ns + 1, np.int64),
               "level": np.zeros(ns, np.int32)}
        L.kkt_symbolic_get(h, ptr(out["perm"], ctypes.c_int), ptr(out["parent"], ctypes.c_int),
                           ptr(out["colcount"], ctypes.c_int), ptr(out["sup_ptr"], ctypes.c_int),
                           ptr(out["sparent"], ctypes.c_int), ptr(out["rows_ptr"], ctypes.c_longlong),
                           None, ptr(out["level"], ctypes.c_int))
        rows = np.zeros(max(int(out["rows_ptr"][-1]), 1), np.int32)
        L.kkt_symbolic_get(h, None, None, None, None, None, None, ptr(rows, ctypes.c_int), None)
        out["rows"] = rows[:int(out["rows_ptr"][-1])]
        out["mask"] = mask
        return out
    finally:
        L.kkt_symbolic_destroy(h)
