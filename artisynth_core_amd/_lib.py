"""ctypes binding of libkkt_mi355x.so (the C ABI declared in include/kkt_mi355x.h).

The library is built in-tree by artisynth_core_amd.build (or __graft_entry__.build()).
There is no fallback: if the library is missing, loading raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkkt_mi355x.so")

KKT_MT_GENERAL, KKT_MT_SYMMETRIC, KKT_MT_SPD = 0, 1, 2
KKT_PREC_FP64, KKT_PREC_FP32_REFINE = 0, 1

ERRORS = {
    0: "OK", -1: "INIT", -2: "ALLOC", -3: "COPY", -4: "PATTERN", -5: "ANALYSIS",
    -6: "FACTORIZATION", -7: "SOLVE", -8: "STATE", -9: "UNKNOWN_MTYPE", -10: "SPMV",
    -11: "REDUCTION", -12: "ITER_BREAKDOWN", -13: "ITER_NO_CONVERGE", -14: "ARGUMENT",
    -15: "DEVICE", -16: "REFINE_STALL",
}


class KktInfo(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("nsuper", ctypes.c_int), ("nlevels", ctypes.c_int),
        ("ndelayed", ctypes.c_int), ("nnz", ctypes.c_longlong), ("nnzL", ctypes.c_longlong),
        ("nnzL_exact", ctypes.c_longlong), ("factor_flops", ctypes.c_double),
        ("front_bytes", ctypes.c_longlong), ("npos", ctypes.c_int), ("nneg", ctypes.c_int),
        ("nperturbed", ctypes.c_int), ("n2x2", ctypes.c_int), ("precision", ctypes.c_int),
        ("refine_steps", ctypes.c_int), ("fp64_fallbacks", ctypes.c_int),
        ("last_rel_residual", ctypes.c_double), ("pivoted_blocks", ctypes.c_int),
        ("device_bytes", ctypes.c_longlong), ("factored_bytes", ctypes.c_longlong),
        ("fp64_fallback_bytes", ctypes.c_longlong), ("precond_half", ctypes.c_int),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class KktBatchSysInfo(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("npos", ctypes.c_int), ("nneg", ctypes.c_int),
        ("nperturbed", ctypes.c_int), ("n2x2", ctypes.c_int), ("refine_steps", ctypes.c_int),
        ("last_rel_residual", ctypes.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class KktServeConfig(ctypes.Structure):
    _fields_ = [("budget_bytes", ctypes.c_longlong), ("nworkers", ctypes.c_int), ("precision", ctypes.c_int),
                ("tol_exp", ctypes.c_int), ("max_iter", ctypes.c_int), ("refine_tol", ctypes.c_double),
                ("group_runners", ctypes.c_int), ("group_max", ctypes.c_int)]


class KktServeResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("kind", ctypes.c_int), ("iterations", ctypes.c_int),
                ("rel_residual", ctypes.c_double), ("submit_ms", ctypes.c_double), ("start_ms", ctypes.c_double),
                ("done_ms", ctypes.c_double), ("message", ctypes.c_char * 160)]


class KktServeStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_longlong), ("completed", ctypes.c_longlong), ("errors", ctypes.c_longlong),
                ("direct", ctypes.c_longlong), ("iterative", ctypes.c_longlong),
                ("admissions", ctypes.c_longlong), ("evictions", ctypes.c_longlong),
                ("queue_depth", ctypes.c_longlong), ("queue_max", ctypes.c_longlong),
                ("budget_bytes", ctypes.c_longlong), ("bytes_resident", ctypes.c_longlong),
                ("bytes_peak", ctypes.c_longlong), ("instances_resident", ctypes.c_int),
                ("service_ms", ctypes.c_double), ("elapsed_ms", ctypes.c_double),
                ("group_calls", ctypes.c_longlong), ("group_requests", ctypes.c_longlong),
                ("releases", ctypes.c_longlong)]


# (name, restype, argtypes) for every symbol of include/kkt_mi355x.h
_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_IP = ctypes.POINTER(ctypes.c_int)
_DP = ctypes.POINTER(ctypes.c_double)
_LLP = ctypes.POINTER(ctypes.c_longlong)
_UCP = ctypes.POINTER(ctypes.c_ubyte)
_PP = ctypes.POINTER(ctypes.c_void_p)  # arrays of handles / device pointers
SIGNATURES = [
    ("kkt_create", _P, []),
    ("kkt_destroy", None, [_P]),
    ("kkt_set_pattern", _I, [_P, _I, _I, _IP, _IP, _I]),
    ("kkt_analyze", _I, [_P]),
    ("kkt_analyze_values", _I, [_P, _DP]),
    ("kkt_factor", _I, [_P, _DP]),
    ("kkt_factor_device", _I, [_P, _P]),
    ("kkt_clear_device_values", _I, [_P]),
    ("kkt_get_device_values", _I, [_P, _DP]),
    ("kkt_add_device_values", _I, [_P, _IP, _DP, _I, _D]),
    ("kkt_add_scaled_diagonal3_device_values", _I, [_P, _IP, _DP, _I, _D]),
    ("kkt_add_scaled_block3_device_values", _I, [_P, _IP, _DP, _DP, _I, _D]),
    ("kkt_add_material_stiffness3_device_values", _I, [_P, _IP, _DP, _DP, _DP, _DP, _DP, _I, _D]),
    ("kkt_add_material_stiffness3_element_device_values", _I,
     [_P, _IP, _IP, _IP, _IP, _IP, _IP, _IP, _IP, _DP, _DP, _DP, _DP, _DP, _I, _D]),
    ("kkt_add_linear_elastic_stiffness3_element_device_values", _I,
     [_P, _IP, _IP, _IP, _IP, _IP, _IP, _DP, _DP, _DP, _I, _D]),
    ("kkt_add_linear_elastic_stiffness3_element_geometry_device_values", _I,
     [_P, _IP, _IP, _IP, _IP, _IP, _IP, _IP, _IP, _DP, _DP, _DP, _DP, _DP, _I, _D]),
    ("kkt_add_dilational_stiffness3_element_device_values", _I,
     [_P, _IP, _IP, _IP, _IP, _IP, _IP, _IP, _IP, _IP, _DP, _DP, _DP, _I, _D]),
    ("kkt_factor_device_values", _I, [_P]),
    ("kkt_autotune_solve", _I, [_P, _I]),
    ("kkt_get_solve_shapes", _I, [_P, _IP, _IP]),
    ("kkt_get_forward_forms", _I, [_P, _IP]),
    ("kkt_set_forward_forms", _I, [_P, _IP]),
    ("kkt_get_forward_row_groups", _I, [_P, _IP]),
    ("kkt_set_forward_row_groups", _I, [_P, _IP]),
    ("kkt_solve", _I, [_P, _DP, _DP]),
    ("kkt_solve_device", _I, [_P, _P, _P]),
    ("kkt_solve_device_noalloc", _I, [_P, _P, _P]),
    ("kkt_solve_multi", _I, [_P, _I, _DP, _DP]),
    ("kkt_solve_multi_device", _I, [_P, _I, _P, _P]),
    ("kkt_multiply", _I, [_P, _DP, _DP]),
    ("kkt_iterative_solve", _I, [_P, _DP, _DP, _DP, _D, _I]),
    ("kkt_iterative_solve_device", _I, [_P, _P, _P, _P, _D, _I]),
    ("kkt_release_factor", _I, [_P]),
    ("kkt_group_create", _P, []),
    ("kkt_group_destroy", None, [_P]),
    ("kkt_group_iterative_solve_device", _I, [_P, _I, _PP, _PP, _PP, _PP, _D, _I, _IP]),
    ("kkt_group_iterative_solve", _I, [_P, _I, _PP, _PP, _PP, _PP, _D, _I, _IP]),
    ("kkt_group_last_error", ctypes.c_char_p, [_P]),
    ("kkt_set_ir_steps", _I, [_P, _I]),
    ("kkt_last_error", ctypes.c_char_p, [_P]),
    ("kkt_version", ctypes.c_char_p, []),
    ("kkt_set_timing_enabled", None, [_I]),
    ("kkt_set_precision", _I, [_P, _I]),
    ("kkt_set_refine_tol", _I, [_P, _D, _I]),
    ("kkt_set_graphs", _I, [_P, _I]),
    ("kkt_set_ordering", _I, [_P, _I]),
    ("kkt_set_max_supernode", _I, [_P, _I]),
    ("kkt_set_iterative_half", _I, [_P, _I]),
    ("kkt_get_info", _I, [_P, ctypes.POINTER(KktInfo)]),
    ("kkt_get_symbolic", _I, [_P, _IP, _IP, _IP, _LLP, _IP]),
    ("kkt_verify_device_structure", _I, [_P]),
    ("kkt_timing_reset", None, [_P, _I]),
    ("kkt_timing_get", _I, [_P, _DP, _LLP, _DP]),
    ("kkt_time_kernels", _I, [_P, _I, _I, _DP, _IP]),
    ("kkt_batch_create", _P, []),
    ("kkt_batch_destroy", None, [_P]),
    ("kkt_batch_set_systems", _I, [_P, _I, _IP, _IP, _IP, _I]),
    ("kkt_batch_analyze", _I, [_P, _DP]),
    ("kkt_batch_factor", _I, [_P, _DP]),
    ("kkt_batch_factor_device", _I, [_P, _P]),
    ("kkt_batch_solve", _I, [_P, _DP, _DP]),
    ("kkt_batch_solve_device", _I, [_P, _P, _P]),
    ("kkt_batch_set_precision", _I, [_P, _I]),
    ("kkt_batch_set_refine_tol", _I, [_P, _D, _I]),
    ("kkt_batch_set_ir_steps", _I, [_P, _I]),
    ("kkt_batch_autotune_solve", _I, [_P, _I]),
    ("kkt_batch_get_info", _I, [_P, ctypes.POINTER(KktInfo)]),
    ("kkt_batch_system_info", _I, [_P, _I, ctypes.POINTER(KktBatchSysInfo)]),
    ("kkt_batch_last_error", ctypes.c_char_p, [_P]),
    ("kkt_serve_default_config", None, [ctypes.POINTER(KktServeConfig)]),
    ("kkt_serve_create", _P, [ctypes.POINTER(KktServeConfig)]),
    ("kkt_serve_destroy", None, [_P]),
    ("kkt_serve_register", _I, [_P, _I, _I, _I, _IP, _IP, _DP, _I]),
    ("kkt_serve_submit", ctypes.c_longlong, [_P, _I, _DP, _DP, _DP]),
    ("kkt_serve_submit_device", ctypes.c_longlong, [_P, _I, _P, _P, _P]),
    ("kkt_serve_wait", _I, [_P, ctypes.c_longlong, ctypes.POINTER(KktServeResult)]),
    ("kkt_serve_get_stats", _I, [_P, ctypes.POINTER(KktServeStats)]),
    ("kkt_serve_last_error", ctypes.c_char_p, [_P]),
    ("kkt_symbolic_create", _P, [_I, _I, _IP, _IP, _UCP, _I]),
    ("kkt_symbolic_info", _I, [_P, ctypes.POINTER(KktInfo)]),
    ("kkt_symbolic_get", _I, [_P, _IP, _IP, _IP, _IP, _IP, _LLP, _IP, _IP]),
    ("kkt_symbolic_destroy", None, [_P]),
]

_lib = None


def lib():
    """Loads the native library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libkkt_mi355x.so is not built (%s); run __graft_entry__.build() or "
                "python -m artisynth_core_amd.build" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a, ctype):
    """Pointer to the data of a C-contiguous numpy array."""
    return a.ctypes.data_as(ctypes.POINTER(ctype))
