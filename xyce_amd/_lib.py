"""ctypes binding of libxg_load.so (the C ABI declared in include/xg_load.h).

The library must have been built (``python -m xyce_amd.build``); importing
this module without it raises immediately: there is no fallback path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libxg_load.so")

XG_OK = 0
XG_E_ARG, XG_E_STATE, XG_E_HIP, XG_E_PATTERN, XG_E_NOMEM, XG_E_PARAM = -1, -2, -3, -4, -5, -6
ERRORS = {
    -1: "XG_E_ARG", -2: "XG_E_STATE", -3: "XG_E_HIP", -4: "XG_E_PATTERN",
    -5: "XG_E_NOMEM", -6: "XG_E_PARAM",
}

c_int32_p = ctypes.POINTER(ctypes.c_int32)
c_int64_p = ctypes.POINTER(ctypes.c_int64)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_char_pp = ctypes.POINTER(ctypes.c_char_p)


class SolverState(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "dcopFlag", "initJctFlag", "initFixFlag", "initTranFlag", "newtonIter",
        "tranopFlag", "acopFlag", "transientFlag")]


class Stats(ctypes.Structure):
    _fields_ = [("n_limited", ctypes.c_int64), ("n_nonfinite", ctypes.c_int64)]


# (name, restype, argtypes) for every exported symbol of include/xg_load.h
SIGNATURES = [
    ("xg_last_error", ctypes.c_char_p, []),
    ("xg_circuit_new", ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    ("xg_circuit_free", None, [ctypes.c_void_p]),
    ("xg_circuit_set_options", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int32, ctypes.c_double, ctypes.c_double]),
    ("xg_add_mos1_model", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, c_char_pp, c_double_p, c_int32_p]),
    ("xg_add_mos1_instance", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, c_char_pp, c_double_p, c_int32_p, ctypes.c_int32, ctypes.c_int32, c_int32_p]),
    ("xg_add_bsim3_model", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, c_char_pp, c_double_p, c_int32_p]),
    ("xg_add_bsim3_instance", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, c_char_pp, c_double_p, c_int32_p, ctypes.c_int32, ctypes.c_int32, c_int32_p]),
    ("xg_add_mos1_instances", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, c_char_pp, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("xg_add_bsim3_instances", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, c_char_pp, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("xg_add_foreign_stamp", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, c_int32_p, ctypes.c_int32, c_int32_p, c_int32_p]),
    ("xg_circuit_set_pattern", ctypes.c_int, [ctypes.c_void_p, c_int32_p, c_int32_p]),
    ("xg_finalize", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    ("xg_finalize_plan", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    ("xg_destroy", None, [ctypes.c_void_p]),
    ("xg_pattern_size", ctypes.c_int, [ctypes.c_void_p, c_int32_p, c_int64_p]),
    ("xg_get_pattern", ctypes.c_int, [ctypes.c_void_p, c_int32_p, c_int32_p]),
    ("xg_get_mos1_offsets", ctypes.c_int, [ctypes.c_void_p, c_int32_p]),
    ("xg_get_bsim3_offsets", ctypes.c_int, [ctypes.c_void_p, c_int32_p]),
    ("xg_get_mos1_fields", ctypes.c_int, [ctypes.c_void_p, c_double_p, c_int32_p]),
    ("xg_get_bsim3_fields", ctypes.c_int, [ctypes.c_void_p, c_double_p, c_int32_p]),
    ("xg_set_solver_state", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SolverState)]),
    ("xg_update_state", ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 5),
    ("xg_update_state_type", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int32] + [ctypes.c_void_p] * 5),
    ("xg_load_vectors", ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 4),
    ("xg_load_matrices", ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 2),
    ("xg_load_all", ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 11),
    ("xg_update_state_host", ctypes.c_int, [ctypes.c_void_p] * 6),
    ("xg_load_vectors_host", ctypes.c_int, [ctypes.c_void_p] * 5),
    ("xg_load_matrices_host", ctypes.c_int, [ctypes.c_void_p] * 3),
    ("xg_load_all_noalloc", ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 11),
    ("xg_get_stats", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Stats)]),
    ("xg_sync", ctypes.c_int, [ctypes.c_void_p]),
    ("xg_set_stream", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("xg_add_resistor", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.c_double]),
    ("xg_add_capacitor", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                         ctypes.c_double, ctypes.c_int32]),
    ("xg_add_resistors", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p]),
    ("xg_add_capacitors", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p]),
    ("xg_add_vsource", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_int32, ctypes.c_double,
                                       ctypes.POINTER(ctypes.c_double)]),
    ("xg_set_time", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]),
    ("xg_rhs_norms", ctypes.c_int, [ctypes.c_void_p] * 3),
    ("xg_tia_predict", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double,
                                       ctypes.c_double] + [ctypes.c_void_p] * 4),
    ("xg_component_limited", ctypes.c_int, [ctypes.c_void_p] * 2),
    ("xg_components", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32),
                                      ctypes.c_void_p]),
    ("xg_tia_assemble", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.c_double] + [ctypes.c_void_p] * 10),
    ("xg_dense_solve", ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 5 +
     [ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]),
    ("xg_add_diode", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                    ctypes.c_int32, c_char_pp, c_double_p, ctypes.c_int32,
                                    c_char_pp, c_double_p, ctypes.c_int32]),
    ("xg_lu_solve", ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 5 +
     [ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]),
    ("xg_lu_info", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("xg_lu_front_set_big", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("xg_lu_front_info", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("xg_lu_front_export", ctypes.c_int, [ctypes.c_void_p] * 4),
    ("xg_circuit_set_lu_mode", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("xg_lu_cache_stats", ctypes.c_int, [ctypes.c_void_p]),
    ("xg_lu_cache_set_limit", ctypes.c_int, [ctypes.c_int64]),
    ("xg_lu_export", ctypes.c_int, [ctypes.c_void_p] * 4),
    ("xg_set_step_source", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("xg_set_active_mask", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("xg_upload", ctypes.c_int, [ctypes.c_void_p]),
    ("xg_tran_create", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                      ctypes.c_void_p]),
    ("xg_tran_dcop", ctypes.c_int, [ctypes.c_void_p]),
    ("xg_tran_uic", ctypes.c_int, [ctypes.c_void_p]),
    ("xg_tran_step", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    ("xg_tran_run", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int64,
                                   ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                   ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]),
    ("xg_tran_prepare_run", ctypes.c_int, [ctypes.c_void_p]),
    ("xg_tran_adaptive_begin", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                              ctypes.c_void_p, ctypes.c_int32, ctypes.c_double,
                                              ctypes.c_double]),
    ("xg_tran_adaptive_step", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("xg_tran_adaptive_info", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("xg_tran_set_eviction", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double]),
    ("xg_tran_evicted", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("xg_tran_dcop_continuation", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("xg_tran_gmin_steps", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("xg_tran_probe", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                     ctypes.c_void_p]),
    ("xg_tran_info", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("xg_tran_set_graphs", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("xg_tran_sync", ctypes.c_int, [ctypes.c_void_p]),
    ("xg_tran_counters", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("xg_tran_restore", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int64,
                                       ctypes.c_int64, ctypes.c_int64]),
    ("xg_instance_state", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                         ctypes.c_void_p]),
    ("xg_tran_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("xg_autotune_info", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("xg_set_eval_variant", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("xg_set_output_mode", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("xg_sizes", ctypes.c_int, [ctypes.c_void_p, c_int32_p, c_int32_p, c_int32_p, c_int32_p, c_int32_p]),
]


class XgError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libxg_load.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise XgError("libxg_load.so is not built: run `python -m xyce_amd.build`")
        # torch first: its HIP runtime (soname libamdhip64.so.7) then serves
        # this library too.  Loaded the other way round, the process holds
        # two HIP runtimes and torch's own fails to find the GPU.
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != XG_OK:
        msg = lib().xg_last_error().decode(errors="replace")
        raise XgError("%s failed: %s (%s)" % (what, ERRORS.get(rc, rc), msg))


_KEY_ARRAYS = {}


def keys_vals(params):
    """dict -> (n, char**, double*) for the name/value parameter lists (the
    name arrays are built once per distinct list of names)."""
    names = tuple(params)
    n = len(names)
    keys = _KEY_ARRAYS.get(names)
    if keys is None:
        keys = (ctypes.c_char_p * max(n, 1))(*[k.encode() for k in names])
        if len(_KEY_ARRAYS) < 4096:
            _KEY_ARRAYS[names] = keys
    vals = (ctypes.c_double * max(n, 1))(*[float(v) for v in params.values()])
    return n, keys, vals
