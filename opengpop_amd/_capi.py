"""ctypes binding of the engine's C ABI (include/ogp.h), loaded from the in-tree libogp.so.

There is no fallback: if the library is missing, or no HIP device is present, the calls raise.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libogp.so")
DATA_DIR = os.path.join(os.path.dirname(PKG), "data")

OK, ERR_ARGUMENT, ERR_DATA, ERR_DEVICE, ERR_STATE, ERR_NONFINITE = range(6)
OBJECT_OK, OBJECT_NONFINITE_INPUT, OBJECT_NONFINITE_RESULT = range(3)
FORCE_GRAVITY, FORCE_NONE, FORCE_CONST, FORCE_POLY7, FORCE_POLY8, FORCE_POLY9, FORCE_EXP, \
    FORCE_DAMPED = range(8)
PRECISION_FP64, PRECISION_SPLIT = 0, 1
QUIRK_RK8_ROW7, QUIRK_BACKWARD_TIMES, QUIRKS_REFERENCE = 1, 2, 3

_P = ctypes.c_void_p
_D = ctypes.c_double
_I64 = ctypes.c_int64


class OgpConfig(ctypes.Structure):
    _fields_ = [("data_dir", ctypes.c_char_p), ("degree", ctypes.c_int),
                ("precision", ctypes.c_int), ("split_degree", ctypes.c_int),
                ("n_corrections", ctypes.c_int), ("quirks", ctypes.c_int),
                ("force", ctypes.c_int), ("year", ctypes.c_int), ("month", ctypes.c_int),
                ("day", ctypes.c_int), ("hour", ctypes.c_int), ("minute", ctypes.c_int),
                ("second", ctypes.c_double)]


class OgpTiming(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("total_ms", "nodes_ms", "startup_ms", "steps_ms", "h2d_ms", "d2h_ms",
                 "step_launches", "object_steps", "gj_object_steps", "launch_ms_sum",
                 "evaluations")]


class OgpSources(ctypes.Structure):
    _fields_ = [("mask", ctypes.c_int), ("drag_bc", ctypes.c_double)]


SOURCE_SUN, SOURCE_MOON, SOURCE_DRAG = 1, 2, 4


class OgpConjTiming(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("propagate_ms", "screen_ms", "refine_ms", "pair_steps", "pairs", "candidates",
                 "screened_pair_steps")]


class OgpAdjointTiming(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("forward_ms", "adjoint_ms", "stm_ms")]


class OgpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[ogp error {code}] {msg}")
        self.code = code


_lib = None


def lib():
    """Loads libogp.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"{LIB_PATH} is missing: build it with python -m opengpop_amd.build")
    # PyTorch ships its own libamdhip64.so (soname libamdhip64.so.7) and loads it by a path of
    # its own. Loaded first, it satisfies libogp.so's libamdhip64.so.7 as well and the process
    # has one HIP runtime; loaded after libogp.so, it brings a second runtime that sees no
    # device, and device tensors passed to the *_device entry points could not be made.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "ogp_default_config": (None, [ctypes.POINTER(OgpConfig)]),
        "ogp_create": (ctypes.c_int, [ctypes.POINTER(OgpConfig), ctypes.POINTER(_P)]),
        "ogp_destroy": (None, [_P]),
        "ogp_last_error": (ctypes.c_char_p, [_P]),
        "ogp_device_bytes": (_I64, [_P]),
        "ogp_gravity_itrs": (ctypes.c_int, [_P, _P, _I64, _P]),
        "ogp_accel": (ctypes.c_int, [_P, _P, _P, _I64, _D, _I64, _P, _I64]),
        "ogp_time_nodes": (ctypes.c_int, [_P, _D, ctypes.c_int, _P, _I64, _P, _P]),
        "ogp_propagate": (ctypes.c_int, [_P, _P, _P, _I64, _P, _P, _P, _P]),
        "ogp_propagate_device": (ctypes.c_int, [_P, _P, _P, _I64, _P, _P, _P, _P]),
        "ogp_perturbed_satellites": (ctypes.c_int, [_P, _P, _P, _I64, ctypes.c_double,
                                                    ctypes.c_uint64, _P, _P, _P, _P]),
        "ogp_perturbed_satellites_device": (ctypes.c_int, [_P, _P, _P, _I64, ctypes.c_double,
                                                           ctypes.c_uint64, _P, _P, _P, _P,
                                                           _I64]),
        "ogp_last_timing": (ctypes.c_int, [_P, ctypes.POINTER(OgpTiming)]),
        "ogp_set_tuning": (ctypes.c_int, [_P, ctypes.c_int, _I64]),
        "ogp_set_slices": (ctypes.c_int, [_P, ctypes.c_int]),
        "ogp_set_fixed_point_exit": (ctypes.c_int, [_P, ctypes.c_int]),
        "ogp_set_lockstep": (ctypes.c_int, [_P, ctypes.c_int]),
        "ogp_set_screen_shape": (ctypes.c_int, [_P, ctypes.c_int]),
        "ogp_set_radial_filter": (ctypes.c_int, [_P, ctypes.c_int]),
        "ogp_get_screen_shape": (ctypes.c_int, [_P, _I64, ctypes.POINTER(ctypes.c_int)]),
        "ogp_get_lockstep": (ctypes.c_int, [_P, _I64, ctypes.POINTER(ctypes.c_int)]),
        "ogp_set_derivative_memory": (ctypes.c_int, [_P, _I64]),
        "ogp_set_sources": (ctypes.c_int, [_P, ctypes.POINTER(OgpSources)]),
        "ogp_get_sources": (ctypes.c_int, [_P, ctypes.POINTER(OgpSources)]),
        "ogp_set_drag_coefficients": (ctypes.c_int, [_P, _P, _I64]),
        "ogp_get_tuning": (ctypes.c_int, [_P, _I64, ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(_I64)]),
        "ogp_prop_init": (ctypes.c_int, [_P, _P, _P, _I64, _D, _I64]),
        "ogp_prop_step": (ctypes.c_int, [_P, _P, _P, _P]),
        "ogp_propagator_create": (ctypes.c_int, [_P, _P, _P, _I64, _D, _I64,
                                                 ctypes.POINTER(_P)]),
        "ogp_propagator_create_host": (ctypes.c_int, [_P, _P, _P, _I64, _D, _I64,
                                                      ctypes.POINTER(_P)]),
        "ogp_propagator_step": (ctypes.c_int, [_P, _P, _P, _P]),
        "ogp_propagator_acceleration": (ctypes.c_int, [_P, _P]),
        "ogp_propagator_copy_state_to_host": (ctypes.c_int, [_P, _P, _P]),
        "ogp_propagator_destroy": (None, [_P]),
        "ogp_startup": (ctypes.c_int, [_P, _P, _P, _I64, _D, _P, _P, _P, _P, _P]),
        "ogp_startup_device": (ctypes.c_int, [_P, _P, _P, _I64, _D, _P, _P, _P, _P, _P]),
        "ogp_predict_device": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _I64, _D, _I64]),
        "ogp_accel_device": (ctypes.c_int, [_P, _P, _P, _I64, _D, _I64, _P, _I64]),
        "ogp_accel_generation": (ctypes.c_int, [_P, _P, _P, _I64, _D, ctypes.c_int,
                                                ctypes.c_int, _P]),
        "ogp_propagate_checked": (ctypes.c_int, [_P, _P, _P, _I64, _P, _P, _P, _P, _P]),
        "ogp_propagate_device_checked": (ctypes.c_int, [_P, _P, _P, _I64, _P, _P, _P, _P,
                                                        _P]),
        "ogp_kat_predict": (ctypes.c_int, [_P, _P, _P, _P, _I64, _D, _P, _P]),
        "ogp_kat_finalize_correction": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _I64, _D]),
        "ogp_conj_init": (ctypes.c_int, [_P, _P, _P, _P, _I64, _D]),
        "ogp_conj_propagate": (ctypes.c_int, [_P, _I64]),
        "ogp_conj_find": (ctypes.c_int, [_P, _D, ctypes.POINTER(_I64)]),
        "ogp_conj_results": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _P]),
        "ogp_conj_find_noalloc": (ctypes.c_int, [_P, _D, _I64, _P, _P, _P, _P, _P, _P, _P,
                                                 ctypes.POINTER(_I64)]),
        "ogp_conj_screen": (ctypes.c_int, [_P, _P, _P]),
        "ogp_conj_num_pairs": (ctypes.c_int, [_P, ctypes.POINTER(_I64)]),
        "ogp_conj_trajectory": (ctypes.c_int, [_P, _P, ctypes.POINTER(_I64)]),
        "ogp_conj_last_timing": (ctypes.c_int, [_P, ctypes.POINTER(OgpConjTiming)]),
        "ogp_gravity_jacobian_itrs": (ctypes.c_int, [_P, _P, _I64, _P, _P]),
        "ogp_stm": (ctypes.c_int, [_P, _P, _P, _I64, _D, _P, _P, _P, _P]),
        "ogp_adjoint_jacobian": (ctypes.c_int, [_P, _P, _P, _I64, _D, _P, _I64, _P, _P, _P]),
        "ogp_adjoint_vjp": (ctypes.c_int, [_P, _P, _P, _I64, _D, _P, _I64, _P, _P, _P, _P]),
        "ogp_adjoint_last_timing": (ctypes.c_int, [_P, ctypes.POINTER(OgpAdjointTiming)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


# Entry points every build must export (checked by the CPU test suite).
def declared_symbols():
    hdr = os.path.join(os.path.dirname(PKG), "include", "ogp.h")
    import re
    text = open(hdr).read()
    return sorted(set(re.findall(r"\b(ogp_[a-z_0-9]+)\s*\(", text)))


def ptr(a):
    return None if a is None else a.ctypes.data


def f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.reshape(shape) if shape is not None else a


def check(ctx_handle, rc):
    if rc != OK:
        msg = lib().ogp_last_error(ctx_handle)
        raise OgpError(rc, msg.decode() if msg else "unknown error")
