"""ctypes binding of libpoutre_amd.so (the C ABI declared in include/poutre_amd.h).

The library is loaded from this package directory only; if it is missing the import
of any operator fails loudly (there is no CPU fallback).
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpoutre_amd.so")

# PType bit values (include/poutre/base/types.hpp:517-527) + uint16 extension
PTYPE_U8, PTYPE_I32, PTYPE_F32, PTYPE_I64, PTYPE_D64, PTYPE_U16 = 4, 8, 16, 32, 64, 128

STATUS = {0: "PM_OK", -1: "PM_E_SIZE", -2: "PM_E_PTYPE", -3: "PM_E_ALIAS", -4: "PM_E_RANK", -5: "PM_E_SE",
          -6: "PM_E_ARG", -7: "PM_E_HIP", -8: "PM_E_UNSUPPORTED", -9: "PM_E_NOCONV"}


class PmImage(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("ptype", ctypes.c_int32), ("rank", ctypes.c_int32),
                ("shape", ctypes.c_int64 * 3)]


class PmOperand(ctypes.Structure):
    """pm_operand of pm_compare: an image, or (image NULL) the scalar value"""
    _fields_ = [("image", ctypes.POINTER(PmImage)), ("value", ctypes.c_double)]


class PoutreError(RuntimeError):
    """Raised when an entry point returns a negative status (the reference throws
    std::runtime_error in the same situations, include/poutre/base/config.hpp:203-215)."""

    def __init__(self, code, msg):
        super().__init__(f"{msg} [{STATUS.get(code, code)}]")
        self.code = code


_lib = None

_P = ctypes.POINTER(PmImage)
_VP = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_D = ctypes.c_double

SIGNATURES = {
    "pm_erode_static": [_P, _I, _I, _P, _VP],
    "pm_dilate_static": [_P, _I, _I, _P, _VP],
    "pm_erode_compound": [_P, _I, _I, _P, _VP],
    "pm_dilate_compound": [_P, _I, _I, _P, _VP],
    "pm_erode_nl": [_P, ctypes.POINTER(_I64), _I, _P, _VP],
    "pm_dilate_nl": [_P, ctypes.POINTER(_I64), _I, _P, _VP],
    "pm_erode_x": [_P, _I64, _P, _VP],
    "pm_dilate_x": [_P, _I64, _P, _VP],
    "pm_erode_y": [_P, _I64, _P, _VP],
    "pm_dilate_y": [_P, _I64, _P, _VP],
    "pm_erode_rect": [_P, _I64, _I64, _P, _VP],
    "pm_dilate_rect": [_P, _I64, _I64, _P, _VP],
    "pm_opening": [_P, _I, _I, _P, _VP],
    "pm_closing": [_P, _I, _I, _P, _VP],
    "pm_gradient": [_P, _I, _I, _P, _VP],
    "pm_reconstruct": [_I, _P, _P, _I, _P, _VP],
    "pm_reconstruct_noalloc": [_I, _P, _P, _I, _P, _VP],
    "pm_debug_set_max_rounds": [_I],
    "pm_debug_fail_allocations": [_I],
    "pm_low_leveling": [_P, _P, _I, _P, _VP],
    "pm_high_leveling": [_P, _P, _I, _P, _VP],
    "pm_leveling": [_P, _P, _I, _P, _VP],
    "pm_h_maxima": [_P, _D, _I, _P, _VP],
    "pm_h_minima": [_P, _D, _I, _P, _VP],
    "pm_h_convex": [_P, _D, _I, _P, _VP],
    "pm_h_concave": [_P, _D, _I, _P, _VP],
    "pm_dynamic_pseudo_opening": [_P, _D, _I, _P, _VP],
    "pm_dynamic_pseudo_closing": [_P, _D, _I, _P, _VP],
    "pm_label_binary": [_P, _I, _P, ctypes.POINTER(_I64), _VP],
    "pm_label_flat_zones": [_P, _I, _P, ctypes.POINTER(_I64), _VP],
    "pm_label_minima": [_P, _I, _P, _VP],
    "pm_label_maxima": [_P, _I, _P, _VP],
    "pm_arith_sup": [_P, _P, _P, _VP],
    "pm_arith_inf": [_P, _P, _P, _VP],
    "pm_arith_ssub_image": [_P, _P, _P, _VP],
    "pm_arith_ssub_const": [_P, _D, _P, _VP],
    "pm_arith_sadd_const": [_P, _D, _P, _VP],
    "pm_arith_sadd_image": [_P, _P, _P, _VP],
    "pm_arith_invert": [_P, _P, _VP],
    "pm_convert_into": [_P, _P, _VP],
    "pm_compare": [_P, _I, ctypes.POINTER(PmOperand), ctypes.POINTER(PmOperand), ctypes.POINTER(PmOperand), _P, _VP],
    "pm_maxplus_fwd": [_VP, _I64, _I64, _I64, _VP, _I, _I, _VP, _VP],
    "pm_maxplus_bwd": [_VP, _VP, _I64, _I64, _I64, _VP, _I, _I, _VP, _VP, _VP],
    "pm_maxplus_fwd_sel": [_VP, _I64, _I64, _I64, _VP, _I, _I, _VP, _VP, _VP],
    "pm_maxplus_bwd_sel": [_VP, _VP, _I64, _I64, _I64, _I, _I, _I, _VP, _VP, _VP],
    "pm_batch_dilate_open": [_I, _P, ctypes.POINTER(_I), _P, _P, _VP],
    "pm_batch_reconstruct": [_I, _I, _P, _P, _I, _P, _VP],
    "pm_batch_plan_create": [_I, _P, ctypes.POINTER(_I), _P, _P, ctypes.POINTER(ctypes.c_void_p)],
    "pm_batch_plan_launch": [ctypes.c_void_p, _VP],
    "pm_batch_plan_destroy": [ctypes.c_void_p],
    "pm_autotune": [_VP],
    "pm_tuning_get": [_I, _I],
    "pm_tuning_set": [_I, _I, _I],
}


def lib():
    """Load libpoutre_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m poutre2_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        L.pm_last_error.restype = ctypes.c_char_p
        L.pm_last_error.argtypes = []
        L.pm_version.restype = ctypes.c_char_p
        L.pm_version.argtypes = []
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise PoutreError(rc, lib().pm_last_error().decode(errors="replace"))
