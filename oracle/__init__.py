"""ctypes wrapper over the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py, as the checker the GPU path is compared against.
The product package (poutre2_amd) never imports it.

Arrays are numpy arrays (C-contiguous); shapes follow the reference's
``{H, W}`` / ``{D, H, W}`` order with the last axis fastest.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

T_CODES = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.int32): 2,
           np.dtype(np.int64): 3, np.dtype(np.float32): 4, np.dtype(np.float64): 5}

# Common_NL_SE (include/poutre/structuring_element/se_types_and_tags.hpp:45-57)
SE = {"SESegmentX1D": 1, "SESquare2D": 2, "SECross2D": 3, "SESegmentX2D": 4, "SESegmentY2D": 5,
      "SESegmentX3D": 6, "SESegmentY3D": 7, "SESegmentZ3D": 8, "SECross3D": 9, "SESquare3D": 10}

_lib = None


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        vp, i64p = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)
        ci, cd, c64 = ctypes.c_int, ctypes.c_double, ctypes.c_int64
        sig = {
            "or_se_offsets": [ci, i64p, ctypes.POINTER(ci)],
            "or_line_x": [ci, vp, vp, ci, i64p, c64, ci],
            "or_line_y": [ci, vp, vp, ci, i64p, c64, ci],
            "or_transpose": [ci, vp, vp, i64p],
            "or_static": [ci, vp, vp, ci, i64p, ci, ci],
            "or_static_iter": [ci, vp, vp, ci, i64p, ci, ci, ci],
            "or_nl": [ci, vp, vp, ci, i64p, i64p, ci, ci],
            "or_compound": [ci, vp, vp, ci, i64p, ci, ci, ci],
            "or_reconstruct": [ci, ci, vp, vp, vp, ci, i64p, ci],
            "or_reconstruct_bf": [ci, ci, vp, vp, vp, ci, i64p, ci],
            "or_low_leveling": [ci, vp, vp, vp, ci, i64p, ci],
            "or_high_leveling": [ci, vp, vp, vp, ci, i64p, ci],
            "or_leveling": [ci, vp, vp, vp, ci, i64p, ci],
            "or_arith_ssub_const": [ci, vp, cd, vp, c64],
            "or_arith_sadd_const": [ci, vp, cd, vp, c64],
            "or_arith_ssub_image": [ci, vp, vp, vp, c64],
            "or_arith_sadd_image": [ci, vp, vp, vp, c64],
            "or_arith_invert": [ci, vp, vp, c64],
            "or_arith_sup": [ci, vp, vp, vp, c64],
            "or_arith_inf": [ci, vp, vp, vp, c64],
            "or_label": [ci, ci, vp, vp, ci, i64p, ci, i64p],
        }
        for name in ("or_h_maxima", "or_h_minima", "or_h_convex", "or_h_concave",
                     "or_dynamic_pseudo_opening", "or_dynamic_pseudo_closing"):
            sig[name] = [ci, vp, cd, vp, ci, i64p, ci]
        for name, args in sig.items():
            f = getattr(_lib, name)
            f.argtypes = args
            f.restype = ci
    return _lib


class OracleError(RuntimeError):
    pass


def _shape(a):
    return (ctypes.c_int64 * max(1, a.ndim))(*a.shape)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(rc, what):
    if rc != 0:
        raise OracleError(f"{what} failed with code {rc}")


def _t(a):
    a = np.ascontiguousarray(a)
    return a, T_CODES[a.dtype]


def se_offsets(se):
    offs = (ctypes.c_int64 * 81)()
    rank = ctypes.c_int(0)
    n = lib().or_se_offsets(SE[se] if isinstance(se, str) else se, offs, ctypes.byref(rank))
    _check(0 if n > 0 else n, "se_offsets")
    return np.array(offs[: n * rank.value], dtype=np.int64).reshape(n, rank.value)


def line_x(img, r, dilate=True):
    img, t = _t(img)
    out = np.empty_like(img)
    _check(lib().or_line_x(t, _p(img), _p(out), img.ndim, _shape(img), r, int(dilate)), "line_x")
    return out


def line_y(img, r, dilate=True):
    img, t = _t(img)
    out = np.empty_like(img)
    _check(lib().or_line_y(t, _p(img), _p(out), img.ndim, _shape(img), r, int(dilate)), "line_y")
    return out


def transpose(img):
    """t_transpose (transpose_op_t.hpp:45-70) as the reference indexes it: a true transpose of
    square images only (the reference rejects other shapes); not a general transpose."""
    img, t = _t(img)
    out = np.empty((img.shape[1], img.shape[0]), dtype=img.dtype)
    _check(lib().or_transpose(t, _p(img), _p(out), _shape(img)), "transpose")
    return out


def static(img, se, dilate=True):
    img, t = _t(img)
    out = np.empty_like(img)
    _check(lib().or_static(t, _p(img), _p(out), img.ndim, _shape(img), SE[se] if isinstance(se, str) else se,
                           int(dilate)), "static")
    return out


def static_iter(img, se, it, dilate=True):
    img, t = _t(img)
    out = np.empty_like(img)
    _check(lib().or_static_iter(t, _p(img), _p(out), img.ndim, _shape(img),
                                SE[se] if isinstance(se, str) else se, it, int(dilate)), "static_iter")
    return out


def nl(img, offsets, dilate=True):
    img, t = _t(img)
    offs = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64).reshape(-1, img.ndim))
    out = np.empty_like(img)
    _check(lib().or_nl(t, _p(img), _p(out), img.ndim, _shape(img),
                       offs.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), offs.shape[0], int(dilate)), "nl")
    return out


def compound(img, compound_se, size, dilate=True):
    img, t = _t(img)
    out = np.empty_like(img)
    _check(lib().or_compound(t, _p(img), _p(out), img.ndim, _shape(img), compound_se, size, int(dilate)),
           "compound")
    return out


def _rec(fn, rec_type, marker, mask, se):
    marker, t = _t(marker)
    mask = np.ascontiguousarray(mask, dtype=marker.dtype)
    out = np.empty_like(mask)
    _check(fn(t, rec_type, _p(marker), _p(mask), _p(out), mask.ndim, _shape(mask),
              SE[se] if isinstance(se, str) else se), "reconstruct")
    return out


def reconstruct(rec_type, marker, mask, se):
    """rec_type: 'dilate' or 'erode' (geo::reconstruction_type)."""
    return _rec(lib().or_reconstruct, 1 if rec_type == "dilate" else 0, marker, mask, se)


def reconstruct_bf(rec_type, marker, mask, se):
    return _rec(lib().or_reconstruct_bf, 1 if rec_type == "dilate" else 0, marker, mask, se)


def _lev(fn, ref, marker, se):
    ref, t = _t(ref)
    marker = np.ascontiguousarray(marker, dtype=ref.dtype)
    out = np.empty_like(ref)
    _check(fn(t, _p(ref), _p(marker), _p(out), ref.ndim, _shape(ref), SE[se] if isinstance(se, str) else se),
           "leveling")
    return out


def low_leveling(ref, marker, se):
    return _lev(lib().or_low_leveling, ref, marker, se)


def high_leveling(ref, marker, se):
    return _lev(lib().or_high_leveling, ref, marker, se)


def leveling(ref, marker, se):
    return _lev(lib().or_leveling, ref, marker, se)


def _ew_const(fn, img, c):
    img, t = _t(img)
    out = np.empty_like(img)
    _check(fn(t, _p(img), float(c), _p(out), img.size), "arith")
    return out


def ssub_const(img, c):
    return _ew_const(lib().or_arith_ssub_const, img, c)


def sadd_const(img, c):
    return _ew_const(lib().or_arith_sadd_const, img, c)


def _ew2(fn, a, b):
    a, t = _t(a)
    b = np.ascontiguousarray(b, dtype=a.dtype)
    out = np.empty_like(a)
    _check(fn(t, _p(a), _p(b), _p(out), a.size), "arith")
    return out


def ssub_image(a, b):
    return _ew2(lib().or_arith_ssub_image, a, b)


def sadd_image(a, b):
    return _ew2(lib().or_arith_sadd_image, a, b)


def invert(img):
    img, t = _t(img)
    out = np.empty_like(img)
    _check(lib().or_arith_invert(t, _p(img), _p(out), img.size), "arith")
    return out


# CompareImage: CompOpType bit values (include/poutre/base/types.hpp:443-453) and the operators
# of compare_op_t.hpp:40-165 (the input pixel on the left)
CMP_OPS = {"Equal": 1, "Diff": 2, "Sup": 4, "SupEqual": 8, "Inf": 16, "InfEqual": 32}
_CMP_FN = {1: np.equal, 2: np.not_equal, 4: np.greater, 8: np.greater_equal, 16: np.less, 32: np.less_equal}


def compare(img, op, comp, vtrue, vfalse, out_dtype):
    """CompareImage, every operand form (compare.hpp:31-84 -> compare.cpp; select of
    compare_op_t.hpp:305-351): out = (img OP comp) ? vtrue : vfalse.  comp is an image of
    img's type or a scalar of that type; vtrue / vfalse are images or scalars of the output
    type.  Elementwise numpy (the comparisons are IEEE's, as C++'s)."""
    img = np.asarray(img)
    out_dtype = np.dtype(out_dtype)
    c = comp if isinstance(comp, np.ndarray) else np.array(comp).astype(img.dtype)
    tv = vtrue if isinstance(vtrue, np.ndarray) else np.array(vtrue).astype(out_dtype)
    fv = vfalse if isinstance(vfalse, np.ndarray) else np.array(vfalse).astype(out_dtype)
    cond = _CMP_FN[int(op)](img, c)
    return np.where(cond, tv, fv).astype(out_dtype)


def sup(a, b):
    return _ew2(lib().or_arith_sup, a, b)


def inf(a, b):
    return _ew2(lib().or_arith_inf, a, b)


def _extrema(name, img, h, se):
    img, t = _t(img)
    out = np.empty_like(img)
    _check(getattr(lib(), "or_" + name)(t, _p(img), float(h), _p(out), img.ndim, _shape(img),
                                        SE[se] if isinstance(se, str) else se), name)
    return out


def h_maxima(img, h, se):
    return _extrema("h_maxima", img, h, se)


def h_minima(img, h, se):
    return _extrema("h_minima", img, h, se)


def h_convex(img, h, se):
    return _extrema("h_convex", img, h, se)


def h_concave(img, h, se):
    return _extrema("h_concave", img, h, se)


def dynamic_pseudo_opening(img, h, se):
    return _extrema("dynamic_pseudo_opening", img, h, se)


def dynamic_pseudo_closing(img, h, se):
    return _extrema("dynamic_pseudo_closing", img, h, se)


def label(img, se, flat=False):
    """label_binary (flat=False) / label_flat_zones (flat=True): (int64 labels, count);
    label_t.hpp t_label_helper, see oracle.h."""
    img, t = _t(img)
    out = np.empty(img.shape, dtype=np.int64)
    cnt = ctypes.c_int64(0)
    _check(lib().or_label(t, int(flat), _p(img), _p(out), img.ndim, _shape(img), SE[se] if isinstance(se, str) else se,
                          ctypes.byref(cnt)), "label")
    return out, cnt.value
