"""poutre2_amd -- MI355X (gfx950) implementation of Poutre2's morphology / geodesy hot path.

Python entry points with the names and argument order of the reference's nanobind
module ``pypoutre`` (python/low_level_morpho/ero_dil.cpp:12-27,
python/geodesy/geodesy.cpp:12-29, python/structuring_element/se_types.cpp:10-36),
operating on device-resident images: contiguous torch tensors on a ROCm device, of
dtype uint8 / uint16 / int32 / int64 / float32 / float64, shaped ``(W,)``, ``(H, W)`` or
``(D, H, W)``.  Every call goes through the C ABI of libpoutre_amd.so on the current
torch stream; there is no CPU fallback.

Differences from pypoutre, all additive:
  * ``out`` may be omitted, in which case a new tensor is returned;
  * ``h_minima`` calls h_minima (the reference binds h_maxima under that name,
    python/geodesy/geodesy.cpp:19);
  * ``reconstruction`` is bound (the reference has no binding for it);
  * ``opening`` / ``closing`` / ``gradient`` / ``erode_rect`` / ``dilate_rect`` are added;
  * images may also be any object exporting DLPack (``__dlpack__``), used in place (zero
    copy); ``factory_image`` / ``from_string`` return device tensors and ``to_string``
    accepts them (python/base/image_interface.cpp:13-31 binds no buffer interop at all).
"""
import ctypes
import enum

import numpy as np

from . import _capi
from ._capi import PoutreError

__all__ = ["static_nl_se", "static_compound_se", "RecType", "erode", "dilate", "erodeX", "dilateX", "erodeY",
           "dilateY", "erode_rect", "dilate_rect", "opening", "closing", "gradient", "reconstruction",
           "h_minima", "h_maxima", "h_concave", "h_convex", "dynamic_pseudo_closing", "dynamic_pseudo_opening",
           "high_leveling", "low_leveling", "leveling", "arith_sup", "arith_inf", "arith_ssub_image",
           "arith_ssub_const", "arith_sadd_const", "batch_dilate_open", "batch_reconstruction", "PoutreError",
           "version", "PType", "CompoundType", "image_ctype", "factory_image", "from_string", "to_string", "as_sizes_compatible",
           "as_types_compatible", "are_images_different", "CompOpType", "compare", "arith_invert", "arith_sadd_image",
           "arith_invert_image", "arith_inf_image", "arith_sup_image", "arith_saturated_add_image",
           "arith_saturated_sub_image", "arith_saturated_add_constant", "arith_saturated_sub_constant", "clone",
           "clone_geometry", "convert", "convert_geometry", "copy_into", "convert_into", "SE", "se_type"]


class static_nl_se(enum.IntEnum):
    """se::Common_NL_SE (se_types_and_tags.hpp:45-57), pypoutre names."""
    segx1d = 1
    square2d = 2
    cross2d = 3
    segx2d = 4
    segy2d = 5
    segx3d = 6
    segy3d = 7
    segz3d = 8
    cross3d = 9
    square3d = 10


class se_type(enum.IntEnum):
    """poutre::se::se_type (python/structuring_element/se_types.cpp:12-17)."""
    runtime = 0
    neighbor_list = 1
    image = 2
    chain = 3


class SE:
    """A runtime structuring element: the neighbour list of poutre::se::neighbor_list_t
    (neighbor_list_se_t.hpp:36-228) with the methods pypoutre's ``SE`` binds
    (python/structuring_element/se_interface.cpp:15-23).  Built from offsets ({dx}, {dy, dx}
    or {dz, dy, dx}); ``dilate`` / ``erode`` take it as the runtime neighbour-list SE."""

    def __init__(self, offsets):
        self._c = [tuple(int(v) for v in o) for o in offsets]
        if len({len(c) for c in self._c}) > 1:
            raise PoutreError(-5, "neighbor_list: coordinate of wrong rank")

    def __iter__(self):
        return iter(self._c)

    def __len__(self):
        return len(self._c)

    def __repr__(self):
        return f"SE({self._c})"

    def _centre(self):
        """the first offset, taken as the centre (neighbor_list_se_t.hpp:96-125)"""
        return self._c[0] if self._c else ()

    def transpose(self):
        """offsets mirrored about the centre: 2 * centre - offset"""
        ctr = self._centre()
        return SE([tuple(2 * a - v for a, v in zip(ctr, c)) for c in self._c])

    def type(self):
        return se_type.neighbor_list

    def remove_center(self):
        ctr = self._centre()
        return SE([c for c in self._c if c != ctr])

    def size(self):
        return len(self._c)

    def clone(self):
        return SE(self._c)

    def is_equal(self, other):
        """the same offsets in the same order"""
        return isinstance(other, SE) and self._c == other._c

    def is_equal_unordered(self, other):
        """the same sets of offsets (sorted, repeats removed: neighbor_list_se_t.hpp:177-195)"""
        return isinstance(other, SE) and sorted(set(self._c)) == sorted(set(other._c))


class static_compound_se(enum.IntEnum):
    """se::Compound_NL_SE (se_types_and_tags.hpp:59-63)."""
    octagon = 1
    rhombicuboctahedron = 2


class RecType(enum.IntEnum):
    """geo::reconstruction_type (mreconstruct.hpp:30-34)."""
    erode = 0
    dilate = 1


def _torch():
    import torch
    return torch


_PTYPES = None
# numpy record layout of pm_image (include/poutre_amd.h): void *data; int32 ptype, rank; int64 shape[3]
_PM_IMAGE_NP = np.dtype([("data", "<u8"), ("ptype", "<i4"), ("rank", "<i4"), ("shape", "<i8", (3,))])


def _ptype(t):
    global _PTYPES
    if _PTYPES is None:
        torch = _torch()
        m = {torch.uint8: _capi.PTYPE_U8, torch.int32: _capi.PTYPE_I32, torch.float32: _capi.PTYPE_F32,
             torch.int64: _capi.PTYPE_I64, torch.float64: _capi.PTYPE_D64}
        if hasattr(torch, "uint16"):
            m[torch.uint16] = _capi.PTYPE_U16
        _PTYPES = m
    p = _PTYPES.get(t.dtype)
    if p is None:
        raise PoutreError(-2, f"Unsupported PType for dtype {t.dtype}")
    return p


def _img(x):
    """x as a torch tensor: itself, or a zero-copy view of a DLPack producer."""
    torch = _torch()
    if isinstance(x, torch.Tensor):
        return x
    if hasattr(x, "__dlpack__"):
        return torch.from_dlpack(x)
    raise PoutreError(-1, f"images must be torch tensors or DLPack producers, not {type(x).__name__}")


def _desc(t):
    if image_ctype(t) != CompoundType.Scalar:
        raise PoutreError(-2, "Unsupported compound type")
    t = _img(t)
    if not t.is_cuda:
        raise PoutreError(-6, "images must be device tensors (HBM-resident)")
    if not t.is_contiguous():
        raise PoutreError(-6, "images must be contiguous")
    if t.dim() < 1 or t.dim() > 3:
        raise PoutreError(-4, f"Unsupported number of dims {t.dim()}")
    d = _capi.PmImage()
    d.data = t.data_ptr()
    d.ptype = _ptype(t)
    d.rank = t.dim()
    for i, s in enumerate(t.shape):
        d.shape[i] = s
    return d


def _descs(ts):
    """pm_image array for a list of device tensors, built as one numpy record array (the
    per-image Python work of a 512-image batch would otherwise rival its kernel time)."""
    if any(image_ctype(t) != CompoundType.Scalar for t in ts):
        raise PoutreError(-2, "Unsupported compound type")
    ts = [_img(t) for t in ts]
    n = len(ts)
    rec = np.zeros(n, dtype=_PM_IMAGE_NP)
    data, ptype, rank, shape = [], [], [], []
    for t in ts:
        if not t.is_cuda:
            raise PoutreError(-6, "images must be device tensors (HBM-resident)")
        if not t.is_contiguous():
            raise PoutreError(-6, "images must be contiguous")
        d = t.dim()
        if d < 1 or d > 3:
            raise PoutreError(-4, f"Unsupported number of dims {d}")
        data.append(t.data_ptr())
        ptype.append(_ptype(t))
        rank.append(d)
        shape.append(tuple(t.shape) + (0,) * (3 - d))
    rec["data"] = data
    rec["ptype"] = ptype
    rec["rank"] = rank
    rec["shape"] = shape
    return rec


def _rec_ptr(rec):
    return ctypes.cast(rec.ctypes.data, ctypes.POINTER(_capi.PmImage))


def _stream(t):
    return ctypes.c_void_p(_torch().cuda.current_stream(_img(t).device).cuda_stream)


def _out(img, out):
    return _torch().empty_like(_img(img)) if out is None else out


def _call(name, *args):
    _capi.check(getattr(_capi.lib(), name)(*args))


def version():
    return _capi.lib().pm_version().decode()


# ---------------------------------------------------------------- image model (python/base)
class PType(enum.IntEnum):
    """poutre::PType bit values (include/poutre/base/types.hpp:517-527, python/base/types.cpp:10-17);
    GrayUINT16 is an addition."""
    GrayUINT8 = 1 << 2
    GrayINT32 = 1 << 3
    F32 = 1 << 4
    GrayINT64 = 1 << 5
    D64 = 1 << 6
    GrayUINT16 = 1 << 7


# poutre::CompoundType (types.hpp; python/base/types.cpp:19-23 names "Scalar", "3Planes", "4Planes")
CompoundType = enum.IntEnum("CompoundType", [("Scalar", 1), ("3Planes", 2), ("4Planes", 4)])

_PT_INFO = {PType.GrayUINT8: ("GUINT8", np.uint8), PType.GrayUINT16: ("GUINT16", np.uint16),
            PType.GrayINT32: ("GINT32", np.int32), PType.GrayINT64: ("GINT64", np.int64),
            PType.F32: ("F32", np.float32), PType.D64: ("D64", np.float64)}


def _torch_dtype(np_dtype):
    torch = _torch()
    return {np.uint8: torch.uint8, np.uint16: getattr(torch, "uint16", None), np.int32: torch.int32,
            np.int64: torch.int64, np.float32: torch.float32, np.float64: torch.float64}[np_dtype]


_PLANES = {CompoundType.Scalar: 1, CompoundType["3Planes"]: 3, CompoundType["4Planes"]: 4}


def _tag(t, ctype):
    """compound images are tensors with a trailing planes axis, tagged with their CompoundType
    (the operators take Scalar images only, like the reference's dispatch)"""
    if ctype != CompoundType.Scalar:
        t._poutre_ctype = ctype
    return t


def image_ctype(img):
    """the CompoundType of an image made by factory_image / from_string (Scalar otherwise)"""
    return getattr(img, "_poutre_ctype", CompoundType.Scalar)


def factory_image(dims, ctype, ptype, device="cuda"):
    """poutre::Create(dims, ctype, ptype) (python/base/image_interface.cpp:27): a zero-filled
    image of that shape and type, as a device tensor; 3Planes / 4Planes images carry a
    trailing axis of 3 / 4 planes."""
    ctype = CompoundType(ctype)
    dims = tuple(int(d) for d in dims)
    if len(dims) < 1 or len(dims) > 4:
        raise PoutreError(-4, f"Unsupported number of dims {len(dims)}")
    shape = dims + ((_PLANES[ctype],) if ctype != CompoundType.Scalar else ())
    return _tag(_torch().zeros(shape, dtype=_torch_dtype(_PT_INFO[PType(ptype)][1]), device=device), ctype)


_INT_RANGE = {np.int32: (-2 ** 31, 2 ** 31 - 1), np.int64: (-2 ** 63, 2 ** 63 - 1)}


def _parse_values(vals, np_dtype):
    """each value read through its PType's text type (types_traits.hpp:118-243): uint8 /
    uint16 as unsigned int (the C++ stream wraps a minus sign modulo 2^32), then cast; int32 /
    int64 as themselves (out of range is a read failure); float32 / float64 as floats"""
    try:
        if np_dtype in (np.uint8, np.uint16):
            ints = []
            for v in vals:
                x = int(v, 10)
                if abs(x) > 2 ** 32 - 1:
                    raise ValueError(v)
                ints.append(x % 2 ** 32)
            return np.array(ints, dtype=np.uint64).astype(np_dtype)
        if np_dtype in (np.int32, np.int64):
            lo, hi = _INT_RANGE[np_dtype]
            ints = [int(v, 10) for v in vals]
            if any(x < lo or x > hi for x in ints):
                raise ValueError("out of range")
            return np.array(ints, dtype=np_dtype)
        return np.array([float(v) for v in vals], dtype=np.float64).astype(np_dtype)
    except ValueError as e:
        raise PoutreError(-1, f"ImageFromString: unable to read current pixel ({e})") from None


def from_string(text, device="cuda"):
    """poutre::ImageFromString (python/base/image_interface.cpp:29; image_interface.cpp:616-690):
    "<Scalar|3Planes|4Planes> <PType> <rank> <dims...> <values...>", rank 1..4, the values of a
    compound pixel consecutive; trailing values are ignored."""
    tok = text.split()
    if len(tok) < 3:
        raise PoutreError(-1, "ImageFromString: unable to read the header")
    ctypes_by_name = {"Scalar": CompoundType.Scalar, "3Planes": CompoundType["3Planes"],
                      "4Planes": CompoundType["4Planes"]}
    if tok[0] not in ctypes_by_name:
        raise PoutreError(-1, "Unable to read CompoundType from stream")
    ctype = ctypes_by_name[tok[0]]
    names = {v[0]: k for k, v in _PT_INFO.items()}
    if tok[1] not in names:
        raise PoutreError(-2, "Unable to read PType from stream")
    pt = names[tok[1]]
    try:
        rank = int(tok[2])
    except ValueError:
        raise PoutreError(-1, "ImageFromString: unable to read NumDims") from None
    if rank < 1:
        raise PoutreError(-4, "ImageFromString: wrong value from NumDims")
    if rank > 4:
        raise PoutreError(-4, "Unsupported number of dims")
    dims = [int(v) for v in tok[3:3 + rank]]
    if len(dims) != rank:
        raise PoutreError(-1, "ImageFromString: unable to read dim")
    planes = _PLANES[ctype]
    n = int(np.prod(dims)) * planes
    vals = tok[3 + rank:3 + rank + n]
    if len(vals) < n:
        raise PoutreError(-1, "ImageFromString: unable to read current pixel")
    arr = _parse_values(vals, _PT_INFO[pt][1]).reshape(dims + ([planes] if planes > 1 else []))
    return _tag(_torch().from_numpy(arr).to(device), ctype)


def to_string(img, ctype=None):
    """poutre::ImageToString (python/base/image_interface.cpp:30; image_interface.cpp:879-909):
    the from_string format with single spaces, integers as integers, floats in the C++ stream
    default (6 significant digits).  `ctype` defaults to the image's own (Scalar for a plain
    tensor); for 3Planes / 4Planes the last axis holds the planes."""
    t = _img(img)
    ctype = CompoundType(ctype if ctype is not None else image_ctype(img))
    pt = PType(_ptype(t))
    shape = list(t.shape)
    if ctype != CompoundType.Scalar:
        if not shape or shape[-1] != _PLANES[ctype]:
            raise PoutreError(-1, f"to_string: a {ctype.name} image needs a last axis of {_PLANES[ctype]} planes")
        shape = shape[:-1]
    a = t.detach().cpu().numpy().reshape(-1)
    if a.dtype.kind == "f":
        vals = ["%g" % float(v) for v in a]
    else:
        vals = [str(int(v)) for v in a]
    return " ".join([ctype.name, _PT_INFO[pt][0], str(len(shape))] + [str(d) for d in shape] + vals)


def convert_into(img, out):
    """poutre::ConvertInto (copy_convert.hpp:52; pypoutre.convert_into): out = img converted
    pixel by pixel to out's type (a C cast), on the device."""
    di, do = _desc(img), _desc(out)
    _call("pm_convert_into", ctypes.byref(di), ctypes.byref(do), _stream(img))
    return out


def copy_into(img, out):
    """poutre::CopyInto (copy_convert.hpp:49): out = img (same shape and type)."""
    a, b = _img(img), _img(out)
    if tuple(a.shape) != tuple(b.shape):
        raise PoutreError(-1, "CopyInto images have not compatible sizes")
    if a.dtype != b.dtype:
        raise PoutreError(-2, "CopyInto images must have compatible types")
    return convert_into(img, out)


def clone(img):
    """poutre::Clone (copy_convert.hpp:31): a new image equal to img."""
    return convert_into(img, _torch().empty_like(_img(img)))


def clone_geometry(img):
    """poutre::CloneGeometry (copy_convert.hpp:34): a zero-filled image of img's shape and type."""
    return _torch().zeros_like(_img(img))


def convert_geometry(img, ctype, ptype):
    """poutre::ConvertGeometry (copy_convert.hpp:42): a zero-filled image of img's shape in ptype."""
    t = _img(img)
    return factory_image(tuple(t.shape), ctype, ptype, device=t.device)


def convert(img, ctype, ptype):
    """poutre::Convert (copy_convert.hpp:38): img converted pixel by pixel to ptype."""
    return convert_into(img, convert_geometry(img, ctype, ptype))


def as_sizes_compatible(a, b):
    """poutre::AsSizesCompatible: same rank and shape."""
    return tuple(_img(a).shape) == tuple(_img(b).shape)


def as_types_compatible(a, b):
    """poutre::AsTypesCompatible: same pixel type (every image here is Scalar)."""
    return _img(a).dtype == _img(b).dtype


def are_images_different(a, b):
    """poutre::AreImagesDifferent: the two images are different memory."""
    return _img(a).data_ptr() != _img(b).data_ptr()


# ---------------------------------------------------------------- low-level morphology
def _erodil(prefix, img, se, it, out):
    out = _out(img, out)
    di, do = _desc(img), _desc(out)
    if isinstance(se, static_compound_se):
        _call(f"pm_{prefix}_compound", ctypes.byref(di), int(se), int(it), ctypes.byref(do), _stream(img))
    elif isinstance(se, (int, static_nl_se)) and not isinstance(se, bool):
        _call(f"pm_{prefix}_static", ctypes.byref(di), int(se), int(it), ctypes.byref(do), _stream(img))
    else:
        # runtime neighbour list: a sequence of offset tuples ({dy,dx} / {dz,dy,dx}), applied once
        offs = [int(v) for o in se for v in o]
        n = len(se)
        arr = (ctypes.c_int64 * max(1, len(offs)))(*offs)
        _call(f"pm_{prefix}_nl", ctypes.byref(di), arr, n, ctypes.byref(do), _stream(img))
    return out


def erode(img, se, iter=1, out=None):
    """poutre::Erode(img, Common_NL_SE | Compound_NL_SE, iter|size, out) (ero_dil.hpp:32-47).
    A list of offset tuples selects the runtime neighbour-list overload (applied once)."""
    return _erodil("erode", img, se, iter, out)


def dilate(img, se, iter=1, out=None):
    """poutre::Dilate(img, Common_NL_SE | Compound_NL_SE, iter|size, out) (ero_dil.hpp:32-47)."""
    return _erodil("dilate", img, se, iter, out)


def _line(name, img, r, out):
    out = _out(img, out)
    di, do = _desc(img), _desc(out)
    _call(name, ctypes.byref(di), int(r), ctypes.byref(do), _stream(img))
    return out


def erodeX(img, size_half_segment, out=None):
    """poutre::ErodeX (ero_dil_line.hpp:32)."""
    return _line("pm_erode_x", img, size_half_segment, out)


def dilateX(img, size_half_segment, out=None):
    """poutre::DilateX (ero_dil_line.hpp:37)."""
    return _line("pm_dilate_x", img, size_half_segment, out)


def erodeY(img, size_half_segment, out=None):
    """poutre::ErodeY (ero_dil_line.hpp:42); any H, W (the reference requires H == W)."""
    return _line("pm_erode_y", img, size_half_segment, out)


def dilateY(img, size_half_segment, out=None):
    """poutre::DilateY (ero_dil_line.hpp:47); any H, W (the reference requires H == W)."""
    return _line("pm_dilate_y", img, size_half_segment, out)


def _rect(name, img, ry, rx, out):
    out = _out(img, out)
    di, do = _desc(img), _desc(out)
    _call(name, ctypes.byref(di), int(ry), int(rx), ctypes.byref(do), _stream(img))
    return out


def erode_rect(img, ry, rx, out=None):
    """Erosion by a (2ry+1) x (2rx+1) rectangle (ErodeX then ErodeY)."""
    return _rect("pm_erode_rect", img, ry, rx, out)


def dilate_rect(img, ry, rx, out=None):
    """Dilation by a (2ry+1) x (2rx+1) rectangle (DilateX then DilateY)."""
    return _rect("pm_dilate_rect", img, ry, rx, out)


def _compose(name, img, se, it, out):
    out = _out(img, out)
    di, do = _desc(img), _desc(out)
    _call(name, ctypes.byref(di), int(se), int(it), ctypes.byref(do), _stream(img))
    return out


def opening(img, se, iter=1, out=None):
    """Dilate(Erode(img, se, iter), se, iter)."""
    return _compose("pm_opening", img, se, iter, out)


def closing(img, se, iter=1, out=None):
    """Erode(Dilate(img, se, iter), se, iter)."""
    return _compose("pm_closing", img, se, iter, out)


def gradient(img, se, iter=1, out=None):
    """Dilate(img, se, iter) - Erode(img, se, iter) (saturated)."""
    return _compose("pm_gradient", img, se, iter, out)


# ---------------------------------------------------------------- geodesy
def reconstruction(rec_type, marker, mask, se, out=None):
    """geo::Reconstruction(rec_type, marker, mask, se, out) (mreconstruct.hpp:37-41)."""
    out = _out(mask, out)
    dm, dk, do = _desc(marker), _desc(mask), _desc(out)
    _call("pm_reconstruct", int(rec_type), ctypes.byref(dm), ctypes.byref(dk), int(se), ctypes.byref(do),
          _stream(mask))
    return out


def _lev(name, ref, marker, se, out):
    out = _out(ref, out)
    dr, dm, do = _desc(ref), _desc(marker), _desc(out)
    _call(name, ctypes.byref(dr), ctypes.byref(dm), int(se), ctypes.byref(do), _stream(ref))
    return out


def low_leveling(ref, marker, se, out=None):
    """geo::low_leveling (leveling.hpp:37)."""
    return _lev("pm_low_leveling", ref, marker, se, out)


def high_leveling(ref, marker, se, out=None):
    """geo::high_leveling (leveling.hpp:45)."""
    return _lev("pm_high_leveling", ref, marker, se, out)


def leveling(ref, marker, se, out=None):
    """geo::leveling (leveling.hpp:52)."""
    return _lev("pm_leveling", ref, marker, se, out)


def _scalar(v):
    """a scalar operand as the double the C ABI carries: integers beyond 2^53 would be rounded,
    so they are refused (the library then checks that the value fits the pixel type)"""
    if isinstance(v, (int, np.integer)) and abs(int(v)) > 2 ** 53:
        raise PoutreError(-6, f"scalar operand {int(v)} is not exactly representable (|value| > 2^53)")
    return float(v)


def _ext(name, img, value, se, out):
    out = _out(img, out)
    di, do = _desc(img), _desc(out)
    _call(name, ctypes.byref(di), _scalar(value), int(se), ctypes.byref(do), _stream(img))
    return out


def h_maxima(img, value, se, out=None):
    """geo::h_maxima (extrema.hpp:36)."""
    return _ext("pm_h_maxima", img, value, se, out)


def h_minima(img, value, se, out=None):
    """geo::h_minima (extrema.hpp:33)."""
    return _ext("pm_h_minima", img, value, se, out)


def h_convex(img, value, se, out=None):
    """geo::h_convex (extrema.hpp:43)."""
    return _ext("pm_h_convex", img, value, se, out)


def h_concave(img, value, se, out=None):
    """geo::h_concave (extrema.hpp:50)."""
    return _ext("pm_h_concave", img, value, se, out)


# ---------------------------------------------------------------- labelling
def _label_out(img, out):
    t, img = _torch(), _img(img)
    return t.empty(img.shape, dtype=t.int64, device=img.device) if out is None else out


def label_binary(img, se, out=None):
    """label::label_binary (label_binary.hpp:35): int64 labels of the connected components of
    the foreground, numbered by the raster position of their first pixel.  Returns
    (out, count); pypoutre returns the count and fills the caller's image."""
    out = _label_out(img, out)
    di, do = _desc(img), _desc(out)
    n = ctypes.c_int64(0)
    _call("pm_label_binary", ctypes.byref(di), int(se), ctypes.byref(do), ctypes.byref(n), _stream(img))
    return out, n.value


def label_flat_zones(img, se, out=None):
    """label::label_flat_zones (label_flat_zones.hpp:35): int64 labels of the flat zones.
    Returns (out, count)."""
    out = _label_out(img, out)
    di, do = _desc(img), _desc(out)
    n = ctypes.c_int64(0)
    _call("pm_label_flat_zones", ctypes.byref(di), int(se), ctypes.byref(do), ctypes.byref(n), _stream(img))
    return out, n.value


def label_minima(img, se, out=None):
    """label::label_minima (label_extrema.hpp:34): int64 labels of the regional minima."""
    out = _label_out(img, out)
    di, do = _desc(img), _desc(out)
    _call("pm_label_minima", ctypes.byref(di), int(se), ctypes.byref(do), _stream(img))
    return out


def label_maxima(img, se, out=None):
    """label::label_maxima (label_extrema.hpp:39): int64 labels of the regional maxima."""
    out = _label_out(img, out)
    di, do = _desc(img), _desc(out)
    _call("pm_label_maxima", ctypes.byref(di), int(se), ctypes.byref(do), _stream(img))
    return out


def dynamic_pseudo_opening(img, value, se, out=None):
    """geo::dynamic_pseudo_opening (extrema.hpp:57)."""
    return _ext("pm_dynamic_pseudo_opening", img, value, se, out)


def dynamic_pseudo_closing(img, value, se, out=None):
    """geo::dynamic_pseudo_closing (extrema.hpp:64)."""
    return _ext("pm_dynamic_pseudo_closing", img, value, se, out)


# ---------------------------------------------------------------- elementwise helpers
def _ew2(name, a, b, out):
    out = _out(a, out)
    da, db, do = _desc(a), _desc(b), _desc(out)
    _call(name, ctypes.byref(da), ctypes.byref(db), ctypes.byref(do), _stream(a))
    return out


def arith_sadd_image(a, b, out=None):
    return _ew2("pm_arith_sadd_image", a, b, out)


def arith_invert(img, out=None):
    """ArithInvertImage (arith.hpp:33): -img, integers modulo 2^bits; out may be img."""
    out = _out(img, out)
    di, do = _desc(img), _desc(out)
    _call("pm_arith_invert", ctypes.byref(di), ctypes.byref(do), _stream(img))
    return out


class CompOpType(enum.IntEnum):
    """poutre::CompOpType (include/poutre/base/types.hpp:443-453; python/base/types.cpp:25-31)."""
    Equal = 1 << 0
    Diff = 1 << 1
    Sup = 1 << 2
    SupEqual = 1 << 3
    Inf = 1 << 4
    InfEqual = 1 << 5


def _operand(x, keep):
    op = _capi.PmOperand()
    if isinstance(x, (int, float, np.integer, np.floating)) and not isinstance(x, bool):
        op.value = float(x)
        return op
    d = _desc(x)
    keep.append(d)
    op.image = ctypes.pointer(d)
    return op


def compare(img, op, comp, vtrue, vfalse, out=None):
    """poutre::CompareImage / pypoutre.compare, every overload (compare.hpp:31-84): out =
    (img OP comp) ? vtrue : vfalse pixel by pixel; comp, vtrue and vfalse are each an image
    or a number.  comp has img's type; vtrue / vfalse / out share a type (out's if given,
    else the first image among vtrue / vfalse, else img's).  out may be one of the inputs."""
    if out is None:
        like = next((v for v in (vtrue, vfalse) if not isinstance(v, (int, float, np.integer, np.floating))), img)
        out = _torch().empty(_img(img).shape, dtype=_img(like).dtype, device=_img(img).device)
    keep = []
    oc, ot, of = _operand(comp, keep), _operand(vtrue, keep), _operand(vfalse, keep)
    di, do = _desc(img), _desc(out)
    _call("pm_compare", ctypes.byref(di), int(op), ctypes.byref(oc), ctypes.byref(ot), ctypes.byref(of),
          ctypes.byref(do), _stream(img))
    return out


def arith_sup(a, b, out=None):
    return _ew2("pm_arith_sup", a, b, out)


def arith_inf(a, b, out=None):
    return _ew2("pm_arith_inf", a, b, out)


def arith_ssub_image(a, b, out=None):
    return _ew2("pm_arith_ssub_image", a, b, out)


def _ewc(name, a, c, out):
    out = _out(a, out)
    da, do = _desc(a), _desc(out)
    _call(name, ctypes.byref(da), _scalar(c), ctypes.byref(do), _stream(a))
    return out


def arith_ssub_const(a, c, out=None):
    return _ewc("pm_arith_ssub_const", a, c, out)


def arith_sadd_const(a, c, out=None):
    return _ewc("pm_arith_sadd_const", a, c, out)


# the names of pypoutre's pixel-processing module (python/pixel_processing/arith.cpp:13-19)
arith_invert_image = arith_invert
arith_inf_image = arith_inf
arith_sup_image = arith_sup
arith_saturated_add_image = arith_sadd_image
arith_saturated_sub_image = arith_ssub_image
arith_saturated_add_constant = arith_sadd_const
arith_saturated_sub_constant = arith_ssub_const


# ---------------------------------------------------------------- batched launches
def batch_dilate_open(images, radii, dil_outs=None, open_outs=None):
    """For every uint8 2D image i: (Dilate(img, square2d, r_i), Dilate(Erode(img, square2d, r_i), square2d, r_i)),
    all images of any sizes in one persistent launch (pm_batch_dilate_open).  Returns the two lists."""
    n = len(images)
    if len(radii) != n:
        raise PoutreError(-6, "batch: one radius per image")
    images = [_img(t) for t in images]
    dil_outs = [_out(im, None) for im in images] if dil_outs is None else [_img(t) for t in dil_outs]
    open_outs = [_out(im, None) for im in images] if open_outs is None else [_img(t) for t in open_outs]
    if n == 0:
        return dil_outs, open_outs
    di, dd, do = _descs(images), _descs(dil_outs), _descs(open_outs)
    rr = np.asarray([int(r) for r in radii], dtype=np.int32)
    _call("pm_batch_dilate_open", n, _rec_ptr(di), rr.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _rec_ptr(dd),
          _rec_ptr(do), _stream(images[0]))
    return dil_outs, open_outs


def _packed(flat, shapes, what):
    """pm_image records of 2D images stored back to back (row-major) in one 1D device tensor."""
    if not flat.is_cuda or flat.dim() != 1 or not flat.is_contiguous():
        raise PoutreError(-6, f"{what}: packed images must be one contiguous 1D device tensor")
    shp = np.asarray(shapes, dtype=np.int64).reshape(-1, 2)
    if (shp < 0).any():
        raise PoutreError(-6, f"{what}: negative image size")
    sizes = shp[:, 0] * shp[:, 1]
    offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    if int(sizes.sum()) > flat.numel():
        raise PoutreError(-6, f"{what}: packed buffer holds {flat.numel()} elements, the images need {int(sizes.sum())}")
    rec = np.zeros(len(shp), dtype=_PM_IMAGE_NP)
    rec["data"] = np.uint64(flat.data_ptr()) + (offsets * flat.element_size()).astype(np.uint64)
    rec["ptype"] = _ptype(flat)
    rec["rank"] = 2
    rec["shape"][:, :2] = shp
    return rec


def batch_dilate_open_packed(flat, shapes, radii, dil_flat=None, open_flat=None):
    """batch_dilate_open for images stored back to back in one 1D uint8 device tensor `flat`
    (image i: shapes[i] = (H_i, W_i), row-major, at offset sum_{j<i} H_j W_j).  The outputs
    use the same packing.  The descriptors are computed with numpy in one pass, so the host
    cost of a call does not grow with per-image Python work.  Returns (dil_flat, open_flat)."""
    n = len(shapes)
    if len(radii) != n:
        raise PoutreError(-6, "batch: one radius per image")
    dil_flat = _out(flat, dil_flat)
    open_flat = _out(flat, open_flat)
    if n == 0:
        return dil_flat, open_flat
    di, dd, do = _packed(flat, shapes, "batch"), _packed(dil_flat, shapes, "batch"), _packed(open_flat, shapes, "batch")
    rr = np.ascontiguousarray(radii, dtype=np.int32)
    _call("pm_batch_dilate_open", n, _rec_ptr(di), rr.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _rec_ptr(dd),
          _rec_ptr(do), _stream(flat))
    return dil_flat, open_flat


def batch_reconstruction_packed(rec_type, marker_flat, mask_flat, shapes, se, out_flat=None):
    """batch_reconstruction for images packed as in batch_dilate_open_packed."""
    n = len(shapes)
    out_flat = _out(mask_flat, out_flat)
    if n == 0:
        return out_flat
    dm, dk, do = (_packed(marker_flat, shapes, "batch"), _packed(mask_flat, shapes, "batch"),
                  _packed(out_flat, shapes, "batch"))
    _call("pm_batch_reconstruct", n, int(rec_type), _rec_ptr(dm), _rec_ptr(dk), int(se), _rec_ptr(do),
          _stream(mask_flat))
    return out_flat


def batch_reconstruction(rec_type, markers, masks, se, outs=None):
    """n uint8 2D reconstructions (geo::Reconstruction semantics) sharing one sequence of global rounds."""
    n = len(markers)
    markers, masks = [_img(t) for t in markers], [_img(t) for t in masks]
    outs = [_out(m, None) for m in masks] if outs is None else [_img(t) for t in outs]
    if n == 0:
        return outs
    dm, dk, do = _descs(markers), _descs(masks), _descs(outs)
    _call("pm_batch_reconstruct", n, int(rec_type), _rec_ptr(dm), _rec_ptr(dk), int(se), _rec_ptr(do),
          _stream(masks[0]))
    return outs


class BatchPlan:
    """batch_dilate_open captured once into a HIP graph (pm_batch_plan_create): its scratch and
    descriptor table are allocated at creation, and run() is one graph launch on the current
    stream with no allocation and no per-image host work.  The plan keeps references to its
    tensors; their contents may change between runs, their storage may not.  A plan owns one
    scratch area (eroded tiles, queue counters): runs of the same plan must be serialised
    (one stream, or an event between streams); runs that overlap race on that area."""

    def __init__(self, images, radii, dil_outs=None, open_outs=None):
        n = len(images)
        if len(radii) != n:
            raise PoutreError(-6, "batch: one radius per image")
        images = [_img(t) for t in images]
        self.dil_outs = [_out(im, None) for im in images] if dil_outs is None else [_img(t) for t in dil_outs]
        self.open_outs = [_out(im, None) for im in images] if open_outs is None else [_img(t) for t in open_outs]
        # the tensors (DLPack producers converted to zero-copy views) stay alive with the plan
        self._keep = (images, self.dil_outs, self.open_outs)
        recs = (_descs(images), _descs(self.dil_outs), _descs(self.open_outs))
        self._create(n, recs, radii, images[0] if n else None)

    @classmethod
    def packed(cls, flat, shapes, radii, dil_flat=None, open_flat=None):
        """Plan over images packed back to back in one 1D uint8 tensor (see batch_dilate_open_packed)."""
        self = cls.__new__(cls)
        dil_flat = _out(flat, dil_flat)
        open_flat = _out(flat, open_flat)
        self.dil_outs, self.open_outs = dil_flat, open_flat
        self._keep = (flat, dil_flat, open_flat)
        recs = (_packed(flat, shapes, "batch"), _packed(dil_flat, shapes, "batch"), _packed(open_flat, shapes, "batch"))
        self._create(len(shapes), recs, radii, flat)
        return self

    def _create(self, n, recs, radii, ref):
        self._plan = None
        self._device = ref.device if ref is not None else None
        rr = np.ascontiguousarray([int(r) for r in radii], dtype=np.int32)
        h = ctypes.c_void_p()
        di, dd, do = recs
        _call("pm_batch_plan_create", n, _rec_ptr(di), rr.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _rec_ptr(dd),
              _rec_ptr(do), ctypes.byref(h))
        self._plan = h

    def run(self):
        if self._plan is None:
            raise PoutreError(-6, "batch plan: closed")
        stream = ctypes.c_void_p(_torch().cuda.current_stream(self._device).cuda_stream)
        _call("pm_batch_plan_launch", self._plan, stream)
        return self.dil_outs, self.open_outs

    def close(self):
        if getattr(self, "_plan", None) is not None:
            _capi.lib().pm_batch_plan_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass


# ---------------------------------------------------------------- autotuner
def tuning():
    """The fused uint8 square kernel's tile width per radius 1..32 on the current device:
    {"aligned": [...], "any": [...]} (rows 16-byte aligned, rows of any alignment)."""
    L = _capi.lib()
    out = {}
    for key, al in (("aligned", 1), ("any", 0)):
        vals = []
        for r in range(1, 33):
            v = L.pm_tuning_get(al, r)
            if v < 0:
                _capi.check(v)
            vals.append(v)
        out[key] = vals
    return out


def autotune(cache=None):
    """Pick the fused square kernel's tile width for every radius by timing (pm_autotune).
    With `cache` (a JSON path): load the table from it when it was written for this GPU's
    name, otherwise measure and write it.  Returns the table (see tuning())."""
    import json
    import os
    torch = _torch()
    name = torch.cuda.get_device_properties(torch.cuda.current_device()).name
    if cache is not None:
        cache = os.path.expanduser(cache)
    if cache is not None and os.path.exists(cache):
        with open(cache) as f:
            data = json.load(f)
        if data.get("device") == name and data.get("version") == version():
            for key, al in (("aligned", 1), ("any", 0)):
                for r, w in enumerate(data["tile_width"][key], start=1):
                    _call("pm_tuning_set", al, r, int(w))
            return tuning()
    _call("pm_autotune", ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    table = tuning()
    if cache is not None:
        d = os.path.dirname(os.path.abspath(cache))
        os.makedirs(d, exist_ok=True)
        with open(cache, "w") as f:
            json.dump({"device": name, "version": version(), "tile_width": table}, f)
    return table
