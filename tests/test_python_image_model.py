"""The image-model names of the reference's Python module (python/base/image_interface.cpp:13-31,
python/base/types.cpp:10-23) on host tensors: factory_image, from_string / to_string (the
fixture format of image_interface.cpp:407-909), the compatibility predicates, DLPack input,
and the SE object of python/structuring_element/se_interface.cpp.  No GPU: these functions run
on the host."""
import numpy as np
import pytest

import poutre2_amd as pm

from . import kats

torch = pytest.importorskip("torch")


def _as_string(d):
    return " ".join(["Scalar", d["ptype"], str(len(d["shape"]))] + [str(v) for v in d["shape"]] +
                    [repr(v) if isinstance(v, float) else str(v) for v in d["values"]])


_IMAGES = [d for c in kats.cases(lambda c: c["expected_complete"]) for d in c["inputs"] + [c["expected"]]
           if d["n_listed"] == int(np.prod(d["shape"]))]


@pytest.mark.parametrize("i", range(len(_IMAGES)))
def test_from_string_matches_kat_images(i):
    d = _IMAGES[i]
    t = pm.from_string(_as_string(d), device="cpu")
    np.testing.assert_array_equal(t.numpy(), kats.image(d))
    assert t.dtype == torch.from_numpy(kats.image(d)).dtype


def test_to_string_round_trip():
    for d in _IMAGES:
        t = pm.from_string(_as_string(d), device="cpu")
        s = pm.to_string(t)
        assert s.split()[:3 + len(d["shape"])] == _as_string(d).split()[:3 + len(d["shape"])]
        np.testing.assert_array_equal(pm.from_string(s, device="cpu").numpy(), kats.image(d))


def test_to_string_float_format():
    # what a C++ std::ostream prints for these values with its default precision (6
    # significant digits), for double and for float
    vals = [1.5, -2, 1e-10, 1234567, 0.1, 100000, 999999.5, 1e16, -0.0, 3.14159265, 2.5e-5, 123456.7, 65504.0]
    exp = "1.5 -2 1e-10 1.23457e+06 0.1 100000 1e+06 1e+16 -0 3.14159 2.5e-05 123457 65504"
    for dt, name in ((torch.float64, "D64"), (torch.float32, "F32")):
        s = pm.to_string(torch.tensor(vals, dtype=dt))
        assert s == f"Scalar {name} 1 {len(vals)} {exp}"


def test_from_string_conversions_and_errors():
    # values go through the PType's text type: "1.9" is not an int (the reference's stream read
    # stops at '.', and the next value read fails), uint8 reads an unsigned int and casts
    with pytest.raises(pm.PoutreError):
        pm.from_string("Scalar GINT32 1 3 1.9 -1.9 7", device="cpu")
    assert pm.from_string("Scalar GUINT8 1 3 -1 300 255", device="cpu").tolist() == [255, 44, 255]
    with pytest.raises(pm.PoutreError):
        pm.from_string("Scalar GINT32 1 1 2147483648", device="cpu")
    big = "Scalar GINT64 1 3 9223372036854775807 -9223372036854775808 -1"
    assert pm.to_string(pm.from_string(big, device="cpu")) == big
    # trailing values are ignored
    assert pm.from_string("Scalar GUINT8 1 2 3 4 5", device="cpu").tolist() == [3, 4]
    with pytest.raises(pm.PoutreError):
        pm.from_string("Vector GUINT8 1 2 3 4", device="cpu")
    with pytest.raises(pm.PoutreError):
        pm.from_string("Scalar GFOO 1 2 3 4", device="cpu")
    with pytest.raises(pm.PoutreError):
        pm.from_string("Scalar GUINT8 2 2 2 1 2 3", device="cpu")
    with pytest.raises(pm.PoutreError):
        pm.from_string("Scalar GUINT8 0", device="cpu")
    # rank 4 images exist (the operators take ranks 1..3)
    r4 = pm.from_string("Scalar GUINT8 4 1 2 1 2 0 255 7 8", device="cpu")
    assert r4.shape == (1, 2, 1, 2) and pm.to_string(r4) == "Scalar GUINT8 4 1 2 1 2 0 255 7 8"


def _iostring_kats():
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "iostring_kats.json")) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("c", _iostring_kats(), ids=lambda c: c["name"])
def test_iostring_kats(c):
    """test/base/image_iostring.cpp:28-242 through the Python names (from_string / to_string)"""
    if c["kind"] == "read":
        t = pm.from_string(c["text"], device="cpu")
        assert pm.image_ctype(t).name == c["compound"]
        planes = {"Scalar": 1, "3Planes": 3, "4Planes": 4}[c["compound"]]
        shape = list(t.shape[:-1]) if planes > 1 else list(t.shape)
        assert shape == c["shape"] and len(shape) == c["rank"]
        assert t.reshape(-1).tolist() == c["values"]
    else:
        dt = {"GUINT8": torch.uint8, "F32": torch.float32, "D64": torch.float64}[c["ptype"]]
        planes = {"Scalar": 1, "3Planes": 3, "4Planes": 4}[c["compound"]]
        shape = tuple(c["shape"]) + ((planes,) if planes > 1 else ())
        t = torch.empty(shape, dtype=dt)
        t[...] = torch.tensor(c["fill"], dtype=dt)
        assert pm.to_string(t, pm.CompoundType[c["compound"]]) == c["expected"]


def test_factory_image_and_predicates():
    a = pm.factory_image([3, 4], pm.CompoundType.Scalar, pm.PType.GrayINT32, device="cpu")
    assert a.shape == (3, 4) and a.dtype == torch.int32 and int(a.abs().sum()) == 0
    b = pm.factory_image((3, 4), pm.CompoundType.Scalar, pm.PType.F32, device="cpu")
    c = pm.factory_image((4, 3), pm.CompoundType.Scalar, pm.PType.GrayINT32, device="cpu")
    assert pm.as_sizes_compatible(a, b) and not pm.as_sizes_compatible(a, c)
    assert pm.as_types_compatible(a, c) and not pm.as_types_compatible(a, b)
    assert pm.are_images_different(a, c) and not pm.are_images_different(a, a)
    c3 = pm.factory_image((3, 4), pm.CompoundType["3Planes"], pm.PType.GrayUINT8, device="cpu")
    assert c3.shape == (3, 4, 3) and pm.image_ctype(c3) == pm.CompoundType["3Planes"]
    assert pm.to_string(c3).startswith("3Planes GUINT8 2 3 4 0 0 0")
    with pytest.raises(pm.PoutreError):  # the operators take Scalar images only
        pm.dilate(c3, pm.static_nl_se.square2d, 1)
    with pytest.raises(pm.PoutreError):
        pm.factory_image((1, 2, 3, 4, 5), pm.CompoundType.Scalar, pm.PType.GrayUINT8, device="cpu")
    assert [int(p) for p in pm.PType] == [4, 8, 16, 32, 64, 128]


def test_dlpack_producers_are_images():
    # a numpy array (a DLPack producer) is used in place: same memory, no copy
    arr = np.arange(6, dtype=np.uint8).reshape(2, 3)
    assert pm.to_string(arr) == "Scalar GUINT8 2 2 3 0 1 2 3 4 5"
    assert pm.as_sizes_compatible(arr, torch.zeros(2, 3))
    assert not pm.are_images_different(arr, torch.from_dlpack(arr))
    with pytest.raises(pm.PoutreError):
        pm.to_string([1, 2, 3])


def test_se_object():
    """pypoutre's SE methods (python/structuring_element/se_interface.cpp:15-23) on a neighbour
    list; unordered equality compares sets (neighbor_list_se_t.hpp:177-195)"""
    # the centre is the first offset, as in every predefined SE (predefined_nl_se.cpp:19-91)
    sq = pm.SE([(0, 0)] + [(dy, dx) for dy in (-1, 0, 1) for dx in (-1, 0, 1) if (dy, dx) != (0, 0)])
    assert sq.size() == 9 and sq.type() == pm.se_type.neighbor_list
    rc = sq.remove_center()
    assert rc.size() == 8 and (0, 0) not in list(rc)
    assert sq.is_equal_unordered(sq.transpose()) and not sq.is_equal(sq.transpose())
    seg = pm.SE([(0, 0), (0, -1), (0, 1), (0, 2)])
    assert list(seg.transpose()) == [(0, 0), (0, 1), (0, -1), (0, -2)]
    # mirrored about the first offset when that is not the origin (neighbor_list_se_t.hpp:96-107)
    assert list(pm.SE([(1, 1), (0, 1)]).transpose()) == [(1, 1), (2, 1)]
    assert not seg.is_equal_unordered(seg.transpose())
    assert pm.SE([(0, 0), (0, 0), (-1, 0)]).is_equal_unordered(pm.SE([(-1, 0), (0, 0)]))
    assert seg.clone().is_equal(seg) and not seg.is_equal([(0, -1), (0, 0), (0, 1), (0, 2)])
    with pytest.raises(pm.PoutreError):
        pm.SE([(0, 0), (0, 0, 1)])
