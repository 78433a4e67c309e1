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
    # values are read as numbers, then converted to the pixel type (truncation toward zero)
    t = pm.from_string("Scalar GINT32 1 3 1.9 -1.9 7", device="cpu")
    assert t.tolist() == [1, -1, 7]
    # trailing values are ignored
    assert pm.from_string("Scalar GUINT8 1 2 3 4 5", device="cpu").tolist() == [3, 4]
    with pytest.raises(pm.PoutreError):
        pm.from_string("3Planes GUINT8 1 2 3 4", device="cpu")
    with pytest.raises(pm.PoutreError):
        pm.from_string("Scalar GFOO 1 2 3 4", device="cpu")
    with pytest.raises(pm.PoutreError):
        pm.from_string("Scalar GUINT8 2 2 2 1 2 3", device="cpu")
    with pytest.raises(pm.PoutreError):
        pm.from_string("Scalar GUINT8 4 1 1 1 1 0", device="cpu")


def test_factory_image_and_predicates():
    a = pm.factory_image([3, 4], pm.CompoundType.Scalar, pm.PType.GrayINT32, device="cpu")
    assert a.shape == (3, 4) and a.dtype == torch.int32 and int(a.abs().sum()) == 0
    b = pm.factory_image((3, 4), pm.CompoundType.Scalar, pm.PType.F32, device="cpu")
    c = pm.factory_image((4, 3), pm.CompoundType.Scalar, pm.PType.GrayINT32, device="cpu")
    assert pm.as_sizes_compatible(a, b) and not pm.as_sizes_compatible(a, c)
    assert pm.as_types_compatible(a, c) and not pm.as_types_compatible(a, b)
    assert pm.are_images_different(a, c) and not pm.are_images_different(a, a)
    with pytest.raises(pm.PoutreError):
        pm.factory_image((3, 4), pm.CompoundType["3Planes"], pm.PType.GrayUINT8, device="cpu")
    with pytest.raises(pm.PoutreError):
        pm.factory_image((1, 2, 3, 4), pm.CompoundType.Scalar, pm.PType.GrayUINT8, device="cpu")
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
    sq = pm.SE([(dy, dx) for dy in (-1, 0, 1) for dx in (-1, 0, 1)])
    assert sq.size() == 9 and sq.type() == pm.se_type.neighbor_list
    rc = sq.remove_center()
    assert rc.size() == 8 and (0, 0) not in list(rc)
    assert rc.is_equal_unordered(rc.transpose()) and not rc.is_equal(rc.transpose())
    seg = pm.SE([(0, -1), (0, 0), (0, 1), (0, 2)])
    assert list(seg.transpose()) == [(0, 1), (0, 0), (0, -1), (0, -2)]
    assert not seg.is_equal_unordered(seg.transpose())
    assert pm.SE([(0, 0), (0, 0), (-1, 0)]).is_equal_unordered(pm.SE([(-1, 0), (0, 0)]))
    assert seg.clone().is_equal(seg) and not seg.is_equal([(0, -1), (0, 0), (0, 1), (0, 2)])
    with pytest.raises(pm.PoutreError):
        pm.SE([(0, 0), (0, 0, 1)])
