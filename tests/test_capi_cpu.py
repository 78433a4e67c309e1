"""CPU-side checks of the C ABI library: it loads without a GPU and exports every
function include/poutre_amd.h declares; the Python binding knows a signature for each."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = []
    for hdr in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not hdr.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", hdr)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(pm_\w+)\s*\(", src, flags=re.M)
    return names


def test_header_declares_entry_points():
    names = declared()
    assert "pm_dilate_static" in names and "pm_reconstruct" in names and "pm_last_error" in names
    assert len(names) == len(set(names))


def test_library_exports_every_declared_symbol():
    from poutre2_amd import _capi
    if not os.path.exists(_capi.LIB_PATH):
        pytest.skip("libpoutre_amd.so not built (run python -m poutre2_amd.build)")
    lib = ctypes.CDLL(_capi.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_compute_entry_points():
    from poutre2_amd import _capi
    compute = [n for n in declared() if n not in ("pm_last_error", "pm_version")]
    assert sorted(set(compute) - set(_capi.SIGNATURES)) == []


def test_version_and_error_without_gpu():
    from poutre2_amd import _capi
    if not os.path.exists(_capi.LIB_PATH):
        pytest.skip("library not built")
    L = _capi.lib()
    assert b"gfx950" in L.pm_version()
    assert L.pm_last_error() == b""
    # argument checks run before any device work: a size mismatch fails on the host
    a, b = _capi.PmImage(), _capi.PmImage()
    a.data, a.ptype, a.rank = 1, _capi.PTYPE_U8, 2
    b.data, b.ptype, b.rank = 2, _capi.PTYPE_U8, 2
    a.shape[0], a.shape[1], b.shape[0], b.shape[1] = 4, 4, 4, 5
    rc = L.pm_dilate_static(ctypes.byref(a), 2, 1, ctypes.byref(b), None)
    assert rc == -1 and b"compatible sizes" in L.pm_last_error()
    b.shape[1] = 4
    b.data = 1
    assert L.pm_dilate_static(ctypes.byref(a), 2, 1, ctypes.byref(b), None) == -3
    b.data = 2
    assert L.pm_dilate_static(ctypes.byref(a), 2, -1, ctypes.byref(b), None) == -6
    assert L.pm_dilate_static(ctypes.byref(a), 10, 1, ctypes.byref(b), None) == -4


def test_batch_descriptor_record_matches_pm_image():
    """The numpy record array used for batched calls has pm_image's exact layout."""
    import ctypes
    import poutre2_amd as pm
    from poutre2_amd import _capi
    assert pm._PM_IMAGE_NP.itemsize == ctypes.sizeof(_capi.PmImage)
    for f in ("data", "ptype", "rank", "shape"):
        assert pm._PM_IMAGE_NP.fields[f][1] == getattr(_capi.PmImage, f).offset
