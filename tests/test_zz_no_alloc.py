"""pm_reconstruct_noalloc (the scratch-free uint8 reconstruction) against pm_reconstruct, both
through the C ABI, bit for bit: two calls in a row on one GPU, the second on a smaller image
(the shared scratch keeps the first call's contents, so the second proves it is cleared),
both reconstruction types, square and cross SEs; plus its argument errors (host-side checks,
no kernel launched)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
pm = pytest.importorskip("poutre2_amd")
from poutre2_amd import _capi, synth  # noqa: E402

DEV = "cuda:0"


def _run(name, rec, marker, mask, se, out):
    dm, dk, do = pm._desc(marker), pm._desc(mask), pm._desc(out)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return getattr(_capi.lib(), name)(int(rec), ctypes.byref(dm), ctypes.byref(dk), int(se), ctypes.byref(do), stream)


def _pair(H, W, seed, rec):
    mask = synth.blurred_noise_mask(H, W, seed, sigma=4.0, radius=12)
    if rec == pm.RecType.dilate:
        marker = synth.hmax_marker(mask, 20)
    else:
        marker = np.minimum(mask.astype(np.int32) + 20, 255).astype(np.uint8)
    return torch.from_numpy(marker).to(DEV), torch.from_numpy(mask).to(DEV)


@pytest.mark.parametrize("rec", [pm.RecType.dilate, pm.RecType.erode])
@pytest.mark.parametrize("se", [pm.static_nl_se.square2d, pm.static_nl_se.cross2d])
def test_noalloc_matches_reconstruct_on_two_calls(rec, se):
    for (H, W), seed in (((700, 913), 3), ((130, 77), 4)):
        print(f"noalloc {rec.name} se {int(se)} {H}x{W}", flush=True)
        marker, mask = _pair(H, W, seed, rec)
        want = torch.empty_like(mask)
        got = torch.full_like(mask, 7)
        _capi.check(_run("pm_reconstruct", rec, marker, mask, se, want))
        _capi.check(_run("pm_reconstruct_noalloc", rec, marker, mask, se, got))
        torch.cuda.synchronize()
        assert torch.equal(got, want)


def test_noalloc_segment_1d():
    marker, mask = _pair(1, 5000, 5, pm.RecType.dilate)
    marker, mask = marker.reshape(-1), mask.reshape(-1)
    want, got = torch.empty_like(mask), torch.empty_like(mask)
    _capi.check(_run("pm_reconstruct", pm.RecType.dilate, marker, mask, pm.static_nl_se.segx1d, want))
    _capi.check(_run("pm_reconstruct_noalloc", pm.RecType.dilate, marker, mask, pm.static_nl_se.segx1d, got))
    torch.cuda.synchronize()
    assert torch.equal(got, want)


def test_noalloc_rejects_what_it_cannot_serve():
    # more tiles than the scratch holds: one row of 62 * 196609 pixels (checked before any launch)
    n = 62 * 196609
    a = torch.zeros((1, n), dtype=torch.uint8, device=DEV)
    b, c = torch.zeros_like(a), torch.empty_like(a)
    assert _run("pm_reconstruct_noalloc", pm.RecType.dilate, a, b, pm.static_nl_se.square2d, c) == -6  # PM_E_ARG
    assert b"limit" in _capi.lib().pm_last_error()
    # other pixel types take the allocating path only
    x = torch.zeros((8, 8), dtype=torch.int32, device=DEV)
    y, z = torch.zeros_like(x), torch.empty_like(x)
    assert _run("pm_reconstruct_noalloc", pm.RecType.dilate, x, y, pm.static_nl_se.square2d, z) == -8  # PM_E_UNSUPPORTED
    # 3D SEs too
    v = torch.zeros((4, 5, 6), dtype=torch.uint8, device=DEV)
    w, o = torch.zeros_like(v), torch.empty_like(v)
    assert _run("pm_reconstruct_noalloc", pm.RecType.dilate, v, w, pm.static_nl_se.cross3d, o) == -8  # PM_E_UNSUPPORTED
