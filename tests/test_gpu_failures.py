"""Error paths the reference's contract (include/poutre/base/config.hpp:203-215: a failing call
throws, nothing half-done is reported as success) asks for, forced from the host so that every
kernel still exits normally:
  * a reconstruction that does not converge within the round cap returns PM_E_NOCONV
    (pm_debug_set_max_rounds), for the uint8 tile kernel, the generic tile kernel and the 3D
    neighbour-list relaxation;
  * a failing scratch allocation returns PM_E_HIP (pm_debug_fail_allocations);
  * a scalar operand that is not a value of the image's pixel type is refused (PM_E_ARG)
    instead of wrapping.
After each forced failure the same call is repeated and must give the oracle's result."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
pm = pytest.importorskip("poutre2_amd")
from poutre2_amd import _capi, synth  # noqa: E402

DEV = "cuda:0"
PM_E_ARG, PM_E_HIP, PM_E_NOCONV = -6, -7, -9


def g(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def h(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _serpentine(H, W, dtype):
    """a mask whose only path winds through the whole image (many tiles, many rounds)"""
    mask = np.zeros((H, W), dtype)
    for i, y in enumerate(range(0, H, 4)):
        mask[y, :] = 200
        x = W - 1 if i % 2 == 0 else 0
        mask[y:y + 4, x] = 200
    marker = np.zeros_like(mask)
    marker[0, 0] = 150
    return marker, mask


@pytest.mark.parametrize("dtype,se,shape", [(np.uint8, "SESquare2D", (300, 260)), (np.int32, "SESquare2D", (300, 260)),
                                            (np.uint8, "SECross3D", (6, 40, 40))])
def test_round_cap_returns_noconv(dtype, se, shape):
    if len(shape) == 2:
        marker, mask = _serpentine(*shape, dtype)
    else:
        marker, mask = _serpentine(shape[1], shape[2], dtype)
        marker = np.repeat(marker[None], shape[0], axis=0)
        mask = np.repeat(mask[None], shape[0], axis=0)
    want = oracle.reconstruct("dilate", marker, mask, se)
    _capi.check(_capi.lib().pm_debug_set_max_rounds(1))
    try:
        with pytest.raises(pm.PoutreError) as e:
            pm.reconstruction(pm.RecType.dilate, g(marker), g(mask), oracle.SE[se])
        assert e.value.code == PM_E_NOCONV and "converge" in str(e.value)
    finally:
        _capi.check(_capi.lib().pm_debug_set_max_rounds(0))
    got = h(pm.reconstruction(pm.RecType.dilate, g(marker), g(mask), oracle.SE[se]))
    np.testing.assert_array_equal(got, want)


def test_noalloc_reconstruction_also_honours_the_cap():
    import ctypes
    marker, mask = _serpentine(300, 260, np.uint8)
    tm, tk = g(marker), g(mask)  # the descriptors hold raw pointers: keep the tensors alive
    out = torch.empty_like(tk)
    dm, dk, do = pm._desc(tm), pm._desc(tk), pm._desc(out)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _capi.check(_capi.lib().pm_debug_set_max_rounds(1))
    try:
        rc = _capi.lib().pm_reconstruct_noalloc(1, ctypes.byref(dm), ctypes.byref(dk), 2, ctypes.byref(do), s)
        assert rc == PM_E_NOCONV
    finally:
        _capi.check(_capi.lib().pm_debug_set_max_rounds(0))


@pytest.mark.parametrize("call", ["leveling", "opening_int32", "h_maxima", "label"])
def test_forced_allocation_failure_returns_hip_error(call):
    mask = synth.blurred_noise_mask(200, 230, 3, sigma=3.0, radius=10)
    f = g(mask)
    run = {
        "leveling": lambda: pm.leveling(f, pm.opening(f, pm.static_nl_se.square2d, 3), pm.static_nl_se.square2d),
        "opening_int32": lambda: pm.opening(f.int(), pm.static_nl_se.cross2d, 2),
        "h_maxima": lambda: pm.h_maxima(f, 20, pm.static_nl_se.square2d),
        "label": lambda: pm.label_binary((f > 128).to(torch.uint8), pm.static_nl_se.square2d)[0],
    }[call]
    want = h(run())
    _capi.check(_capi.lib().pm_debug_fail_allocations(1))
    try:
        with pytest.raises(pm.PoutreError) as e:
            run()
        assert e.value.code == PM_E_HIP and "allocation" in str(e.value)
    finally:
        _capi.check(_capi.lib().pm_debug_fail_allocations(0))
    np.testing.assert_array_equal(h(run()), want)


def test_scalar_operands_must_fit_the_pixel_type():
    img = g(synth.uniform_u8((40, 50), 1))
    for bad in (256, -1, 2.5):
        with pytest.raises(pm.PoutreError) as e:
            pm.h_maxima(img, bad, pm.static_nl_se.square2d)
        assert e.value.code == PM_E_ARG
    with pytest.raises(pm.PoutreError):
        pm.arith_sadd_const(img, 300)
    with pytest.raises(pm.PoutreError):
        pm.h_minima(img.int(), 0.5, pm.static_nl_se.square2d)
    # in range: accepted, and equal to the oracle
    np.testing.assert_array_equal(h(pm.h_maxima(img, 255, pm.static_nl_se.square2d)),
                                  oracle.h_maxima(h(img), 255, "SESquare2D"))
    f = g(np.linspace(-3, 3, 40 * 50, dtype=np.float32).reshape(40, 50))
    with pytest.raises(pm.PoutreError):
        pm.h_maxima(f, 1e39, pm.static_nl_se.square2d)  # beyond float32
    assert h(pm.h_maxima(f, 0.25, pm.static_nl_se.square2d)).shape == (40, 50)
