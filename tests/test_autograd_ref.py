"""CPU: the numpy tap-by-tap references (tests/autograd_ref.py) used for the full-size
autograd digests agree with the torch CPU operator (F.unfold + first-occurrence max / min)
and with torch CPU autograd of it, including plateaus (ties) and image borders."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
from tests import autograd_ref  # noqa: E402


def _x(shape, seed, plateau=True):
    g = np.random.default_rng(seed)
    x = g.standard_normal(shape).astype(np.float32)
    if plateau:
        x[..., 2:6, 3:9] = np.float32(0.25)  # ties resolved by the raster rule
        x[..., -1, :] = np.float32(-0.5)
    return x


@pytest.mark.parametrize("k,with_w,dilate", [(3, False, True), (3, False, False), (7, False, True),
                                             (5, True, True), (5, True, False), (9, True, True)])
def test_numpy_reference_matches_torch_autograd(k, with_w, dilate):
    x = _x((2, 3, 13, 17), k)
    w = (0.1 * np.random.default_rng(100 + k).standard_normal((k, k))).astype(np.float32) if with_w else None
    gy = np.random.default_rng(7).integers(-3, 4, size=x.shape).astype(np.float32)
    y, tap = autograd_ref.forward_with_taps(x, k=k, w=w, dilate=dilate)
    xt = torch.from_numpy(x).requires_grad_(True)
    wt = torch.from_numpy(w).requires_grad_(True) if with_w else None
    yt = autograd_ref.reference_cpu(xt, k=k, w=wt, dilate=dilate)
    np.testing.assert_array_equal(y, yt.detach().numpy())
    (yt * torch.from_numpy(gy)).sum().backward()
    gx, gw = autograd_ref.backward_from_taps(tap, gy, k)
    np.testing.assert_array_equal(gx.astype(np.float32), xt.grad.numpy())
    if with_w:
        sign = 1.0 if dilate else -1.0
        np.testing.assert_allclose(sign * gw, wt.grad.numpy().astype(np.float64), rtol=0, atol=0)


def test_taps_point_at_the_first_extremum():
    x = np.zeros((1, 4, 4), np.float32)  # every window is a plateau: the first tap in raster order
    _, tap = autograd_ref.forward_with_taps(x, k=3)
    assert tap[0, 0, 0] == 4  # (dy, dx) = (0, 0): the taps above and left are outside the image
    assert tap[0, 1, 1] == 0 and tap[0, 3, 3] == 0
