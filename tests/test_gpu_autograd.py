"""Differentiable flat / max-plus dilation and erosion (poutre2_amd.autograd).

Forward: bit-exact against the torch CPU operator (F.unfold + max/min, first occurrence).
Backward: against torch CPU autograd of that operator (rtol 1e-5) and against central finite
differences (eps = 1e-3) on tie-free windows (rtol 1e-3: float32 forward, float64 FD sums).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
pm = pytest.importorskip("poutre2_amd")
from poutre2_amd import autograd as ag  # noqa: E402
from poutre2_amd import synth  # noqa: E402
from tests import autograd_ref  # noqa: E402

DEV = "cuda:0"


def _x(shape, seed):
    return torch.from_numpy(synth.normal_f32(shape, seed))


def _w(k, seed):
    g = np.random.default_rng(seed)
    return torch.from_numpy((0.1 * g.standard_normal((k, k))).astype(np.float32))


CASES = [(3, False), (7, False), (5, True), (3, True), (9, False)]


@pytest.mark.parametrize("k,nonflat", CASES)
@pytest.mark.parametrize("dilate", [True, False])
@pytest.mark.parametrize("shape", [(2, 3, 37, 45), (1, 1, 16, 64), (3, 5, 7)])
def test_forward_bit_exact(k, nonflat, dilate, shape):
    x = _x(shape, 7)
    w = _w(k, 3) if nonflat else None
    ref = autograd_ref.reference_cpu(x, k=k, w=w, dilate=dilate)
    f = ag.dilation2d if dilate else ag.erosion2d
    got = f(x.to(DEV), k=k, w=None if w is None else w.to(DEV))
    torch.cuda.synchronize()
    assert torch.equal(got.cpu(), ref)


@pytest.mark.parametrize("k,nonflat", CASES)
@pytest.mark.parametrize("dilate", [True, False])
def test_backward_vs_cpu_autograd(k, nonflat, dilate):
    shape = (2, 3, 40, 70)
    x = _x(shape, 11)
    w = _w(k, 5) if nonflat else None
    g = _x(shape, 12)
    xc = x.clone().requires_grad_(True)
    wc = w.clone().requires_grad_(True) if w is not None else None
    (autograd_ref.reference_cpu(xc, k=k, w=wc, dilate=dilate) * g).sum().backward()
    xd = x.to(DEV).requires_grad_(True)
    wd = w.to(DEV).requires_grad_(True) if w is not None else None
    f = ag.dilation2d if dilate else ag.erosion2d
    (f(xd, k=k, w=wd) * g.to(DEV)).sum().backward()
    torch.testing.assert_close(xd.grad.cpu(), xc.grad, rtol=1e-5, atol=1e-6)
    if w is not None:
        torch.testing.assert_close(wd.grad.cpu(), wc.grad, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("k,nonflat,dilate", [(3, False, True), (7, False, False), (5, True, True), (5, True, False)])
def test_finite_differences_tie_free(k, nonflat, dilate):
    eps = 1e-3
    x = _x((1, 1, 32, 32), 21).to(DEV)
    w = _w(k, 22).to(DEV) if nonflat else None
    g = _x((1, 1, 32, 32), 23).to(DEV)
    f = ag.dilation2d if dilate else ag.erosion2d
    xg = x.clone().requires_grad_(True)
    wg = w.clone().requires_grad_(True) if w is not None else None
    (f(xg, k=k, w=wg) * g).sum().backward()
    # windows whose best and second-best candidates are more than 2*eps apart
    r = (k - 1) // 2
    xp = torch.nn.functional.pad(x.cpu().reshape(1, 1, 32, 32), (r, r, r, r),
                                 value=float("-inf") if dilate else float("inf"))
    cand = torch.nn.functional.unfold(xp, k)[0]
    if w is not None:
        cand = cand + w.cpu().reshape(-1, 1) if dilate else cand - w.cpu().reshape(-1, 1)
    srt = cand.sort(dim=0, descending=dilate).values
    gap = (srt[0] - srt[1]).abs().reshape(32, 32)
    tie = (gap <= 2 * eps).float().reshape(1, 1, 32, 32)
    # an input pixel is checked only if no window that can select it has a near tie
    near = torch.nn.functional.max_pool2d(tie, k, stride=1, padding=r).reshape(-1) > 0
    ok = np.flatnonzero(~near.numpy())
    assert ok.size > 100

    def F(xx):
        return float((f(xx, k=k, w=w).double() * g.double()).sum())

    rng = np.random.default_rng(0)
    for idx in rng.choice(ok, size=12, replace=False):
        d = torch.zeros_like(x).view(-1)
        d[int(idx)] = eps
        d = d.view_as(x)
        fd = (F(x + d) - F(x - d)) / (2 * eps)
        an = float(xg.grad.view(-1)[int(idx)])
        assert abs(fd - an) <= 1e-3 * max(1.0, abs(an)), (idx, fd, an)
    if w is not None and not near.any():
        for t in range(k * k):
            dw = torch.zeros_like(w).view(-1)
            dw[t] = eps
            dw = dw.view_as(w)
            fp = float((f(x, k=k, w=w + dw).double() * g.double()).sum())
            fm = float((f(x, k=k, w=w - dw).double() * g.double()).sum())
            fd = (fp - fm) / (2 * eps)
            an = float(wg.grad.view(-1)[t])
            assert abs(fd - an) <= 1e-3 * max(1.0, abs(an)), (t, fd, an)


def test_flat_matches_library_dilate():
    """The flat operator equals Dilate(SESquare2D, iter=(k-1)/2) of the non-differentiable path."""
    x = _x((2, 3, 50, 60), 31).to(DEV)
    y = ag.dilation2d(x, k=7)
    ref = torch.stack([pm.dilate(x[i, j].contiguous(), pm.static_nl_se.square2d, 3)
                       for i in range(2) for j in range(3)]).reshape(x.shape)
    torch.cuda.synchronize()
    assert torch.equal(y, ref)


def test_bad_arguments():
    x = _x((1, 1, 8, 8), 1).to(DEV)
    with pytest.raises(pm.PoutreError):
        ag.dilation2d(x, k=4)
    with pytest.raises(TypeError):
        ag.dilation2d(x.double(), k=3)
    # a structuring function in host memory is rejected before any launch
    with pytest.raises(TypeError):
        ag.dilation2d(x, w=torch.zeros((3, 3)))
    with pytest.raises(TypeError):
        ag.erosion2d(x, w=np.zeros((3, 3), np.float32))


@pytest.mark.parametrize("k,with_w,dilate", [(3, False, True), (5, True, False), (7, False, False), (3, True, True), (7, True, True), (9, True, False)])
def test_stored_taps_match_recomputed_taps(k, with_w, dilate):
    """the C ABI pair that stores the selected taps (pm_maxplus_fwd_sel / _bwd_sel, used by the
    autograd Function) and the pair that re-selects them from x (pm_maxplus_fwd / _bwd) select
    the same taps: y agrees bit for bit.  gx from stored taps is summed over tap rows, then
    tap columns, where the re-selecting backward runs the taps in raster order, and gw sums
    per-wave partials where the other sums per-tile ones: those agree to fp32 rounding of the
    order (rtol 1e-5), each path being deterministic on its own"""
    import ctypes
    from poutre2_amd import _capi
    g = torch.Generator().manual_seed(40 + k)
    x = torch.randn(2, 3, 67, 130, generator=g).to(DEV)
    x[0, 0, 10:20, 10:20] = 0.5  # plateaus: ties resolved by the raster rule
    gy = torch.randn(x.shape, generator=g).to(DEV)
    w = (0.1 * torch.randn(k, k, generator=g)).to(DEV) if with_w else None
    B, H, W = 6, 67, 130
    p = (lambda t: ctypes.c_void_p(t.data_ptr() if t is not None else 0))
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    lib = _capi.lib()
    y1, y2 = torch.empty_like(x), torch.empty_like(x)
    sel = torch.empty(x.shape, dtype=torch.uint8, device=DEV)
    _capi.check(lib.pm_maxplus_fwd(p(x), B, H, W, p(w), k, int(dilate), p(y1), s))
    _capi.check(lib.pm_maxplus_fwd_sel(p(x), B, H, W, p(w), k, int(dilate), p(y2), p(sel), s))
    gx1, gx2 = torch.empty_like(x), torch.empty_like(x)
    gw1 = torch.empty(k, k, device=DEV) if with_w else None
    gw2 = torch.empty(k, k, device=DEV) if with_w else None
    _capi.check(lib.pm_maxplus_bwd(p(x), p(gy), B, H, W, p(w), k, int(dilate), p(gx1), p(gw1), s))
    _capi.check(lib.pm_maxplus_bwd_sel(p(sel), p(gy), B, H, W, k, int(dilate), int(with_w), p(gx2), p(gw2), s))
    torch.cuda.synchronize()
    assert int(sel.max()) < k * k
    assert torch.equal(y1, y2)
    torch.testing.assert_close(gx1, gx2, rtol=1e-5, atol=1e-5)
    if with_w:
        torch.testing.assert_close(gw1, gw2, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("k", [3, 7])
def test_stored_taps_of_a_structuring_function_without_gw(k):
    """pm_maxplus_bwd_sel with with_w = 0 sums gx separably when every output column agrees on
    the tap column (taps of a flat forward); taps of a structuring function break that in
    most tiles, which then gather directly: gx still equals the re-selecting backward's (to
    fp32 rounding, rtol 1e-5: a tile where the columns happen to agree sums separably)"""
    import ctypes
    from poutre2_amd import _capi
    x = _x((3, 70, 140), 60 + k).to(DEV)
    w = _w(k, 61 + k).to(DEV)
    gy = _x((3, 70, 140), 62 + k).to(DEV)
    B, H, W = 3, 70, 140
    p = (lambda t: ctypes.c_void_p(t.data_ptr() if t is not None else 0))
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    lib = _capi.lib()
    y = torch.empty_like(x)
    sel = torch.empty(x.shape, dtype=torch.uint8, device=DEV)
    _capi.check(lib.pm_maxplus_fwd_sel(p(x), B, H, W, p(w), k, 1, p(y), p(sel), s))
    gx1, gx2 = torch.empty_like(x), torch.empty_like(x)
    gw1 = torch.empty(k, k, device=DEV)
    _capi.check(lib.pm_maxplus_bwd(p(x), p(gy), B, H, W, p(w), k, 1, p(gx1), p(gw1), s))
    _capi.check(lib.pm_maxplus_bwd_sel(p(sel), p(gy), B, H, W, k, 1, 0, p(gx2), p(None), s))
    torch.cuda.synchronize()
    torch.testing.assert_close(gx1, gx2, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("k,with_w", [(3, False), (7, True)])
def test_recorded_and_plain_forward_agree(k, with_w):
    """a call autograd records stores the selected taps; one it does not (no_grad, or no input
    requiring grad) runs the max-only kernel: both give the same y bit for bit"""
    x = _x((2, 45, 77), 7).to(DEV)
    w = _w(k, 8).to(DEV) if with_w else None
    plain = ag.dilation2d(x, k=k, w=w)
    with torch.no_grad():
        nograd = ag.erosion2d(x.clone().requires_grad_(True), k=k, w=w)
    xr = x.clone().requires_grad_(True)
    rec = ag.dilation2d(xr, k=k, w=w)
    rec_e = ag.erosion2d(xr, k=k, w=w)
    assert rec.requires_grad and not nograd.requires_grad
    assert torch.equal(plain, rec.detach()) and torch.equal(nograd, rec_e.detach())


@pytest.mark.parametrize("k,with_w,dilate", [(3, False, True), (5, True, True), (7, True, False)])
def test_backward_is_deterministic(k, with_w, dilate):
    """gx and gw come out bit for bit the same on every run: every sum runs in a fixed order
    (no atomics), over tap rows, tap columns, lanes and waves"""
    f = ag.dilation2d if dilate else ag.erosion2d
    x = _x((2, 3, 150, 270), 90 + k).to(DEV)
    gy = _x((2, 3, 150, 270), 91 + k).to(DEV)
    w0 = _w(k, 92 + k).to(DEV) if with_w else None
    outs = []
    for _ in range(3):
        xr = x.clone().requires_grad_(True)
        wr = w0.clone().requires_grad_(True) if with_w else None
        f(xr, k=k, w=wr).backward(gy)
        outs.append((xr.grad.clone(), wr.grad.clone() if with_w else None))
    for gx_, gw_ in outs[1:]:
        assert torch.equal(gx_, outs[0][0])
        if with_w:
            assert torch.equal(gw_, outs[0][1])


@pytest.mark.parametrize("shape", [(1, 1), (2, 3), (1, 70), (70, 1), (5, 130), (33, 2), (67, 200)])
@pytest.mark.parametrize("k,nonflat,dilate", [(3, False, True), (7, False, False), (9, False, True), (5, True, False)])
def test_backward_small_and_ragged_with_plateaus(shape, k, nonflat, dilate):
    """images smaller than the window, one row or one column, widths across several strips of
    64 - 2R columns and heights across 32-row strips; integer values in 0..3 make plateaus, so
    the first-maximiser rule decides most routes.  Against torch CPU autograd (rtol 1e-5)."""
    g0 = np.random.default_rng(hash((shape, k)) % 2**32)
    x = torch.from_numpy(g0.integers(0, 4, (2,) + shape).astype(np.float32))
    gy = torch.from_numpy(g0.standard_normal((2,) + shape).astype(np.float32))
    w = _w(k, 9) if nonflat else None
    xc = x.clone().requires_grad_(True)
    wc = w.clone().requires_grad_(True) if w is not None else None
    yc = autograd_ref.reference_cpu(xc, k=k, w=wc, dilate=dilate)
    (yc * gy).sum().backward()
    xd = x.to(DEV).requires_grad_(True)
    wd = w.to(DEV).requires_grad_(True) if w is not None else None
    f = ag.dilation2d if dilate else ag.erosion2d
    yd = f(xd, k=k, w=wd)
    (yd * gy.to(DEV)).sum().backward()
    assert torch.equal(yd.detach().cpu(), yc.detach())
    torch.testing.assert_close(xd.grad.cpu(), xc.grad, rtol=1e-5, atol=1e-5)
    if w is not None:
        torch.testing.assert_close(wd.grad.cpu(), wc.grad, rtol=1e-5, atol=1e-5)
