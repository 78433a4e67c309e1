"""CPU references for the differentiable float32 operators of poutre2_amd.autograd (test
infrastructure only).

reference_cpu: torch CPU form of the same operator (F.unfold + max / min, first occurrence),
the autograd reference of tests/test_gpu_autograd.py.  Out-of-image taps are -inf (+inf), so
they are never selected.

forward_with_taps / backward_from_taps: numpy forms over whole [..., H, W] tensors, tap by
tap in raster order with a strict comparison (the first maximiser / minimiser wins), used for
the full-size digests of the [64, 32, 256, 256] workload: the forward value of every output,
the selected tap, and the gradient routed back to the selected input pixel.
"""
import numpy as np


def reference_cpu(x, k=3, w=None, dilate=True):
    import torch.nn.functional as F
    H, W = x.shape[-2], x.shape[-1]
    if w is not None:
        k = w.shape[-1]
    r = (k - 1) // 2
    pad = float("-inf") if dilate else float("inf")
    xb = x.reshape(-1, 1, H, W)
    xp = F.pad(xb, (r, r, r, r), value=pad)
    u = F.unfold(xp, k)  # (B, k*k, H*W), taps in raster order
    if w is not None:
        wv = w.reshape(1, k * k, 1)
        u = u + wv if dilate else u - wv
    y = u.max(dim=1).values if dilate else u.min(dim=1).values
    return y.reshape(x.shape)


def _shifted(xp, r, dy, dx, H, W):
    """view of the padded tensor aligned so that element q reads x[q + (dy, dx)]"""
    return xp[..., r + dy:r + dy + H, r + dx:r + dx + W]


def forward_with_taps(x, k=3, w=None, dilate=True):
    """(y, tap) for float32 x of shape (..., H, W): y[q] = max_t x[q+t] + w[t] (dilation) or
    min_t x[q+t] - w[t] (erosion), tap = dy*k + dx of the first extremum in raster order;
    float32 adds (correctly rounded, as on the device)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    if w is not None:
        k = w.shape[-1]
        w = np.asarray(w, dtype=np.float32)
    r = (k - 1) // 2
    H, W = x.shape[-2:]
    pad = np.float32(-np.inf if dilate else np.inf)
    xp = np.full(x.shape[:-2] + (H + 2 * r, W + 2 * r), pad, dtype=np.float32)
    xp[..., r:r + H, r:r + W] = x
    best = np.full(x.shape, pad, dtype=np.float32)
    tap = np.zeros(x.shape, dtype=np.uint8)
    for t in range(k * k):
        dy, dx = t // k - r, t % k - r
        v = _shifted(xp, r, dy, dx, H, W)
        if w is not None:
            v = v + w[t // k, t % k] if dilate else v - w[t // k, t % k]
        upd = v > best if dilate else v < best
        np.copyto(best, v, where=upd)
        np.copyto(tap, np.uint8(t), where=upd)
    return best, tap


def backward_from_taps(tap, gy, k):
    """gx[p] = sum of gy[q] over the outputs q whose selected tap lands on p; gw[t] = sum of
    gy[q] over the outputs that selected tap t.  Summed in float64 (exact for the
    integer-valued gy the digests use)."""
    r = (k - 1) // 2
    H, W = tap.shape[-2:]
    gy = np.asarray(gy, dtype=np.float64)
    gxp = np.zeros(tap.shape[:-2] + (H + 2 * r, W + 2 * r), dtype=np.float64)
    gw = np.zeros((k, k), dtype=np.float64)
    for t in range(k * k):
        sel = np.where(tap == t, gy, 0.0)
        dy, dx = t // k - r, t % k - r
        _shifted(gxp, r, dy, dx, H, W)[...] += sel
        gw[t // k, t % k] = sel.sum()
    return gxp[..., r:r + H, r:r + W], gw
