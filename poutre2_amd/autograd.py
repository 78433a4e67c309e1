"""Differentiable float32 morphology on PyTorch-ROCm, backed by the hand-written kernels of
libpoutre_amd.so.  When autograd records the call (grad mode on and x or w requiring grad) the
forward keeps the selected tap of every output, one byte, for the backward
(pm_maxplus_fwd_sel / pm_maxplus_bwd_sel); otherwise it runs the plain max / min kernel
(pm_maxplus_fwd) and stores nothing.

    y = dilation2d(x, k=3)            flat dilation by a k x k square
    y = erosion2d(x, k=7)             flat erosion
    y = dilation2d(x, w=w)            max-plus dilation, y[q] = max_t x[q+t] + w[t]
    y = erosion2d(x, w=w)             y[q] = min_t x[q+t] - w[t]

x: float32 tensor (..., H, W) on a ROCm device (contiguous); w: (K, K) float32 with K odd in
{3, 5, 7, 9}.  Taps outside the image are skipped.  The gradient goes to the first
maximiser (minimiser) in raster order of the window (torch.max's first occurrence over an
F.unfold window); w receives the sum of the routed output gradients (negated for erosion).
"""
import ctypes

import torch

from . import _capi


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _stream(t):
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _dims(x):
    if x.dim() < 2:
        raise ValueError("expected a tensor of shape (..., H, W)")
    H, W = x.shape[-2], x.shape[-1]
    B = x.numel() // (H * W) if H * W else 0
    return B, H, W


class _MorphFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, w, k, dilate, record):
        if x.dtype != torch.float32 or not x.is_cuda:
            raise TypeError("x must be a float32 device tensor")
        x = x.contiguous()
        if w is not None:
            w = w.contiguous().to(torch.float32)
            k = w.shape[-1]
            if w.shape != (k, k):
                raise ValueError("w must be (K, K)")
        B, H, W = _dims(x)
        y = torch.empty_like(x)
        ctx.k, ctx.dilate, ctx.with_w = int(k), int(dilate), w is not None
        if not record:
            _capi.check(_capi.lib().pm_maxplus_fwd(_ptr(x), B, H, W, _ptr(w), int(k), int(dilate), _ptr(y),
                                                   _stream(x)))
            return y
        # the selected tap of every output (1 byte) is what the backward needs, not x
        sel = torch.empty(x.shape, dtype=torch.uint8, device=x.device)
        _capi.check(_capi.lib().pm_maxplus_fwd_sel(_ptr(x), B, H, W, _ptr(w), int(k), int(dilate), _ptr(y), _ptr(sel),
                                                   _stream(x)))
        ctx.save_for_backward(sel)
        return y

    @staticmethod
    def backward(ctx, gy):
        (sel,) = ctx.saved_tensors
        gy = gy.contiguous()
        B, H, W = _dims(sel)
        gx = torch.empty(sel.shape, dtype=torch.float32, device=sel.device)
        gw = torch.empty((ctx.k, ctx.k), dtype=torch.float32, device=sel.device) if ctx.with_w else None
        _capi.check(_capi.lib().pm_maxplus_bwd_sel(_ptr(sel), _ptr(gy), B, H, W, ctx.k, ctx.dilate, int(ctx.with_w),
                                                   _ptr(gx), _ptr(gw), _stream(sel)))
        return gx, (gw if (ctx.with_w and ctx.needs_input_grad[1]) else None), None, None, None


def _check(x, w):
    """the kernels read x and w through device pointers: both must be tensors on one device"""
    if not isinstance(x, torch.Tensor) or x.dtype != torch.float32 or not x.is_cuda:
        raise TypeError("x must be a float32 device tensor")
    if w is not None and (not isinstance(w, torch.Tensor) or not w.is_cuda or w.device != x.device):
        raise TypeError("w must be a tensor on the same device as x")


def _recorded(x, w):
    """True when autograd will record the call, so that a backward can follow"""
    return torch.is_grad_enabled() and (x.requires_grad or (w is not None and w.requires_grad))


def dilation2d(x, k=3, w=None):
    """Flat (k x k square) or max-plus (structuring function w) dilation, differentiable."""
    _check(x, w)
    return _MorphFn.apply(x, w, k, True, _recorded(x, w))


def erosion2d(x, k=3, w=None):
    """Flat (k x k square) or max-plus (structuring function w) erosion, differentiable."""
    _check(x, w)
    return _MorphFn.apply(x, w, k, False, _recorded(x, w))

