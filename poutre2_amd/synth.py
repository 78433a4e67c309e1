"""Deterministic synthetic inputs for the benchmark workloads and the parity tests.

All generators are seeded numpy PCG64 streams (numpy.random.Generator(PCG64(seed))),
so every run and every machine sees the same bytes.

Workloads (sizes as used by bench.py):
  * uniform_u8(shape, seed)          iid uniform [0, 255] images (8192x8192, 1024x1024, 512^3)
  * blurred_noise_mask(H, W, seed)   white noise N(0,1), separable Gaussian blur sigma=8
                                     (kernel radius 32, mirror border, float64), min-max
                                     rescaled to [0, 255] and rounded half-to-even -> uint8
  * hmax_marker(mask, h)             max(mask - h, 0): the h-maxima marker
"""
import numpy as np


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def uniform_u8(shape, seed):
    return rng(seed).integers(0, 256, size=shape, dtype=np.uint8)


def gaussian_kernel(sigma=8.0, radius=32):
    x = np.arange(-radius, radius + 1, dtype=np.float64)
    k = np.exp(-0.5 * (x / sigma) ** 2)
    return k / k.sum()


def _blur_axis(a, k, axis):
    """Correlate along `axis` with mirror (reflect-without-edge-repeat) border, float64."""
    r = (len(k) - 1) // 2
    n = a.shape[axis]
    pad = [(0, 0)] * a.ndim
    pad[axis] = (r, r)
    mode = "reflect" if n > r else "symmetric"
    ap = np.pad(a, pad, mode=mode)
    out = np.zeros_like(a, dtype=np.float64)
    for i, w in enumerate(k):
        sl = [slice(None)] * a.ndim
        sl[axis] = slice(i, i + n)
        out += w * ap[tuple(sl)]
    return out


def blurred_noise_mask(H, W, seed, sigma=8.0, radius=32):
    noise = rng(seed).standard_normal((H, W))
    k = gaussian_kernel(sigma, radius)
    b = _blur_axis(_blur_axis(noise, k, 0), k, 1)
    return _rescale_u8(b)


def blurred_noise_mask_fast(H, W, seed, sigma=8.0, radius=32):
    """Same recipe as blurred_noise_mask with scipy's separable correlation (mode "mirror"),
    for the large batched workloads; float64 summation order differs, so the two may differ
    by one grey level at exact .5 rounding ties."""
    from scipy.ndimage import correlate1d
    noise = rng(seed).standard_normal((H, W))
    k = gaussian_kernel(sigma, radius)
    b = correlate1d(correlate1d(noise, k, axis=0, mode="mirror"), k, axis=1, mode="mirror")
    return _rescale_u8(b)


def _rescale_u8(b):
    lo, hi = b.min(), b.max()
    scaled = (b - lo) * (255.0 / (hi - lo)) if hi > lo else np.zeros_like(b)
    return np.rint(scaled).clip(0, 255).astype(np.uint8)


def hmax_marker(mask, h):
    return np.maximum(mask.astype(np.int32) - int(h), 0).astype(mask.dtype)


def normal_f32(shape, seed, jitter=1e-4):
    g = rng(seed)
    x = g.standard_normal(shape, dtype=np.float32)
    x += np.float32(jitter) * g.random(shape, dtype=np.float32)
    return x
