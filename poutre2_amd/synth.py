"""Deterministic synthetic inputs for the benchmark workloads and the parity tests.

All generators are seeded numpy PCG64 streams (numpy.random.Generator(PCG64(seed))),
so every run and every machine sees the same bytes.

Workloads (sizes as used by bench.py and the config-size parity tests):
  * uniform_u8(shape, seed)          iid uniform [0, 255] images (8192x8192 seed 2, 1024x1024
                                     seed 1, 512^3 seed 3)
  * blurred_noise_mask(H, W, seed)   white noise N(0,1), separable Gaussian blur sigma=8
                                     (kernel radius 32, mirror border, float64), min-max
                                     rescaled to [0, 255] and rounded half-to-even -> uint8
                                     (4096x4096 seed 4)
  * hmax_marker(mask, h)             max(mask - h, 0): the h-maxima marker
  * batch_dilate_open_workload()     512 uint8 images, H, W ~ U{64..2048}, square side
                                     ~ U{3, 5, .., 31} (seed 5)
  * batch_reconstruction_workload()  256 blurred-noise masks, H, W ~ U{128..1024} (seed 6,
                                     mask i from seed 1000 + i), h = 16
  * normal_f32(shape, seed)          N(0,1) + 1e-4 U(0,1) float32 ([64,32,256,256] seed 7)

blurred_noise_mask(..., device=d) runs the same float64 operations in the same order on a
torch device (one rounding per multiply and per add, as numpy does), so it returns the same
bytes as the host form in a fraction of the time; the parity tests check that by digest.
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


def _reflect_index(n, r):
    """source index of each padded position: mirror border without edge repeat (numpy's
    "reflect"), or with it ("symmetric") when the axis is not longer than the radius"""
    return np.pad(np.arange(n), (r, r), mode="reflect" if n > r else "symmetric")


def _blur_axis(a, k, axis):
    """Correlate along `axis` with mirror border, float64: out = sum_i k[i] * a[pad(i + .)],
    accumulated tap by tap from zero (one multiply and one add per tap and element)."""
    r = (len(k) - 1) // 2
    n = a.shape[axis]
    ap = np.take(a, _reflect_index(n, r), axis=axis)
    out = np.zeros_like(a, dtype=np.float64)
    for i, w in enumerate(k):
        sl = [slice(None)] * a.ndim
        sl[axis] = slice(i, i + n)
        out += w * ap[tuple(sl)]
    return out


def _blur_axis_torch(a, k, axis):
    import torch
    r = (len(k) - 1) // 2
    n = a.shape[axis]
    ap = a.index_select(axis, torch.from_numpy(_reflect_index(n, r)).to(a.device))
    out = torch.zeros_like(a)
    for i, w in enumerate(k):
        out += float(w) * ap.narrow(axis, i, n)
    return out


def blurred_noise_mask(H, W, seed, sigma=8.0, radius=32, device=None):
    noise = rng(seed).standard_normal((H, W))
    k = gaussian_kernel(sigma, radius)
    if device is None:
        return _rescale_u8(_blur_axis(_blur_axis(noise, k, 0), k, 1))
    import torch
    t = torch.from_numpy(noise).to(device)
    b = _blur_axis_torch(_blur_axis_torch(t, k, 0), k, 1)
    lo, hi = b.min().item(), b.max().item()
    if hi > lo:
        scaled = (b - lo) * (255.0 / (hi - lo))
    else:
        scaled = torch.zeros_like(b)
    return torch.round(scaled).clamp(0, 255).to(torch.uint8)


def _rescale_u8(b):
    lo, hi = b.min(), b.max()
    scaled = (b - lo) * (255.0 / (hi - lo)) if hi > lo else np.zeros_like(b)
    return np.rint(scaled).clip(0, 255).astype(np.uint8)


def hmax_marker(mask, h):
    """max(mask - h, 0) for a numpy uint8 mask, or a torch uint8 tensor"""
    if isinstance(mask, np.ndarray):
        return np.maximum(mask.astype(np.int32) - int(h), 0).astype(mask.dtype)
    return (mask.int() - int(h)).clamp_min(0).to(mask.dtype)


def normal_f32(shape, seed, jitter=1e-4):
    g = rng(seed)
    x = g.standard_normal(shape, dtype=np.float32)
    x += np.float32(jitter) * g.random(shape, dtype=np.float32)
    return x


def batch_dilate_open_workload(seed=5, n=512, lo=64, hi=2048, max_radius=15):
    """(shapes, radii, images): n uint8 images with H, W ~ U{lo..hi} drawn first, then the
    square radii ~ U{1..max_radius} (sides 3..31), then the pixels, all from one stream."""
    g = rng(seed)
    shapes = [(int(g.integers(lo, hi + 1)), int(g.integers(lo, hi + 1))) for _ in range(n)]
    radii = [int(g.integers(1, max_radius + 1)) for _ in range(n)]
    imgs = [g.integers(0, 256, size=s, dtype=np.uint8) for s in shapes]
    return shapes, radii, imgs


def batch_reconstruction_workload(seed=6, n=256, lo=128, hi=1024, h=16, device=None):
    """(markers, masks): n blurred-noise masks (the 4096^2 recipe, sigma 8) with H, W ~ U{lo..hi}
    from `seed`, mask i from seed 1000 + i, and their h-maxima markers max(mask - h, 0)."""
    g = rng(seed)
    shapes = [(int(g.integers(lo, hi + 1)), int(g.integers(lo, hi + 1))) for _ in range(n)]
    masks = [blurred_noise_mask(H, W, 1000 + i, device=device) for i, (H, W) in enumerate(shapes)]
    return [hmax_marker(m, h) for m in masks], masks
