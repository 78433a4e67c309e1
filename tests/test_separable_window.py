"""The flat autograd kernels (poutre2_amd/csrc/maxplus.hip: mp_fwd_flat_kernel,
mp_bwd_flat_kernel) rest on two identities, checked here on the CPU by brute force:
  * forward: the first maximiser (minimiser) in raster order of a K x K window is found by
    a row pass (each row's extremum and its first arg dx) followed by a column pass taking
    the first row that strictly beats the running best -- same tap, same value bits, with
    ties, NaN and signed zeros;
  * backward: gx[p] = sum over q with q + tap(q) = p of gy[q] equals the two-stage gather
    (stage A: per input row and output column, the sum over the q whose tap row lands there
    and their tap column; stage B: along the row, the columns whose tap column lands on p),
    whenever every (row, column) sees one tap column -- and taps drawn at random are
    detected as mixing columns.
Pure-Python loops on small images: the statement of the algorithm, not the kernels."""
import numpy as np
import pytest


def _pad(x, r, value):
    H, W = x.shape
    xp = np.full((H + 2 * r, W + 2 * r), value, x.dtype)
    xp[r:r + H, r:r + W] = x
    return xp


def _better(v, best, dil):
    return (v > best) if dil else (v < best)


def raster_scan(x, K, dil):
    R = (K - 1) // 2
    pad = -np.inf if dil else np.inf
    xp = _pad(x, R, pad)
    y = np.empty_like(x)
    s = np.zeros(x.shape, np.int64)
    for i in range(x.shape[0]):
        for j in range(x.shape[1]):
            best, sel = np.float32(pad), 0
            for dy in range(K):
                for dx in range(K):
                    v = xp[i + dy, j + dx]
                    if _better(v, best, dil):
                        best, sel = v, dy * K + dx
            y[i, j], s[i, j] = best, sel
    return y, s


def separable_scan(x, K, dil):
    R = (K - 1) // 2
    pad = -np.inf if dil else np.inf
    xp = _pad(x, R, pad)
    H, W = x.shape
    hm = np.empty((H + 2 * R, W), x.dtype)
    ha = np.zeros((H + 2 * R, W), np.int64)
    for i in range(H + 2 * R):
        for j in range(W):
            best, a = np.float32(pad), 0
            for dx in range(K):
                v = xp[i, j + dx]
                if _better(v, best, dil):
                    best, a = v, dx
            hm[i, j], ha[i, j] = best, a
    y = np.empty_like(x)
    s = np.zeros(x.shape, np.int64)
    for i in range(H):
        for j in range(W):
            best, sel = np.float32(pad), 0
            for dy in range(K):
                if _better(hm[i + dy, j], best, dil):
                    best, sel = hm[i + dy, j], dy * K + ha[i + dy, j]
            y[i, j], s[i, j] = best, sel
    return y, s


def scatter_backward(s, g, K):
    H, W = s.shape
    R = (K - 1) // 2
    gx = np.zeros((H, W))
    for qi in range(H):
        for qj in range(W):
            dy, dx = divmod(int(s[qi, qj]), K)
            pi, pj = qi + dy - R, qj + dx - R
            if 0 <= pi < H and 0 <= pj < W:
                gx[pi, pj] += g[qi, qj]
    return gx


def two_stage_backward(s, g, K):
    H, W = s.shape
    R = (K - 1) // 2
    sp = _pad(s, R, 255)
    gp = _pad(g, R, 0.0)
    QC = W + 2 * R
    gr = np.zeros((H, QC))
    ad = np.full((H, QC), 255)
    mixed = False
    for i in range(H):
        for c in range(QC):
            a = -1
            for dy in range(K):
                d = sp[i + 2 * R - dy, c] - dy * K  # tap column of q when its tap row is dy
                if 0 <= d < K:
                    gr[i, c] += gp[i + 2 * R - dy, c]
                    mixed |= a >= 0 and a != d
                    a = d
            ad[i, c] = 255 if a < 0 else a
    gx = np.zeros((H, W))
    for i in range(H):
        for j in range(W):
            for dx in range(K):
                if ad[i, j + 2 * R - dx] == dx:
                    gx[i, j] += gr[i, j + 2 * R - dx]
    return gx, mixed


def _images(rng, shape):
    yield rng.integers(0, 4, shape).astype(np.float32)  # plateaus: ties everywhere
    x = rng.standard_normal(shape).astype(np.float32)
    x[1:4, 2:7] = np.nan
    yield x
    yield np.where(rng.random(shape) < 0.5, np.float32(0.0), np.float32(-0.0))
    yield np.full(shape, -np.inf, np.float32)  # every tap equals the padding


@pytest.mark.parametrize("K", [3, 5, 7])
@pytest.mark.parametrize("dil", [True, False])
def test_separable_scan_selects_the_raster_tap(K, dil):
    rng = np.random.default_rng(K * 2 + dil)
    for x in _images(rng, (9, 13)):
        y1, s1 = raster_scan(x, K, dil)
        y2, s2 = separable_scan(x, K, dil)
        assert np.array_equal(s1, s2)
        assert np.array_equal(y1.view(np.uint32), y2.view(np.uint32))


@pytest.mark.parametrize("K", [3, 5, 7])
def test_two_stage_backward_is_the_scatter(K):
    rng = np.random.default_rng(40 + K)
    for dil in (True, False):
        for x in _images(rng, (9, 13)):
            _, s = separable_scan(x, K, dil)
            g = rng.standard_normal(x.shape)
            gx, mixed = two_stage_backward(s, g, K)
            assert not mixed
            np.testing.assert_allclose(gx, scatter_backward(s, g, K), rtol=1e-12, atol=1e-12)
    # taps that a flat forward cannot produce: detected
    s = rng.integers(0, K * K, (9, 13))
    assert two_stage_backward(s, rng.standard_normal((9, 13)), K)[1]
