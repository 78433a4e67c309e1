"""GPU parity: every entry point through the C ABI, compared bit-exactly with the oracle
(and, for the reference's own known-answer tests, with the reference's expected output)."""
import numpy as np
import pytest

import oracle
from tests import kats

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
pm = pytest.importorskip("poutre2_amd")

DEV = "cuda:0"
DTYPES = [np.uint8, np.uint16, np.int32, np.int64, np.float32, np.float64]
SE2D = ["SESquare2D", "SECross2D", "SESegmentX2D", "SESegmentY2D"]
SE3D = ["SESegmentX3D", "SESegmentY3D", "SESegmentZ3D", "SECross3D", "SESquare3D"]
PM_SE = {n: v for n, v in oracle.SE.items()}


def g(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def h(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def rand(shape, dtype, seed, lo=0, hi=256):
    r = np.random.default_rng(seed)
    a = r.integers(lo, hi, size=shape)
    if np.dtype(dtype).kind == "f":
        a = a.astype(np.float64) + r.random(shape) * 0.5 - 0.25
    if np.dtype(dtype).kind in "iu" and np.dtype(dtype).itemsize >= 4:
        a = a * 1000003 - 128 * 1000003  # exercise negative values and wide ranges
    return a.astype(dtype)


# ----------------------------------------------------------------------------- KATs
def _gpu_kat(c):
    ins = [g(kats.image(d)) for d in c["inputs"]]
    op, se = c["op"], c.get("se")
    if op in ("t_Dilate", "t_Erode"):
        f = pm.dilate if op == "t_Dilate" else pm.erode
        if c["runtime_se"]:
            return f(ins[0], [tuple(int(v) for v in o) for o in oracle.se_offsets(se)])
        return f(ins[0], PM_SE[se], 1)
    if op in ("t_DilateX", "t_ErodeX", "t_DilateY", "t_ErodeY"):
        return getattr(pm, {"t_DilateX": "dilateX", "t_ErodeX": "erodeX", "t_DilateY": "dilateY",
                            "t_ErodeY": "erodeY"}[op])(ins[0], c["r"])
    if op == "Reconstruction":
        return pm.reconstruction(pm.RecType[c["rec_type"]], ins[0], ins[1], PM_SE[se])
    if op in ("low_leveling", "high_leveling", "leveling"):
        return getattr(pm, op)(ins[0], ins[1], PM_SE[se])
    if op in ("h_minima", "h_maxima", "h_concave", "h_convex", "dynamic_pseudo_closing", "dynamic_pseudo_opening"):
        return getattr(pm, op)(ins[0], c["h"], PM_SE[se])
    if op in ("label_binary", "label_flat_zones"):
        return getattr(pm, op)(ins[0], PM_SE[se])[0]
    if op in ("label_minima", "label_maxima"):
        return getattr(pm, op)(ins[0], PM_SE[se])
    pytest.skip(f"{op} is not a device operation")


@pytest.mark.parametrize("case", kats.cases(lambda c: c["op"] != "t_transpose"), ids=kats.case_id)
def test_gpu_matches_reference_kat(case):
    got = h(_gpu_kat(case))
    np.testing.assert_array_equal(got, kats.image(case["expected"]))


# ----------------------------------------------------------------------------- erosion / dilation
@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("se", SE2D)
@pytest.mark.parametrize("it", [1, 2, 5])
def test_static_2d(dtype, se, it):
    img = rand((37, 53), dtype, 1)
    for dil in (True, False):
        f = pm.dilate if dil else pm.erode
        np.testing.assert_array_equal(h(f(g(img), PM_SE[se], it)), oracle.static_iter(img, se, it, dil))


@pytest.mark.parametrize("dtype", [np.uint8, np.int32, np.float32, np.int64])
@pytest.mark.parametrize("se", SE3D)
@pytest.mark.parametrize("it", [1, 2, 3])
def test_static_3d(dtype, se, it):
    img = rand((9, 10, 11), dtype, 2)
    for dil in (True, False):
        f = pm.dilate if dil else pm.erode
        np.testing.assert_array_equal(h(f(g(img), PM_SE[se], it)), oracle.static_iter(img, se, it, dil))


@pytest.mark.parametrize("se", ["SECross3D", "SESquare3D", "SESegmentZ3D"])
@pytest.mark.parametrize("shape,it", [((11, 13, 48), 1), ((11, 13, 48), 2), ((7, 9, 32), 3), ((20, 5, 16), 8),
                                      ((20, 5, 16), 9), ((3, 17, 64), 4), ((1, 6, 16), 2), ((4, 2, 32), 2),
                                      ((2, 3, 16), 5), ((40, 70, 272), 2), ((40, 70, 272), 3), ((33, 65, 144), 4),
                                      ((17, 40, 256), 6), ((16, 32, 128), 3)])
def test_static_3d_u8_volume_kernels(se, shape, it):
    """uint8 volumes with W % 16 == 0 take the z-streaming kernels (z radius <= 8; 9 falls back);
    the larger shapes cover several 2D tiles and partial ones on every axis."""
    img = rand(shape, np.uint8, 5 + it)
    for dil in (True, False):
        f = pm.dilate if dil else pm.erode
        np.testing.assert_array_equal(h(f(g(img), PM_SE[se], it)), oracle.static_iter(img, se, it, dil))


@pytest.mark.parametrize("n,it", [(1, 1), (101, 1), (101, 4), (5000, 40)])
def test_static_1d(n, it):
    img = rand((n,), np.uint8, 3)
    for dil in (True, False):
        f = pm.dilate if dil else pm.erode
        np.testing.assert_array_equal(h(f(g(img), pm.static_nl_se.segx1d, it)),
                                      oracle.static_iter(img, "SESegmentX1D", it, dil))


def test_static_iter0_copies():
    img = rand((8, 9), np.uint8, 4)
    np.testing.assert_array_equal(h(pm.dilate(g(img), pm.static_nl_se.square2d, 0)), img)


@pytest.mark.parametrize("dtype", [np.uint8, np.int32, np.float64])
@pytest.mark.parametrize("shape", [(64, 64), (70, 300), (1, 17), (513, 7)])
@pytest.mark.parametrize("r", [1, 3, 7, 31, 100, 3000])
def test_line_x(dtype, shape, r):
    img = rand(shape, dtype, 5)
    for dil in (True, False):
        f = pm.dilateX if dil else pm.erodeX
        np.testing.assert_array_equal(h(f(g(img), r)), oracle.line_x(img, r, dil))


@pytest.mark.parametrize("dtype", [np.uint8, np.int32, np.float64])
@pytest.mark.parametrize("shape", [(64, 64), (300, 70), (17, 1), (7, 513)])
@pytest.mark.parametrize("r", [1, 3, 7, 31, 100, 3000])
def test_line_y(dtype, shape, r):
    img = rand(shape, dtype, 6)
    for dil in (True, False):
        f = pm.dilateY if dil else pm.erodeY
        exp = oracle.static_iter(img, "SESegmentY2D", min(r, shape[0]), dil)
        if shape[0] == shape[1]:
            np.testing.assert_array_equal(oracle.line_y(img, r, dil), exp)
        np.testing.assert_array_equal(h(f(g(img), r)), exp)


def test_line_x_1d_and_r0():
    img = rand((1000,), np.uint8, 7)
    np.testing.assert_array_equal(h(pm.dilateX(g(img), 5)), oracle.line_x(img, 5, True))
    np.testing.assert_array_equal(h(pm.erodeX(g(img), 0)), img)


@pytest.mark.parametrize("ry,rx", [(1, 1), (3, 0), (0, 5), (7, 2), (31, 31), (40, 3)])
def test_rect(ry, rx):
    img = rand((150, 170), np.uint8, 8)
    for dil in (True, False):
        f = pm.dilate_rect if dil else pm.erode_rect
        exp = oracle.static_iter(oracle.line_x(img, rx, dil), "SESegmentY2D", ry, dil)
        np.testing.assert_array_equal(h(f(g(img), ry, rx)), exp)


@pytest.mark.parametrize("size", [0, 1, 2, 3, 5])
def test_compound_octagon(size):
    img = rand((40, 45), np.int32, 9)
    for dil in (True, False):
        f = pm.dilate if dil else pm.erode
        np.testing.assert_array_equal(h(f(g(img), pm.static_compound_se.octagon, size)),
                                      oracle.compound(img, 1, size, dil))


@pytest.mark.parametrize("size", [0, 2, 3])
def test_compound_rhombicuboctahedron(size):
    img = rand((8, 9, 10), np.uint8, 10)
    for dil in (True, False):
        f = pm.dilate if dil else pm.erode
        np.testing.assert_array_equal(h(f(g(img), pm.static_compound_se.rhombicuboctahedron, size)),
                                      oracle.compound(img, 2, size, dil))


def test_compound_rhombicuboctahedron_size1_throws():
    img = g(rand((4, 5, 6), np.uint8, 11))
    with pytest.raises(pm.PoutreError):
        pm.dilate(img, pm.static_compound_se.rhombicuboctahedron, 1)


@pytest.mark.parametrize("offs", [[(0, 0), (2, -1), (-3, 0)], [(1, 1)], [(0, -4), (0, 4), (5, 5), (-1, 2)]])
@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
def test_runtime_nl_2d(offs, dtype):
    img = rand((33, 41), dtype, 12)
    for dil in (True, False):
        f = pm.dilate if dil else pm.erode
        np.testing.assert_array_equal(h(f(g(img), offs)), oracle.nl(img, offs, dil))


def test_runtime_nl_3d():
    img = rand((6, 7, 8), np.int32, 13)
    offs = [(0, 0, 0), (1, -1, 2), (-2, 0, 0)]
    np.testing.assert_array_equal(h(pm.dilate(g(img), offs)), oracle.nl(img, offs, True))


@pytest.mark.parametrize("se,it", [("SESquare2D", 1), ("SESquare2D", 3), ("SECross2D", 2), ("SESegmentX2D", 4)])
def test_open_close_gradient(se, it):
    img = rand((60, 70), np.uint8, 14)
    e = oracle.static_iter(img, se, it, False)
    d = oracle.static_iter(img, se, it, True)
    np.testing.assert_array_equal(h(pm.opening(g(img), PM_SE[se], it)), oracle.static_iter(e, se, it, True))
    np.testing.assert_array_equal(h(pm.closing(g(img), PM_SE[se], it)), oracle.static_iter(d, se, it, False))
    np.testing.assert_array_equal(h(pm.gradient(g(img), PM_SE[se], it)), oracle.ssub_image(d, e))


@pytest.mark.parametrize("off,n", [(0, 4096), (3, 1), (5, 15), (7, 100), (1, 70 * 61 + 3)])
def test_gradient_iter0_fills_zero(off, n):
    """iter = 0 gives an all-zero gradient: the fill kernel writes exactly the view's bytes
    (unaligned head and tail included) and nothing around it."""
    buf = torch.full((n + 32,), 9, dtype=torch.uint8, device=DEV)
    src = torch.full((n,), 200, dtype=torch.uint8, device=DEV)
    out = buf[off:off + n]
    pm.gradient(src, pm.static_nl_se.segx1d, 0, out=out)
    b = buf.cpu().numpy()
    assert (b[off:off + n] == 0).all()
    assert (b[:off] == 9).all() and (b[off + n:] == 9).all()


# ----------------------------------------------------------------------------- geodesy
def _recon_inputs(shape, dtype, seed, sparse=True):
    r = np.random.default_rng(seed)
    mask = r.integers(0, 60, size=shape).astype(dtype)
    marker = r.integers(0, 60, size=shape).astype(dtype)
    if sparse:
        marker = np.where(r.random(shape) < 0.02, marker, 0).astype(dtype)
    return marker, mask


@pytest.mark.parametrize("dtype", [np.uint8, np.int32, np.int64, np.float32, np.uint16])
@pytest.mark.parametrize("se", ["SESquare2D", "SECross2D"])
@pytest.mark.parametrize("shape", [(64, 64), (130, 200), (5, 300), (257, 33)])
@pytest.mark.parametrize("rec", ["dilate", "erode"])
def test_reconstruction_2d(dtype, se, shape, rec):
    marker, mask = _recon_inputs(shape, dtype, 21, sparse=(rec == "dilate"))
    got = h(pm.reconstruction(pm.RecType[rec], g(marker), g(mask), PM_SE[se]))
    np.testing.assert_array_equal(got, oracle.reconstruct(rec, marker, mask, se))


def test_reconstruction_scratch_reuse_across_calls():
    """uint8 reconstructions reuse their device scratch from call to call, its round tags
    offset per call instead of cleared: alternating sizes (the larger buffer kept, a smaller
    call reusing it, a larger one taking a new one), both operators, speculative leveling
    calls (whose first round may abort) and the scratch-free entry point in between, every
    result equal to the oracle's"""
    import ctypes

    from poutre2_amd import _capi, synth
    seq = [((300, 310), "dilate", "SESquare2D"), ((64, 70), "erode", "SECross2D"),
           ((300, 310), "erode", "SESquare2D"), ((520, 130), "dilate", "SECross2D"),
           ((64, 70), "dilate", "SESquare2D"), ((300, 310), "dilate", "SECross2D")]
    for k, (shape, rec, se) in enumerate(seq):
        marker, mask = _recon_inputs(shape, np.uint8, 40 + k, sparse=(rec == "dilate"))
        got = h(pm.reconstruction(pm.RecType[rec], g(marker), g(mask), PM_SE[se]))
        np.testing.assert_array_equal(got, oracle.reconstruct(rec, marker, mask, se))
        f = synth.blurred_noise_mask(shape[0], shape[1], 50 + k, sigma=3.0, radius=10)
        for marker_img in (oracle.static_iter(f, "SESquare2D", 2, False), f ^ 1):  # below f; neither
            got = h(pm.leveling(g(f), g(marker_img), PM_SE[se]))
            np.testing.assert_array_equal(got, oracle.leveling(f, marker_img, se))
        tm, tk = g(marker), g(mask)  # the descriptors hold raw pointers: keep the tensors alive
        out = torch.empty_like(tk)
        dm, dk, do = pm._desc(tm), pm._desc(tk), pm._desc(out)
        s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _capi.check(_capi.lib().pm_reconstruct_noalloc(int(pm.RecType[rec]), ctypes.byref(dm), ctypes.byref(dk),
                                                       int(PM_SE[se]), ctypes.byref(do), s))
        np.testing.assert_array_equal(h(out), oracle.reconstruct(rec, marker, mask, se))


def test_reconstruction_long_geodesic():
    """A serpentine corridor forces propagation through many tiles and rounds."""
    H, W = 256, 256
    mask = np.zeros((H, W), np.uint8)
    for y in range(0, H, 4):
        mask[y, :] = 200
        x = W - 1 if (y // 4) % 2 == 0 else 0
        mask[y:y + 4, x] = 200
    marker = np.zeros_like(mask)
    marker[0, 0] = 200
    got = h(pm.reconstruction(pm.RecType.dilate, g(marker), g(mask), pm.static_nl_se.cross2d))
    np.testing.assert_array_equal(got, oracle.reconstruct("dilate", marker, mask, "SECross2D"))


@pytest.mark.parametrize("se,shape", [("SESegmentX1D", (777,)), ("SECross3D", (7, 8, 9)), ("SESquare3D", (6, 7, 8))])
@pytest.mark.parametrize("rec", ["dilate", "erode"])
def test_reconstruction_1d_3d(se, shape, rec):
    marker, mask = _recon_inputs(shape, np.uint8, 22, sparse=(rec == "dilate"))
    got = h(pm.reconstruction(pm.RecType[rec], g(marker), g(mask), PM_SE[se]))
    np.testing.assert_array_equal(got, oracle.reconstruct_bf(rec, marker, mask, se))


@pytest.mark.parametrize("op", ["low_leveling", "high_leveling", "leveling"])
@pytest.mark.parametrize("dtype", [np.uint8, np.int32, np.float32])
@pytest.mark.parametrize("se,shape", [("SESquare2D", (90, 110)), ("SECross2D", (70, 65)), ("SESegmentX1D", (300,))])
def test_leveling(op, dtype, se, shape):
    r = np.random.default_rng(31)
    ref = r.integers(0, 40, size=shape).astype(dtype)
    marker = r.integers(0, 40, size=shape).astype(dtype)
    got = h(getattr(pm, op)(g(ref), g(marker), PM_SE[se]))
    np.testing.assert_array_equal(got, getattr(oracle, op)(ref, marker, se))


@pytest.mark.parametrize("op", ["low_leveling", "high_leveling", "leveling"])
@pytest.mark.parametrize("dtype", [np.uint8, np.int32, np.float32])
@pytest.mark.parametrize("order", ["opening", "closing", "equal"])
def test_leveling_ordered_marker(op, dtype, order):
    """markers below the reference (an opening: the benchmark's case), above it (a closing)
    and equal to it take the library's single-reconstruction forms; the oracle is the
    priority-queue restatement"""
    from poutre2_amd import synth
    ref = synth.blurred_noise_mask(150, 170, 12, sigma=3.0, radius=10).astype(dtype)
    if order == "equal":
        marker = ref.copy()
    else:
        first, second = (False, True) if order == "opening" else (True, False)
        marker = oracle.static_iter(oracle.static_iter(ref, "SESquare2D", 3, first), "SESquare2D", 3, second)
    got = h(getattr(pm, op)(g(ref), g(marker), PM_SE["SESquare2D"]))
    np.testing.assert_array_equal(got, getattr(oracle, op)(ref, marker, "SESquare2D"))


@pytest.mark.parametrize("pos,vertical", [((0, 61), True), ((0, 62), True), ((149, 123), True), ((61, 124), True),
                                          ((62, 1), True), ((149, 168), True), ((75, 61), True), ((61, 0), False),
                                          ((62, 169), False), ((123, 0), False), ((124, 169), False), ((1, 0), False),
                                          ((148, 169), False)])
def test_leveling_marker_above_at_one_pixel(pos, vertical):
    """The uint8 leveling assumes a marker below the reference and lets the reconstruction's
    first round check that on every pixel: here the marker exceeds the reference at ONE pixel
    (tile corners and edges, image borders), a gap in a low wall of the reference through which
    the true leveling (low(f, g) = f ^ R_{f v g}(g)) floods the far side, while the assumed form
    R_f(g) cannot: a missed pixel gives a different image."""
    H, W = 150, 170
    y, x = pos
    f = np.full((H, W), 100, np.uint8)
    m = np.zeros((H, W), np.uint8)
    if vertical:
        f[:, x] = 10
        m[:, :x] = 100
    else:
        f[y, :] = 10
        m[:y, :] = 100
    m[y, x] = 100
    expected = oracle.leveling(f, m, "SESquare2D")
    assert (expected != oracle.reconstruct("dilate", m, f, "SESquare2D")).any()
    for op in ("leveling", "low_leveling"):
        np.testing.assert_array_equal(h(getattr(pm, op)(g(f), g(m), PM_SE["SESquare2D"])),
                                      getattr(oracle, op)(f, m, "SESquare2D"))


@pytest.mark.parametrize("op", ["h_maxima", "h_minima", "h_convex", "h_concave", "dynamic_pseudo_opening",
                                "dynamic_pseudo_closing"])
@pytest.mark.parametrize("dtype", [np.uint8, np.int32])
def test_extrema(op, dtype):
    from poutre2_amd import synth
    f = synth.blurred_noise_mask(96, 120, 41, sigma=3.0, radius=12).astype(dtype)
    got = h(getattr(pm, op)(g(f), 16, pm.static_nl_se.square2d))
    np.testing.assert_array_equal(got, getattr(oracle, op)(f, 16, "SESquare2D"))


# ----------------------------------------------------------------------------- errors
def test_errors_follow_reference_checks():
    a = g(rand((8, 8), np.uint8, 50))
    with pytest.raises(pm.PoutreError) as e:
        pm.dilate(a, pm.static_nl_se.square2d, 1, out=g(rand((8, 9), np.uint8, 1)))
    assert e.value.code == -1
    with pytest.raises(pm.PoutreError) as e:
        pm.dilate(a, pm.static_nl_se.square2d, 1, out=g(rand((8, 8), np.int32, 1)))
    assert e.value.code == -2
    with pytest.raises(pm.PoutreError) as e:
        pm.dilate(a, pm.static_nl_se.square2d, 1, out=a)
    assert e.value.code == -3
    with pytest.raises(pm.PoutreError) as e:
        pm.dilate(a, pm.static_nl_se.square3d, 1)
    assert e.value.code == -4
    with pytest.raises(pm.PoutreError) as e:
        pm.dilate(a, 42, 1)
    assert e.value.code == -5
    with pytest.raises(pm.PoutreError) as e:
        pm.dilate(a, pm.static_nl_se.square2d, -1)
    assert e.value.code == -6
    with pytest.raises(pm.PoutreError):
        pm.dilateY(g(rand((10,), np.uint8, 2)), 1)


# ----------------------------------------------------------------------------- fused uint8 square kernel
@pytest.mark.parametrize("shape", [(256, 272), (130, 48), (517, 1008), (16, 16), (300, 16),
                                   (131, 37), (77, 1001), (3, 5), (129, 130), (260, 259)])
@pytest.mark.parametrize("r", [1, 2, 3, 4, 7, 8, 15, 16, 31, 32, 33, 70])
def test_square_u8_fused(shape, r):
    img = rand(shape, np.uint8, 60 + r)
    for dil in (True, False):
        f = pm.dilate if dil else pm.erode
        exp = oracle.static_iter(oracle.line_x(img, r, dil), "SESegmentY2D", min(r, shape[0]), dil)
        np.testing.assert_array_equal(h(f(g(img), pm.static_nl_se.square2d, r)), exp)


@pytest.mark.parametrize("shape", [(256, 272), (130, 48), (77, 1001), (3, 5), (129, 130), (300, 16), (20, 33)])
@pytest.mark.parametrize("r", [1, 2, 3, 4, 7, 8, 15, 16, 31, 32, 33])
def test_gradient_u8_fused(shape, r):
    """uint8 gradient by a square (one fused launch for r <= 32) == dilation - erosion."""
    img = rand(shape, np.uint8, 80 + r)
    d = oracle.static_iter(oracle.line_x(img, r, True), "SESegmentY2D", min(r, shape[0]), True)
    e = oracle.static_iter(oracle.line_x(img, r, False), "SESegmentY2D", min(r, shape[0]), False)
    np.testing.assert_array_equal(h(pm.gradient(g(img), pm.static_nl_se.square2d, r)), oracle.ssub_image(d, e))


def test_gradient_u8_fused_unaligned_view():
    base = torch.from_numpy(rand((70 * 61 + 8,), np.uint8, 90)).to(DEV)
    view = base[3:3 + 70 * 61].view(70, 61)
    obuf = torch.full((70 * 61 + 8,), 5, dtype=torch.uint8, device=DEV)
    oview = obuf[1:1 + 70 * 61].view(70, 61)
    img = view.cpu().numpy()
    pm.gradient(view, pm.static_nl_se.square2d, 2, out=oview)
    d = oracle.static_iter(img, "SESquare2D", 2, True)
    e = oracle.static_iter(img, "SESquare2D", 2, False)
    np.testing.assert_array_equal(h(oview), oracle.ssub_image(d, e))
    rest = obuf.cpu().numpy()
    assert rest[0] == 5 and (rest[1 + 70 * 61:] == 5).all()


def _sq_oracle(img, r, dil):
    return oracle.static_iter(oracle.line_x(img, r, dil), "SESegmentY2D", min(r, img.shape[0]), dil)


@pytest.mark.parametrize("shape", [(256, 272), (130, 48), (517, 1008), (77, 1001), (3, 5), (129, 130),
                                   (300, 16), (20, 33), (700, 520), (125, 257), (16, 4096)])
@pytest.mark.parametrize("r", [1, 2, 3, 4, 5, 7, 8, 9, 12, 15, 16])
def test_open_close_u8_fused(shape, r):
    """uint8 opening / closing by a square (one fused launch for r <= 7: the band of the
    first operator stays in LDS) == the two-operator composition of the oracle."""
    img = rand(shape, np.uint8, 100 + r)
    e, d = _sq_oracle(img, r, False), _sq_oracle(img, r, True)
    np.testing.assert_array_equal(h(pm.opening(g(img), pm.static_nl_se.square2d, r)), _sq_oracle(e, r, True))
    np.testing.assert_array_equal(h(pm.closing(g(img), pm.static_nl_se.square2d, r)), _sq_oracle(d, r, False))


@pytest.mark.parametrize("ioff,ooff", [(3, 1), (0, 2), (4, 0)])
def test_open_close_u8_fused_unaligned_view(ioff, ooff):
    H, W = 150, 301
    base = torch.from_numpy(rand((H * W + 8,), np.uint8, 110)).to(DEV)
    view = base[ioff:ioff + H * W].view(H, W)
    img = view.cpu().numpy()
    for opening in (True, False):
        obuf = torch.full((H * W + 8,), 5, dtype=torch.uint8, device=DEV)
        oview = obuf[ooff:ooff + H * W].view(H, W)
        (pm.opening if opening else pm.closing)(view, pm.static_nl_se.square2d, 3, out=oview)
        first = _sq_oracle(img, 3, not opening)
        np.testing.assert_array_equal(h(oview), _sq_oracle(first, 3, opening))
        rest = obuf.cpu().numpy()
        assert (rest[:ooff] == 5).all() and (rest[ooff + H * W:] == 5).all()


@pytest.mark.parametrize("shape", [(1, 16), (2, 32), (3, 1024), (9, 1040), (17, 2064), (70, 4096), (257, 160),
                                   (8, 1008), (16, 1024), (25, 3072)])
def test_square_u8_streamed_3x3(shape):
    """uint8 3x3 square, opening, closing and gradient on rows of a multiple of 16 bytes (the
    register-streamed kernels: bands of 8 rows, 64 chunks of 16 bytes per wave): partial bands,
    partial and single-chunk wave column groups, the words beside each wave's chunks, image
    borders (where the first operator of a composition is padded with the second's neutral)"""
    img = rand(shape, np.uint8, 140 + shape[0])
    d = oracle.static_iter(img, "SESquare2D", 1, True)
    e = oracle.static_iter(img, "SESquare2D", 1, False)
    for dil in (True, False):
        f = pm.dilate if dil else pm.erode
        np.testing.assert_array_equal(h(f(g(img), pm.static_nl_se.square2d, 1)), d if dil else e)
    # the compositions on the same streams (two operator stages, or max and min at once)
    np.testing.assert_array_equal(h(pm.opening(g(img), pm.static_nl_se.square2d, 1)),
                                  oracle.static_iter(e, "SESquare2D", 1, True))
    np.testing.assert_array_equal(h(pm.closing(g(img), pm.static_nl_se.square2d, 1)),
                                  oracle.static_iter(d, "SESquare2D", 1, False))
    np.testing.assert_array_equal(h(pm.gradient(g(img), pm.static_nl_se.square2d, 1)), d - e)


@pytest.mark.parametrize("dtype", [np.uint16, np.int32, np.int64, np.float32, np.float64])
@pytest.mark.parametrize("shape", [(64, 64), (65, 70), (200, 130), (3, 5), (130, 257), (1, 300), (300, 1),
                                   (72, 136)])
@pytest.mark.parametrize("r", [1, 2, 3, 4, 7, 8, 9, 15, 16, 17, 40])
def test_square_generic_fused(dtype, shape, r):
    """uint16 / int32 / int64 / float32 / float64 square (one fused launch per radius up to 16,
    chained launches above) == the oracle's x line pass then r y segments.  (72, 136): rows a
    multiple of 16 bytes for every type, partial tiles on both axes.  uint16 takes its full
    range, so that both bytes of every pixel take part."""
    img = rand(shape, dtype, 120 + r, hi=65536 if dtype == np.uint16 else 256)
    for dil in (True, False):
        f = pm.dilate if dil else pm.erode
        np.testing.assert_array_equal(h(f(g(img), pm.static_nl_se.square2d, r)), _sq_oracle(img, r, dil))


@pytest.mark.parametrize("dtype", [np.uint16, np.int32, np.float64])
@pytest.mark.parametrize("W,r", [(93, 5), (96, 1), (96, 12), (7, 2)])
def test_square_generic_fused_views(dtype, W, r):
    """the fused square kernels on views at element offsets (no alignment assumed; W = 96
    makes rows a multiple of 16 bytes on an unaligned base, W = 7 rows narrower than 16
    bytes); the elements around the output view are left untouched"""
    H = 70
    base = torch.from_numpy(rand((H * W + 5,), dtype, 130, hi=65536 if dtype == np.uint16 else 256)).to(DEV)
    view = base[3:3 + H * W].view(H, W)
    img = view.cpu().numpy()
    for dil in (True, False):
        obuf = torch.zeros((H * W + 5,), dtype=base.dtype, device=DEV)
        oview = obuf[1:1 + H * W].view(H, W)
        (pm.dilate if dil else pm.erode)(view, pm.static_nl_se.square2d, r, out=oview)
        np.testing.assert_array_equal(h(oview), _sq_oracle(img, r, dil))
        rest = obuf.cpu().numpy()
        assert rest[0] == 0 and (rest[1 + H * W:] == 0).all()


@pytest.mark.parametrize("it", [1, 2, 5])
def test_square3d_u8_fused(it):
    img = rand((6, 40, 48), np.uint8, 70)
    for dil in (True, False):
        f = pm.dilate if dil else pm.erode
        np.testing.assert_array_equal(h(f(g(img), pm.static_nl_se.square3d, it)),
                                      oracle.static_iter(img, "SESquare3D", it, dil))


@pytest.mark.parametrize("w", [64, 61, 150])
@pytest.mark.parametrize("off_in,off_out", [(1, 0), (2, 3), (3, 1), (0, 2), (5, 7)])
@pytest.mark.parametrize("r", [1, 3, 9, 20])
def test_square_u8_unaligned_views(w, off_in, off_out, r):
    """Input and output views at any byte offset (rows not 4- or 16-byte aligned), same bits;
    the bytes of the output buffer around the view are left untouched."""
    H = 67
    base = torch.from_numpy(rand((H + 1, w), np.uint8, 71)).to(DEV)
    view = base.view(-1)[off_in:off_in + H * w].view(H, w)
    img = view.cpu().numpy()
    obuf = torch.full(((H + 1) * w,), 77, dtype=torch.uint8, device=DEV)
    oview = obuf[off_out:off_out + H * w].view(H, w)
    for dil in (True, False):
        f = pm.dilate if dil else pm.erode
        f(view, pm.static_nl_se.square2d, r, out=oview)
        np.testing.assert_array_equal(h(oview), oracle.static_iter(img, "SESquare2D", r, dil))
        rest = obuf.cpu().numpy()
        assert (rest[:off_out] == 77).all() and (rest[off_out + H * w:] == 77).all()


# ----------------------------------------------------------------------------- labelling
@pytest.mark.parametrize("se,shape", [("SESquare2D", (300, 517)), ("SECross2D", (129, 257)), ("SESegmentX2D", (64, 90)),
                                      ("SESegmentY2D", (90, 64)), ("SESegmentX1D", (5000,)), ("SECross3D", (20, 33, 41)),
                                      ("SESquare3D", (17, 30, 29)), ("SESegmentZ3D", (9, 10, 11))])
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.int32, np.int64, np.float32, np.float64])
def test_label_binary_and_flat_zones(se, shape, dtype):
    rng = np.random.default_rng(200)
    if np.dtype(dtype).kind == "i":
        img = rng.integers(-1, 2, size=shape).astype(dtype)  # negatives: foreground not linkable from below
    else:
        img = (rng.random(shape) < 0.5).astype(dtype)
    out, n = pm.label_binary(g(img), PM_SE[se])
    exp, m = oracle.label(img, se)
    np.testing.assert_array_equal(h(out), exp)
    assert n == m
    zones = rng.integers(0, 4, size=shape).astype(dtype)
    out, n = pm.label_flat_zones(g(zones), PM_SE[se])
    exp, m = oracle.label(zones, se, flat=True)
    np.testing.assert_array_equal(h(out), exp)
    assert n == m


@pytest.mark.parametrize("shape", [(1024, 1024), (1, 1), (3, 2049), (700, 3)])
def test_label_binary_large_blobs(shape):
    """long-range components (thresholded blurred noise: a few big, winding blobs) and
    degenerate shapes"""
    from poutre2_amd import synth
    if min(shape) >= 64:
        img = (synth.blurred_noise_mask(shape[0], shape[1], 9) > 127).astype(np.uint8)
    else:
        img = (np.random.default_rng(9).random(shape) < 0.6).astype(np.uint8)
    for se in ("SESquare2D", "SECross2D"):
        out, n = pm.label_binary(g(img), PM_SE[se])
        exp, m = oracle.label(img, se)
        np.testing.assert_array_equal(h(out), exp)
        assert n == m


def test_label_bench_workload_4096():
    """the bench's labelling workload at its size (4096^2 blurred-noise mask, seed 4,
    thresholded at 128 and quantised to 8 levels): binary by SESquare2D and SECross2D and flat
    zones, equal to the oracle, each three times (the union-find's plain loads race with its
    atomicMin updates by design; every run must give the same labels)"""
    from poutre2_amd import synth
    mask = h(synth.blurred_noise_mask(4096, 4096, 4, device=DEV))
    binary = (mask > 127).astype(np.uint8)
    zones = (mask // 32).astype(np.uint8)
    for img, fn, se, flat in ((binary, pm.label_binary, "SESquare2D", False),
                              (binary, pm.label_binary, "SECross2D", False),
                              (zones, pm.label_flat_zones, "SESquare2D", True)):
        exp, m = oracle.label(img, se, flat=flat)
        gi = g(img)
        for _ in range(3):
            out, n = fn(gi, PM_SE[se])
            assert n == m, (se, flat, n, m)
            np.testing.assert_array_equal(h(out), exp)


def _serpentine(H, W):
    """one 1-pixel-wide path through every row pair: rows 0, 2, 4, ... are foreground and
    joined alternately at the right and the left end"""
    img = np.zeros((H, W), np.uint8)
    img[0::2, :] = 1
    for y in range(1, H, 2):
        img[y, W - 1 if (y // 2) % 2 == 0 else 0] = 1
    return img


@pytest.mark.parametrize("pattern,shape", [("serpentine", (1001, 2100)), ("comb", (900, 4100)),
                                           ("checker", (777, 1100)), ("blobs", (2048, 4100))])
def test_label_chains_across_tiles(pattern, shape):
    """components that cross many tiles in both directions (every border level and partial
    tiles): one path snaking through the whole image, vertical bars joined by one row at
    the bottom, a checkerboard (one diagonal component by SESquare2D, isolated pixels by
    SECross2D), large blobs; binary and flat zones"""
    from poutre2_amd import synth
    H, W = shape
    if pattern == "serpentine":
        img = _serpentine(H, W)
    elif pattern == "comb":
        img = np.zeros(shape, np.uint8)
        img[:, ::3] = 1
        img[H - 1, :] = 1
    elif pattern == "checker":
        img = ((np.arange(H)[:, None] + np.arange(W)[None, :]) % 2).astype(np.uint8)
    else:
        img = (synth.blurred_noise_mask(H, W, 11) > 127).astype(np.uint8)
    for se in ("SESquare2D", "SECross2D"):
        out, n = pm.label_binary(g(img), PM_SE[se])
        exp, m = oracle.label(img, se)
        assert n == m, (se, n, m)
        np.testing.assert_array_equal(h(out), exp)
    zones = (img * 3 + (np.arange(W)[None, :] // 700).astype(np.uint8)).astype(np.uint8)
    out, n = pm.label_flat_zones(g(zones), PM_SE["SESquare2D"])
    exp, m = oracle.label(zones, "SESquare2D", flat=True)
    assert n == m
    np.testing.assert_array_equal(h(out), exp)


@pytest.mark.parametrize("op", ["label_minima", "label_maxima"])
def test_label_extrema(op):
    img = rand((120, 97), np.int32, 210, 0, 6)
    got = h(getattr(pm, op)(g(img), PM_SE["SESquare2D"]))
    f = oracle.h_concave if op == "label_minima" else oracle.h_convex
    exp, _ = oracle.label(f(img, 1, "SESquare2D"), "SESquare2D")
    np.testing.assert_array_equal(got, exp)


def test_label_errors():
    a = g(rand((8, 9), np.uint8, 1))
    with pytest.raises(pm.PoutreError):
        pm.label_binary(a, PM_SE["SESquare2D"], out=g(np.zeros((8, 9), np.int32)))  # output must be int64
    with pytest.raises(pm.PoutreError):
        pm.label_binary(a, PM_SE["SESquare2D"], out=g(np.zeros((8, 8), np.int64)))
    with pytest.raises(pm.PoutreError):
        pm.label_binary(a, PM_SE["SECross3D"])


class _DLPackOnly:
    """an image that only exports DLPack (as a CuPy or JAX device array would)"""

    def __init__(self, t):
        self._t = t

    def __dlpack__(self, *args, **kw):
        return self._t.__dlpack__(*args, **kw)

    def __dlpack_device__(self):
        return self._t.__dlpack_device__()


def test_dlpack_images_zero_copy():
    """DLPack producers as inputs and outputs (used in place) and a factory_image output"""
    img = rand((300, 517), np.uint8, 301)
    out = pm.factory_image(img.shape, pm.CompoundType.Scalar, pm.PType.GrayUINT8)
    r = pm.dilate(_DLPackOnly(g(img)), PM_SE["SESquare2D"], 3, out=_DLPackOnly(out))
    assert isinstance(r, _DLPackOnly)
    np.testing.assert_array_equal(h(out), oracle.static_iter(img, "SESquare2D", 3, True))
    fg = (img > 128).astype(np.uint8)
    lab, n = pm.label_binary(_DLPackOnly(g(fg)), PM_SE["SESquare2D"])
    exp, m = oracle.label(fg, "SESquare2D")
    assert n == m
    np.testing.assert_array_equal(h(lab), exp)
    assert pm.to_string(pm.from_string("Scalar GUINT8 2 2 2 1 2 3 4")) == "Scalar GUINT8 2 2 2 1 2 3 4"


# ----------------------------------------------------------------------------- CompareImage / arith
from tests.test_oracle import COMPARE_KATS, compare_operands  # noqa: E402


@pytest.mark.parametrize("c", COMPARE_KATS, ids=lambda c: f"compare.cpp:{c['line']}:{c['name']}")
def test_compare_kats(c):
    exp = kats.image(c["expected"])
    ops = [g(o) if isinstance(o, np.ndarray) else o for o in compare_operands(c)]
    out = g(np.zeros_like(exp))
    got = pm.compare(g(kats.image(c["input"])), pm.CompOpType[c["op"]], *ops, out=out)
    np.testing.assert_array_equal(h(got), exp)


@pytest.mark.parametrize("tin,tout", [(np.uint8, np.uint8), (np.float32, np.uint8), (np.int64, np.float64),
                                      (np.uint16, np.int32)])
@pytest.mark.parametrize("form", ["sss", "iss", "sis", "ssi", "sii", "iii", "iis", "isi"])
def test_compare_all_ops_forms_types(tin, tout, form):
    shape = (37, 259)
    a = rand(shape, tin, 400, 0, 6)
    imgs = {"c": rand(shape, tin, 401, 0, 6), "t": rand(shape, tout, 402, 0, 100), "f": rand(shape, tout, 403, 0, 100)}
    scal = {"c": 3, "t": 7, "f": 2}
    opnds = [imgs[k] if kind == "i" else scal[k] for k, kind in zip("ctf", form)]
    for op in pm.CompOpType:
        out = g(np.zeros(shape, tout))
        pm.compare(g(a), op, *[g(o) if isinstance(o, np.ndarray) else o for o in opnds], out=out)
        np.testing.assert_array_equal(h(out), oracle.compare(a, int(op), *opnds, tout), err_msg=str(op))


def test_compare_nan_inplace_and_errors():
    x = np.array([[np.nan, 1.0, 2.0, -np.inf, np.inf, 0.0]], np.float32)
    y = np.array([[1.0, np.nan, 2.0, -np.inf, 0.0, -0.0]], np.float32)
    for op in pm.CompOpType:
        got = pm.compare(g(x), op, g(y), 1.5, -1.0)
        np.testing.assert_array_equal(h(got), oracle.compare(x, int(op), y, 1.5, -1.0, np.float32))
    a = rand((50, 60), np.int32, 404)
    t = g(a)
    pm.compare(t, pm.CompOpType.Sup, 0, t, 0, out=t)  # out is the input and the true image
    np.testing.assert_array_equal(h(t), np.where(a > 0, a, 0))
    with pytest.raises(pm.PoutreError):
        pm.compare(g(a), 3, 0, 1, 0)  # not a single CompOpType bit
    with pytest.raises(pm.PoutreError):
        pm.compare(g(a), pm.CompOpType.Equal, g(a[:, :10].copy()), 1, 0)
    with pytest.raises(pm.PoutreError):
        pm.compare(g(a), pm.CompOpType.Equal, g(a.astype(np.float32)), 1, 0)  # comparand of another type


@pytest.mark.parametrize("dtype", DTYPES)
def test_sadd_image_and_invert(dtype):
    a = rand((130, 257), dtype, 405)
    b = rand((130, 257), dtype, 406)
    if np.dtype(dtype).kind == "i":
        a, b = np.abs(a // 2).astype(dtype), (b // 2).astype(dtype)  # sums stay above the type's minimum
    np.testing.assert_array_equal(h(pm.arith_saturated_add_image(g(a), g(b))), oracle.sadd_image(a, b))
    np.testing.assert_array_equal(h(pm.arith_invert_image(g(a))), oracle.invert(a))
    t = g(a)
    pm.arith_invert(t, out=t)
    np.testing.assert_array_equal(h(t), oracle.invert(a))


@pytest.mark.parametrize("tin", DTYPES)
def test_convert_copy_clone(tin):
    """ConvertInto / Convert / Clone / CopyInto (copy_convert.hpp:31-52): a C cast per pixel
    (numpy's astype is one) for values every type holds; negative integers into unsigned types
    wrap modulo 2^bits"""
    r = np.random.default_rng(407)
    a = (r.integers(-100, 200, size=(33, 70)) if np.dtype(tin).kind == "i" else r.integers(0, 200, size=(33, 70)))
    a = (a + r.random((33, 70)) * 0.9 if np.dtype(tin).kind == "f" else a).astype(tin)
    for tout, pt in ((np.uint8, pm.PType.GrayUINT8), (np.uint16, pm.PType.GrayUINT16), (np.int32, pm.PType.GrayINT32),
                     (np.int64, pm.PType.GrayINT64), (np.float32, pm.PType.F32), (np.float64, pm.PType.D64)):
        if np.dtype(tin).kind == "f" and np.dtype(tout).kind in "iu":
            src = np.abs(a)  # float to integer: values the integer type holds
        else:
            src = a
        got = pm.convert(g(src), pm.CompoundType.Scalar, pt)
        np.testing.assert_array_equal(h(got), src.astype(tout), err_msg=f"{tin} -> {tout}")
    c = pm.clone(g(a))
    np.testing.assert_array_equal(h(c), a)
    z = pm.clone_geometry(g(a))
    assert z.shape == a.shape and z.dtype == c.dtype and int((z != 0).sum()) == 0
    d = pm.factory_image(a.shape, pm.CompoundType.Scalar, pm.PType(pm._ptype(c)))
    pm.copy_into(c, d)
    np.testing.assert_array_equal(h(d), a)
    with pytest.raises(pm.PoutreError):
        pm.copy_into(c, g(np.zeros((33, 71), tin)))
    with pytest.raises(pm.PoutreError):
        pm.convert_into(c, g(np.zeros((33, 71), np.float64)))


def test_se_object_is_a_runtime_neighbour_list():
    img = rand((40, 50), np.int32, 408)
    se = pm.SE([(0, -1), (0, 0), (0, 1), (-1, 2)])
    for dil in (True, False):
        f = pm.dilate if dil else pm.erode
        np.testing.assert_array_equal(h(f(g(img), se)), oracle.nl(img, list(se), dil))
