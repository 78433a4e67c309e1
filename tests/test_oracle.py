"""Pins the CPU oracle before anything is compared against it.

1. Every known-answer test of the reference for this path (tests/golden/kats.json).
2. Independent checks the mathematics fixes: brute-force window max/min, composition
   identities of iterated structuring elements, the brute-force fixed point of
   reconstruction, the closed forms of leveling, saturating arithmetic.
Each check is chosen so that a plausible mistake in the oracle (off-by-one window,
wrong neutral, non-causal scan split, missing clamp) fails one of them.
"""
import numpy as np
import pytest

import oracle
from tests import kats


def _run_kat(c):
    ins = [kats.image(d) for d in c["inputs"]]
    op = c["op"]
    if op in ("t_Dilate", "t_Erode"):
        return oracle.static(ins[0], c["se"], dilate=(op == "t_Dilate"))
    if op in ("t_DilateX", "t_ErodeX"):
        return oracle.line_x(ins[0], c["r"], dilate=(op == "t_DilateX"))
    if op in ("t_DilateY", "t_ErodeY"):
        return oracle.line_y(ins[0], c["r"], dilate=(op == "t_DilateY"))
    if op == "Reconstruction":
        return oracle.reconstruct(c["rec_type"], ins[0], ins[1], c["se"])
    if op == "low_leveling":
        return oracle.low_leveling(ins[0], ins[1], c["se"])
    if op == "high_leveling":
        return oracle.high_leveling(ins[0], ins[1], c["se"])
    if op == "leveling":
        return oracle.leveling(ins[0], ins[1], c["se"])
    if op in ("h_minima", "h_maxima", "h_concave", "h_convex", "dynamic_pseudo_closing",
              "dynamic_pseudo_opening"):
        return getattr(oracle, op)(ins[0], c["h"], c["se"])
    if op == "t_transpose":
        return oracle.transpose(ins[0])
    if op in ("label_binary", "label_flat_zones"):
        return oracle.label(ins[0], c["se"], flat=(op == "label_flat_zones"))[0]
    if op == "label_minima":  # src/label/label_extrema.cpp: label_binary(h_concave(img, 1))
        return oracle.label(oracle.h_concave(ins[0], 1, c["se"]), c["se"])[0]
    if op == "label_maxima":  # label_binary(h_convex(img, 1))
        return oracle.label(oracle.h_convex(ins[0], 1, c["se"]), c["se"])[0]
    raise AssertionError(op)


@pytest.mark.parametrize("case", kats.cases(), ids=kats.case_id)
def test_oracle_matches_reference_kat(case):
    got = _run_kat(case)
    exp = kats.image(case["expected"])
    assert got.shape == exp.shape
    np.testing.assert_array_equal(got, exp)


def test_kat_count():
    # 20 static + 20 runtime + 6 line + 4 reconstruction + 4 leveling + 6 extrema + 1 transpose
    # + 4 label_binary + 2 label_flat_zones + 2 label_minima/maxima
    assert len(kats.cases()) == 69


# ---------------------------------------------------------------- labelling pins
def _label_brute(img, se, flat):
    """Connected components of the link graph by breadth-first search, numbered by the
    raster position of each component's first pixel -- an independent statement of what the
    one-pass union-find of the reference computes (links: a pixel to its raster-earlier SE
    neighbours; binary: both pixels != 0 and the earlier one > 0; flat zones: equal values)."""
    offs = [tuple(o) for o in oracle.se_offsets(se)]
    upper = [o for o in offs if next((x for x in o if x != 0), 0) < 0]
    shp = img.shape
    flat_img = img.reshape(-1)
    N = flat_img.size

    def acc(v):
        return True if flat else (abs(v) > 1e-10 if img.dtype.kind == "f" else v != 0)

    def link(v, w):  # v at the later pixel, w at the earlier one
        return (w == v) if flat else (acc(v) and acc(w) and w > 0)
    adj = [[] for _ in range(N)]
    for i in range(N):
        p = np.unravel_index(i, shp)
        for o in upper:
            q = tuple(a + b for a, b in zip(p, o))
            if all(0 <= c < d for c, d in zip(q, shp)):
                j = int(np.ravel_multi_index(q, shp))
                if link(flat_img[i], flat_img[j]):
                    adj[i].append(j)
                    adj[j].append(i)
    out = np.zeros(N, dtype=np.int64)
    n = 0
    for i in range(N):
        if out[i] or not acc(flat_img[i]):
            continue
        n += 1
        out[i] = n
        stack = [i]
        while stack:
            a = stack.pop()
            for b in adj[a]:
                if not out[b]:
                    out[b] = n
                    stack.append(b)
    return out.reshape(shp), n


@pytest.mark.parametrize("se,shape", [("SESquare2D", (9, 13)), ("SECross2D", (12, 7)), ("SESegmentX2D", (5, 9)),
                                      ("SESegmentY2D", (8, 6)), ("SESegmentX1D", (40,)), ("SECross3D", (4, 5, 6)),
                                      ("SESquare3D", (3, 6, 5)), ("SESegmentZ3D", (5, 3, 4))])
@pytest.mark.parametrize("dtype", [np.uint8, np.int32, np.float64])
def test_label_matches_components(se, shape, dtype):
    rng = np.random.default_rng(11)
    for trial in range(4):
        if dtype == np.int32:
            img = rng.integers(-1, 2, size=shape).astype(dtype)  # negatives exercise the one-sided link
        else:
            img = (rng.random(shape) < 0.45).astype(dtype)
        got, n = oracle.label(img, se)
        exp, m = _label_brute(img, se, False)
        np.testing.assert_array_equal(got, exp)
        assert n == m
        zones = rng.integers(0, 3, size=shape).astype(dtype)
        got, n = oracle.label(zones, se, flat=True)
        exp, m = _label_brute(zones, se, True)
        np.testing.assert_array_equal(got, exp)
        assert n == m


# ---------------------------------------------------------------- brute-force pins
def _window_brute(img, ry, rx, dilate):
    """max/min over the in-bounds part of a (2ry+1)x(2rx+1) window, numpy, no padding trick."""
    H, W = img.shape
    out = np.empty_like(img)
    f = np.max if dilate else np.min
    for y in range(H):
        for x in range(W):
            out[y, x] = f(img[max(0, y - ry):y + ry + 1, max(0, x - rx):x + rx + 1])
    return out


RNG = np.random.default_rng(1234)


@pytest.mark.parametrize("dtype", [np.uint8, np.int32, np.float32, np.int64, np.uint16, np.float64])
@pytest.mark.parametrize("shape,r", [((7, 13), 1), ((9, 5), 2), ((6, 17), 4), ((3, 4), 9), ((1, 1), 3)])
def test_line_x_brute(dtype, shape, r):
    img = (RNG.integers(0, 200, size=shape) - (50 if np.dtype(dtype).kind in "if" else 0)).astype(dtype)
    for d in (True, False):
        np.testing.assert_array_equal(oracle.line_x(img, r, d), _window_brute(img, 0, r, d))


@pytest.mark.parametrize("n,r", [(11, 2), (12, 5), (3, 7)])
def test_line_y_brute_and_square_only(n, r):
    img = RNG.integers(0, 256, size=(n, n)).astype(np.uint8)
    for d in (True, False):
        np.testing.assert_array_equal(oracle.line_y(img, r, d), _window_brute(img, r, 0, d))
    with pytest.raises(oracle.OracleError):
        oracle.line_y(RNG.integers(0, 256, size=(n, n + 1)).astype(np.uint8), r)


def test_line_r0_copies():
    img = RNG.integers(0, 256, size=(5, 6)).astype(np.uint8)
    np.testing.assert_array_equal(oracle.line_x(img, 0), img)
    np.testing.assert_array_equal(oracle.line_x(img, -3, False), img)


@pytest.mark.parametrize("se", ["SESquare2D", "SECross2D", "SESegmentX2D", "SESegmentY2D"])
def test_static_2d_brute(se):
    img = RNG.integers(0, 256, size=(9, 11)).astype(np.uint8)
    offs = oracle.se_offsets(se)
    H, W = img.shape
    for d in (True, False):
        exp = np.empty_like(img)
        for y in range(H):
            for x in range(W):
                v = [img[y + dy, x + dx] for dy, dx in offs if 0 <= y + dy < H and 0 <= x + dx < W]
                exp[y, x] = max(v) if d else min(v)
        np.testing.assert_array_equal(oracle.static(img, se, d), exp)


def test_se_geometry():
    sq = {tuple(o) for o in oracle.se_offsets("SESquare2D")}
    assert sq == {(dy, dx) for dy in (-1, 0, 1) for dx in (-1, 0, 1)}
    cr = {tuple(o) for o in oracle.se_offsets("SECross2D")}
    assert cr == {(0, 0), (1, 0), (-1, 0), (0, 1), (0, -1)}
    assert len(oracle.se_offsets("SESquare3D")) == 27
    assert {tuple(o) for o in oracle.se_offsets("SECross3D")} == {
        (0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)}
    assert {tuple(o) for o in oracle.se_offsets("SESegmentZ3D")} == {(0, 0, 0), (1, 0, 0), (-1, 0, 0)}
    assert {tuple(o) for o in oracle.se_offsets("SESegmentX1D")} == {(0,), (1,), (-1,)}


@pytest.mark.parametrize("it", [0, 1, 2, 3, 4])
def test_static_iter_square_equals_window(it):
    """Iterated 3x3 square on a box = (2it+1)^2 window (L-inf geodesic distance = L-inf distance)."""
    img = RNG.integers(0, 256, size=(13, 10)).astype(np.uint8)
    for d in (True, False):
        np.testing.assert_array_equal(oracle.static_iter(img, "SESquare2D", it, d), _window_brute(img, it, it, d))


@pytest.mark.parametrize("it", [2, 3])
def test_static_iter_cross_equals_diamond(it):
    img = RNG.integers(0, 256, size=(11, 12)).astype(np.int32)
    diamond = [(dy, dx) for dy in range(-it, it + 1) for dx in range(-it, it + 1) if abs(dy) + abs(dx) <= it]
    for d in (True, False):
        np.testing.assert_array_equal(oracle.static_iter(img, "SECross2D", it, d), oracle.nl(img, diamond, d))


def test_static_iter_negative_rejected():
    img = np.zeros((3, 3), np.uint8)
    with pytest.raises(oracle.OracleError):
        oracle.static_iter(img, "SESquare2D", -1)


# Number of square dilations m in the compound SE of size n: the rule of
# ero_dil_compound_static_se_t.hpp:47-50 (m = round-half-up(n / (1 + sqrt 2))), written out
# by hand for n = 2..12 (n / 2.41421356...: 0.83 1.24 1.66 2.07 2.49 2.90 3.31 3.73 4.14 4.56 4.97).
_COMPOUND_SQUARES = {2: 1, 3: 1, 4: 2, 5: 2, 6: 2, 7: 3, 8: 3, 9: 4, 10: 4, 11: 5, 12: 5}


def _minkowski(a, b):
    return sorted({tuple(x + y for x, y in zip(p, q)) for p in a for q in b})


def _ball(rank, r, norm):
    import itertools
    pts = itertools.product(range(-r, r + 1), repeat=rank)
    if norm == 1:
        return [p for p in pts if sum(abs(c) for c in p) <= r]
    return list(pts)


@pytest.mark.parametrize("n", range(0, 13))
def test_compound_octagon_is_minkowski_sum(n):
    """The octagon of size n (ero_dil_compound_static_se_t.hpp:33-65) is the cross applied
    n - m times then the square m times; on a box domain with out-of-image taps ignored,
    chained dilations by these L1 / L-inf balls equal ONE dilation by their Minkowski sum
    diamond(n - m) + square(m), taken here over the explicitly enumerated offsets by the
    KAT-pinned neighbour-list brute force.  n = 1 is the cross itself (:42-45), n = 0 a copy."""
    img = RNG.integers(0, 256, size=(23, 29)).astype(np.uint8)
    if n == 0:
        offs = [(0, 0)]
    elif n == 1:
        offs = _ball(2, 1, 1)
    else:
        m = _COMPOUND_SQUARES[n]
        offs = _minkowski(_ball(2, n - m, 1), _ball(2, m, 0))
    for d in (True, False):
        np.testing.assert_array_equal(oracle.compound(img, 1, n, d), oracle.nl(img, offs, d))


@pytest.mark.parametrize("n", [2, 3, 4, 5])
def test_compound_rhombicuboctahedron_is_minkowski_sum(n):
    """3D analogue (:67-99): SECross3D n - m times, SESquare3D m times = one dilation by
    octahedron(n - m) + cube(m)."""
    img = RNG.integers(0, 256, size=(9, 11, 10)).astype(np.int32)
    m = _COMPOUND_SQUARES[n]
    offs = _minkowski(_ball(3, n - m, 1), _ball(3, m, 0))
    for d in (True, False):
        np.testing.assert_array_equal(oracle.compound(img, 2, n, d), oracle.nl(img, offs, d))


def test_compound_errors():
    """size 1 on a volume applies SECross2D to a 3D image, which throws (:77-79); each compound
    SE needs its own rank (ero_dil_compound_static.cpp:68-70,134-136); negative size rejected."""
    with pytest.raises(oracle.OracleError):
        oracle.compound(np.zeros((3, 4, 5), np.uint8), 2, 1)
    with pytest.raises(oracle.OracleError):
        oracle.compound(np.zeros((3, 4, 5), np.uint8), 1, 2)
    with pytest.raises(oracle.OracleError):
        oracle.compound(np.zeros((4, 5), np.uint8), 2, 2)
    with pytest.raises(oracle.OracleError):
        oracle.compound(np.zeros((4, 5), np.uint8), 1, -1)


def test_static_3d_vs_window():
    img = RNG.integers(0, 256, size=(5, 6, 7)).astype(np.uint8)
    out = oracle.static(img, "SESquare3D", True)
    D, H, W = img.shape
    exp = np.empty_like(img)
    for z in range(D):
        for y in range(H):
            for x in range(W):
                exp[z, y, x] = img[max(0, z - 1):z + 2, max(0, y - 1):y + 2, max(0, x - 1):x + 2].max()
    np.testing.assert_array_equal(out, exp)


# ---------------------------------------------------------------- reconstruction pins
def _rand_pair(shape, dtype, rng, hi=50):
    mask = rng.integers(0, hi, size=shape).astype(dtype)
    marker = rng.integers(0, hi, size=shape).astype(dtype)
    return marker, mask


@pytest.mark.parametrize("dtype", [np.uint8, np.int32, np.int64])
@pytest.mark.parametrize("se,shape", [("SESquare2D", (17, 23)), ("SECross2D", (19, 14)),
                                      ("SESegmentX1D", (40,)), ("SECross3D", (5, 6, 7)),
                                      ("SESquare3D", (4, 5, 6))])
@pytest.mark.parametrize("rec", ["dilate", "erode"])
def test_reconstruct_hybrid_equals_fixed_point(dtype, se, shape, rec):
    rng = np.random.default_rng(7)
    for _ in range(3):
        marker, mask = _rand_pair(shape, dtype, rng)
        # sparse marker, like an h-maxima seed
        if rec == "dilate":
            marker = np.where(rng.random(shape) < 0.05, marker, 0).astype(dtype)
        hyb = oracle.reconstruct(rec, marker, mask, se)
        bf = oracle.reconstruct_bf(rec, marker, mask, se)
        np.testing.assert_array_equal(hyb, bf)
        lo = np.minimum(marker, mask) if rec == "dilate" else mask
        hi = mask if rec == "dilate" else np.maximum(marker, mask)
        assert (hyb >= lo).all() and (hyb <= hi).all()
        np.testing.assert_array_equal(oracle.reconstruct(rec, hyb, mask, se), hyb)  # idempotent


def test_reconstruct_bf_closed_form_plateau():
    """A single seed inside a connected plateau floods it and nothing else (8- vs 4-connectivity)."""
    mask = np.zeros((5, 5), np.uint8)
    mask[1, 1] = mask[2, 2] = mask[3, 3] = 9  # diagonal chain: 8-connected, not 4-connected
    marker = np.zeros_like(mask)
    marker[1, 1] = 9
    sq = oracle.reconstruct_bf("dilate", marker, mask, "SESquare2D")
    cr = oracle.reconstruct_bf("dilate", marker, mask, "SECross2D")
    np.testing.assert_array_equal(sq, mask)
    assert cr[1, 1] == 9 and cr[2, 2] == 0 and cr[3, 3] == 0


# ---------------------------------------------------------------- leveling pins
@pytest.mark.parametrize("dtype", [np.uint8, np.int32])
@pytest.mark.parametrize("se,shape", [("SESquare2D", (16, 21)), ("SECross2D", (13, 9)), ("SESegmentX1D", (50,))])
def test_leveling_priority_queue_equals_closed_forms(dtype, se, shape):
    """low(r,g) = r ^ R^d_{r v g}(g); high(r,g) = r v R^e_{r ^ g}(g); leveling = low(high(f,g),g)."""
    rng = np.random.default_rng(11)
    for _ in range(4):
        r = rng.integers(0, 30, size=shape).astype(dtype)
        g = rng.integers(0, 30, size=shape).astype(dtype)
        low = oracle.low_leveling(r, g, se)
        np.testing.assert_array_equal(
            low, np.minimum(r, oracle.reconstruct_bf("dilate", g, np.maximum(r, g), se)))
        high = oracle.high_leveling(r, g, se)
        np.testing.assert_array_equal(
            high, np.maximum(r, oracle.reconstruct_bf("erode", g, np.minimum(r, g), se)))
        lev = oracle.leveling(r, g, se)
        np.testing.assert_array_equal(lev, oracle.low_leveling(high, g, se))


# ---------------------------------------------------------------- arithmetic pins
def test_saturating_arith_closed_forms():
    a = np.arange(256, dtype=np.uint8)
    np.testing.assert_array_equal(oracle.ssub_const(a, 32), np.maximum(a.astype(int) - 32, 0).astype(np.uint8))
    np.testing.assert_array_equal(oracle.sadd_const(a, 32), np.minimum(a.astype(int) + 32, 255).astype(np.uint8))
    b = a[::-1].copy()
    np.testing.assert_array_equal(oracle.ssub_image(a, b),
                                  np.maximum(a.astype(int) - b.astype(int), 0).astype(np.uint8))
    i = np.array([-2**31, -5, 0, 2**31 - 1], np.int32)
    np.testing.assert_array_equal(oracle.ssub_const(i, 10), np.array([-2**31, -15, -10, 2**31 - 11], np.int32))
    # float saturated subtraction floors at FLT_MIN (a positive number), as the reference's
    # numeric_limits<float>::min() does
    f = np.array([1.0, 0.5, -3.0], np.float32)
    got = oracle.ssub_const(f, 1.0)
    assert got[0] == np.finfo(np.float32).tiny and got[1] == np.finfo(np.float32).tiny


def test_h_maxima_is_ssub_then_reconstruct():
    rng = np.random.default_rng(3)
    f = rng.integers(0, 256, size=(20, 20)).astype(np.uint8)
    exp = oracle.reconstruct_bf("dilate", np.maximum(f.astype(int) - 32, 0).astype(np.uint8), f, "SESquare2D")
    np.testing.assert_array_equal(oracle.h_maxima(f, 32, "SESquare2D"), exp)


def test_saturated_add_image_and_invert():
    """ArithSaturatedAddImage / ArithInvertImage: against Python integers (arbitrary precision,
    no accumulator type), the involution -(-a) = a, and the reference's own values
    (test/pixel_processing/arith.cpp:26-45: an int64 image of 1 negates to -1;
    arith.cpp:657-678: uint8 250 + 150 saturates at 255)."""
    rng = np.random.default_rng(12)
    for dt in (np.uint8, np.uint16, np.int32, np.int64):
        info = np.iinfo(dt)
        a = rng.integers(info.min, info.max, size=500, dtype=dt, endpoint=True)
        b = rng.integers(info.min, info.max, size=500, dtype=dt, endpoint=True)
        if info.min < 0:  # the reference clamps at the maximum only: keep sums above the minimum
            a, b = np.abs(a // 2), b // 2
        s = oracle.sadd_image(a, b)
        exp = [min(int(x) + int(y), int(info.max)) for x, y in zip(a, b)]
        assert [int(v) for v in s] == exp
        inv = oracle.invert(a)
        m = 1 << info.bits  # -a taken back into the type's range modulo 2^bits
        assert [int(v) for v in inv] == [(-int(x) - int(info.min)) % m + int(info.min) for x in a]
        np.testing.assert_array_equal(oracle.invert(inv), a)
    f = np.array([1.5, -0.0, 3e38, -2.0], np.float32)
    np.testing.assert_array_equal(oracle.invert(f), -f)
    assert np.signbit(oracle.invert(f)[1]) == False  # noqa: E712  (-(-0.0) = +0.0)
    big = np.array([3e38, 1.0], np.float32)
    assert oracle.sadd_image(big, big)[0] == np.finfo(np.float32).max
    np.testing.assert_array_equal(oracle.invert(np.ones((5, 6), np.int64)), -np.ones((5, 6), np.int64))
    np.testing.assert_array_equal(oracle.sadd_image(np.full((5, 6), 250, np.uint8), np.full((5, 6), 150, np.uint8)),
                                  np.full((5, 6), 255, np.uint8))


def _compare_kats():
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "compare_kats.json")
    with open(path) as f:
        return json.load(f)["cases"]


COMPARE_KATS = _compare_kats()


def compare_operands(c):
    return [kats.image(o["image"]) if "image" in o else o["value"] for o in (c["comp"], c["true"], c["false"])]


@pytest.mark.parametrize("c", COMPARE_KATS, ids=lambda c: f"compare.cpp:{c['line']}:{c['name']}")
def test_oracle_compare_kats(c):
    """every whole-image CompareImage KAT of the reference (8 operand forms x 6 operators)"""
    exp = kats.image(c["expected"])
    got = oracle.compare(kats.image(c["input"]), oracle.CMP_OPS[c["op"]], *compare_operands(c), exp.dtype)
    np.testing.assert_array_equal(got, exp)


def test_oracle_compare_ieee_and_types():
    """NaN compares false except for Diff (Python's float comparisons decide each pixel), and
    the output type is the true / false values' type, not the input's"""
    import operator
    x = np.array([np.nan, 1.0, 2.0, -np.inf, np.inf], np.float64)
    y = np.array([1.0, np.nan, 2.0, -np.inf, 0.0], np.float64)
    fns = {1: operator.eq, 2: operator.ne, 4: operator.gt, 8: operator.ge, 16: operator.lt, 32: operator.le}
    for op, fn in fns.items():
        got = oracle.compare(x, op, y, np.int32(7), np.int32(-1), np.int32)
        assert got.dtype == np.int32
        assert got.tolist() == [7 if fn(float(a), float(b)) else -1 for a, b in zip(x, y)]
