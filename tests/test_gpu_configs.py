"""Config-size parity: the five benchmark workloads at their full sizes (tests/configs.py),
run through the device path and compared bit for bit with the oracle's outputs on the same
inputs, by sha256 digest (tests/golden/config_digests.json, written in the build container by
tests/golden/make_config_digests.py from the CPU oracle).

Every test first checks the digest of its regenerated input, so a mismatch in input
generation is reported as such and never as a device error.  C5 also checks finite
differences on sampled tie-free pixels of the full tensor.
"""
import json
import os

import numpy as np
import pytest

from tests import configs

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
pm = pytest.importorskip("poutre2_amd")
from poutre2_amd import synth  # noqa: E402

DEV = "cuda:0"
D = configs.digest
_GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config_digests.json")


@pytest.fixture(scope="module")
def gold():
    with open(_GOLD) as f:
        return json.load(f)


def _h(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def _say(*a):
    print(*a, flush=True)


# ----------------------------------------------------------------------- C1
@pytest.mark.parametrize("op", ["dilate", "erode"])
def test_c1_square3x3_1024(gold, op):
    img = synth.uniform_u8((1024, 1024), 1)
    assert D(img) == gold["C1"]["input"], "input generation differs from the golden run"
    out = getattr(pm, op)(torch.from_numpy(img).to(DEV), pm.static_nl_se.square2d, 1)
    assert D(_h(out)) == gold["C1"][op]


# ----------------------------------------------------------------------- C2
@pytest.fixture(scope="module")
def img8k(gold):
    img = synth.uniform_u8((8192, 8192), 2)
    assert D(img) == gold["C2_2d"]["input"], "input generation differs from the golden run"
    return torch.from_numpy(img).to(DEV)


@pytest.mark.parametrize("side", configs.SIDES)
@pytest.mark.parametrize("op", configs.OPS_2D)
def test_c2_8192_square(gold, img8k, op, side):
    _say(f"C2 8192^2 {op} side {side}")
    out = getattr(pm, op)(img8k, pm.static_nl_se.square2d, (side - 1) // 2)
    assert D(_h(out)) == gold["C2_2d"][f"{op}_{side}"]


@pytest.fixture(scope="module")
def vol512(gold):
    vol = synth.uniform_u8((512, 512, 512), 3)
    assert D(vol) == gold["C2_3d"]["input"], "input generation differs from the golden run"
    return torch.from_numpy(vol).to(DEV)


@pytest.mark.parametrize("se", configs.SES_3D)
@pytest.mark.parametrize("op", ["dilate", "erode"])
@pytest.mark.parametrize("it", configs.ITERS_3D)
def test_c2_512cube(gold, vol512, se, op, it):
    _say(f"C2 512^3 {op} {se} iter {it}")
    pse = {"SECross3D": pm.static_nl_se.cross3d, "SESquare3D": pm.static_nl_se.square3d}[se]
    out = getattr(pm, op)(vol512, pse, it)
    assert D(_h(out)) == gold["C2_3d"][f"{op}_{se}_{it}"]


# ----------------------------------------------------------------------- C3
@pytest.fixture(scope="module")
def mask4k(gold):
    m = synth.blurred_noise_mask(4096, 4096, 4, device=DEV)
    assert D(_h(m)) == gold["C3"]["mask"], "device mask generation differs from the golden (numpy) run"
    return m


def test_c3_reconstruction_hmax32(gold, mask4k):
    marker = synth.hmax_marker(mask4k, 32)
    assert D(_h(marker)) == gold["C3"]["marker"]
    out = pm.reconstruction(pm.RecType.dilate, marker, mask4k, pm.static_nl_se.square2d)
    assert D(_h(out)) == gold["C3"]["reconstruct"]
    # h_maxima(mask, 32) = the same reconstruction (extrema.cpp:39-49)
    out2 = pm.h_maxima(mask4k, 32, pm.static_nl_se.square2d)
    assert D(_h(out2)) == gold["C3"]["reconstruct"]


def test_c3_leveling_open15(gold, mask4k):
    g = pm.opening(mask4k, pm.static_nl_se.square2d, 7)
    assert D(_h(g)) == gold["C3"]["opening_7"]
    out = pm.leveling(mask4k, g, pm.static_nl_se.square2d)
    assert D(_h(out)) == gold["C3"]["leveling"]


# ----------------------------------------------------------------------- C4
@pytest.fixture(scope="module")
def batch512(gold):
    shapes, radii, imgs = synth.batch_dilate_open_workload()
    g = gold["C4"]["dilate_open"]
    assert configs.digest_of_digests([D(a) for a in imgs]) == g["inputs"], "input generation differs"
    return shapes, radii, [torch.from_numpy(a).to(DEV) for a in imgs]


def _check_each(outs, each, total, what):
    ds = [D(_h(o)) for o in outs]
    bad = [i for i, (d, e) in enumerate(zip(ds, each)) if d[:16] != e]
    assert not bad, f"{what}: {len(bad)} images differ, first {bad[:8]}"
    assert configs.digest_of_digests(ds) == total


def test_c4_batch_dilate_open_512(gold, batch512):
    shapes, radii, imgs = batch512
    g = gold["C4"]["dilate_open"]
    dils, opens = pm.batch_dilate_open(imgs, radii)
    _check_each(dils, g["dilate_each"], g["dilate"], "dilation")
    _check_each(opens, g["opening_each"], g["opening"], "opening")


def test_c4_batch_plan_packed_512(gold, batch512):
    shapes, radii, imgs = batch512
    g = gold["C4"]["dilate_open"]
    flat = torch.cat([a.reshape(-1) for a in imgs])
    plan = pm.BatchPlan.packed(flat, shapes, radii)
    try:
        for rep in range(2):  # replays with eager work in between
            plan.dil_outs.zero_()
            plan.open_outs.zero_()
            d, o = plan.run()
            sizes = [a.numel() for a in imgs]
            dl = list(torch.split(d, sizes))
            ol = list(torch.split(o, sizes))
            _check_each([x.view(s) for x, s in zip(dl, shapes)], g["dilate_each"], g["dilate"], f"dilation rep {rep}")
            _check_each([x.view(s) for x, s in zip(ol, shapes)], g["opening_each"], g["opening"], f"opening rep {rep}")
    finally:
        plan.close()


def test_c4_batch_reconstruction_256(gold, batch512):
    g = gold["C4"]["reconstruct"]
    markers, masks = synth.batch_reconstruction_workload(device=DEV)
    assert configs.digest_of_digests([D(_h(m)) for m in masks]) == g["masks"], "device mask generation differs"
    assert configs.digest_of_digests([D(_h(m)) for m in markers]) == g["markers"]
    outs = pm.batch_reconstruction(pm.RecType.dilate, markers, masks, pm.static_nl_se.square2d)
    _check_each(outs, g["out_each"], g["out"], "reconstruction")
    # the packed form over the same images
    shapes = [tuple(m.shape) for m in masks]
    of = pm.batch_reconstruction_packed(pm.RecType.dilate, torch.cat([m.reshape(-1) for m in markers]),
                                        torch.cat([m.reshape(-1) for m in masks]), shapes, pm.static_nl_se.square2d)
    parts = torch.split(of, [m.numel() for m in masks])
    _check_each([p.view(s) for p, s in zip(parts, shapes)], g["out_each"], g["out"], "packed reconstruction")


# ----------------------------------------------------------------------- C5
@pytest.fixture(scope="module")
def c5_data(gold):
    x = configs.c5_input()
    w = configs.c5_w()
    gy = configs.c5_gy()
    g = gold["C5"]
    assert D(x) == g["input"] and D(w) == g["w"] and D(gy) == g["gy"], "input generation differs"
    return x, w, torch.from_numpy(x).to(DEV), torch.from_numpy(w).to(DEV), torch.from_numpy(gy).to(DEV).float()


def _tie_free(x, w, k, dil, p, eps):
    """every output whose window holds input p has a top-2 gap above 2 eps (so +-eps at p
    moves no selection); x numpy [B, C, H, W], p = (b, c, i, j)"""
    b, c, i, j = p
    H, W = x.shape[-2:]
    r = (k - 1) // 2
    sgn = 1.0 if dil else -1.0
    for qi in range(max(0, i - r), min(H, i + r + 1)):
        for qj in range(max(0, j - r), min(W, j + r + 1)):
            vals = []
            for dy in range(-r, r + 1):
                for dx in range(-r, r + 1):
                    yy, xx = qi + dy, qj + dx
                    if 0 <= yy < H and 0 <= xx < W:
                        wv = float(w[dy + r, dx + r]) if w is not None else 0.0
                        vals.append(sgn * (float(x[b, c, yy, xx]) + sgn * wv))
            vals.sort(reverse=True)
            if len(vals) > 1 and vals[0] - vals[1] <= 2 * eps:
                return False
    return True


@pytest.mark.parametrize("name,k,with_w,dil", configs.C5_OPS, ids=[o[0] for o in configs.C5_OPS])
def test_c5_autograd_64x32x256x256(gold, c5_data, name, k, with_w, dil):
    from poutre2_amd import autograd as ag
    _say(f"C5 {name}")
    x_np, w_np, x, w, gy = c5_data
    g = gold["C5"][name]
    f = ag.dilation2d if dil else ag.erosion2d
    xr = x.detach().clone().requires_grad_(True)
    wr = w.detach().clone().requires_grad_(True) if with_w else None
    y = f(xr, k=k, w=wr)
    assert D(_h(y)) == g["y"], "forward differs from the oracle"
    y.backward(gy)
    assert D(_h(xr.grad)) == g["gx"], "input gradient differs from the oracle"
    if with_w:
        np.testing.assert_array_equal(_h(wr.grad).astype(np.float64), np.asarray(g["gw"], dtype=np.float64))
    # the forward of an unrecorded call gives the same bits
    with torch.no_grad():
        assert D(_h(f(x, k=k, w=w if with_w else None))) == g["y"]

    # central finite differences on sampled tie-free pixels of the full tensor
    eps = 1e-3
    gx = xr.grad
    rng = np.random.default_rng(1234 + k)
    B, C, H, W = x_np.shape
    r = (k - 1) // 2
    xs = x.detach().clone()
    checked = 0
    for _ in range(40):
        p = (int(rng.integers(B)), int(rng.integers(C)), int(rng.integers(H)), int(rng.integers(W)))
        if not _tie_free(x_np, w_np if with_w else None, k, dil, p, eps):
            continue
        b, c, i, j = p
        y0, y1 = max(0, i - r), min(H, i + r + 1)
        x0, x1 = max(0, j - r), min(W, j + r + 1)
        orig = xs[p].item()
        loss = []
        for s in (+1, -1):
            xs[p] = float(np.float32(orig + s * eps))
            with torch.no_grad():
                yy = f(xs, k=k, w=w if with_w else None)
            loss.append(float((yy[b, c, y0:y1, x0:x1].double() * gy[b, c, y0:y1, x0:x1].double()).sum()))
        xs[p] = orig
        fd = (loss[0] - loss[1]) / (float(np.float32(orig + eps)) - float(np.float32(orig - eps)))
        an = gx[p].item()
        assert abs(fd - an) <= 1e-3 * abs(an) + 2e-2, f"FD {fd} vs analytic {an} at {p}"
        checked += 1
    assert checked >= 10
