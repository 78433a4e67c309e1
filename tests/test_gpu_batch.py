"""Batched launches: many uint8 images of uneven sizes and mixed radii in one launch,
bit-exact against the oracle per image (and against the per-image GPU path)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
pm = pytest.importorskip("poutre2_amd")
from poutre2_amd import synth  # noqa: E402

DEV = "cuda:0"


def _batch(n, seed, hmax=300, rmax=15):
    rng = np.random.default_rng(seed)
    imgs, radii = [], []
    for i in range(n):
        H, W = int(rng.integers(1, hmax)), int(rng.integers(1, hmax))
        imgs.append(rng.integers(0, 256, size=(H, W)).astype(np.uint8))
        radii.append(int(rng.integers(1, rmax + 1)))
    return imgs, radii


def _square(img, r, dil):
    return oracle.static_iter(oracle.line_x(img, r, dil), "SESegmentY2D", min(r, img.shape[0]), dil)


@pytest.mark.parametrize("seed,n,hmax,rmax", [(1, 12, 300, 15), (2, 5, 40, 32), (3, 1, 129, 7), (4, 30, 70, 3),
                                            (5, 8, 700, 7), (6, 6, 700, 15)])
def test_batch_dilate_open_matches_oracle(seed, n, hmax, rmax):
    imgs, radii = _batch(n, seed, hmax, rmax)
    dils, opens = pm.batch_dilate_open([torch.from_numpy(a).to(DEV) for a in imgs], radii)
    torch.cuda.synchronize()
    for a, r, d, o in zip(imgs, radii, dils, opens):
        np.testing.assert_array_equal(d.cpu().numpy(), _square(a, r, True))
        np.testing.assert_array_equal(o.cpu().numpy(), _square(_square(a, r, False), r, True))


def test_batch_dilate_open_equals_per_image_path():
    imgs, radii = _batch(8, 9, 520, 15)
    ts = [torch.from_numpy(a).to(DEV) for a in imgs]
    dils, opens = pm.batch_dilate_open(ts, radii)
    for t, r, d, o in zip(ts, radii, dils, opens):
        assert torch.equal(d, pm.dilate(t, pm.static_nl_se.square2d, r))
        assert torch.equal(o, pm.opening(t, pm.static_nl_se.square2d, r))


def test_batch_rejects_bad_radius():
    a = torch.zeros((10, 10), dtype=torch.uint8, device=DEV)
    with pytest.raises(pm.PoutreError):
        pm.batch_dilate_open([a], [0])


@pytest.mark.parametrize("se", ["SESquare2D", "SECross2D"])
@pytest.mark.parametrize("rec", ["dilate", "erode"])
def test_batch_reconstruction_matches_oracle(se, rec):
    rng = np.random.default_rng(5)
    markers, masks = [], []
    for i in range(9):
        H, W = int(rng.integers(5, 260)), int(rng.integers(5, 260))
        mask = synth.blurred_noise_mask(H, W, 100 + i, sigma=3.0, radius=12)
        marker = synth.hmax_marker(mask, 16) if rec == "dilate" else np.minimum(mask.astype(int) + 16, 255).astype(np.uint8)
        markers.append(marker)
        masks.append(mask)
    outs = pm.batch_reconstruction(pm.RecType[rec], [torch.from_numpy(m).to(DEV) for m in markers],
                                   [torch.from_numpy(m).to(DEV) for m in masks], oracle.SE[se])
    torch.cuda.synchronize()
    for mk, ms, o in zip(markers, masks, outs):
        np.testing.assert_array_equal(o.cpu().numpy(), oracle.reconstruct(rec, mk, ms, se))


@pytest.mark.parametrize("seed,n,hmax,rmax", [(11, 10, 300, 15), (12, 6, 40, 32)])
def test_batch_packed_matches_list_api(seed, n, hmax, rmax):
    """Images packed back to back in one buffer give the same bits as the list API."""
    imgs, radii = _batch(n, seed, hmax, rmax)
    flat = torch.from_numpy(np.concatenate([a.reshape(-1) for a in imgs])).to(DEV)
    shapes = [a.shape for a in imgs]
    d_flat, o_flat = pm.batch_dilate_open_packed(flat, shapes, radii)
    dils, opens = pm.batch_dilate_open([torch.from_numpy(a).to(DEV) for a in imgs], radii)
    torch.cuda.synchronize()
    off = 0
    for a, d, o in zip(imgs, dils, opens):
        k = a.size
        assert torch.equal(d_flat[off:off + k].view(a.shape), d)
        assert torch.equal(o_flat[off:off + k].view(a.shape), o)
        off += k


def test_batch_reconstruction_packed_matches_oracle():
    rng = np.random.default_rng(9)
    markers, masks = [], []
    for i in range(7):
        H, W = int(rng.integers(5, 200)), int(rng.integers(5, 200))
        mask = synth.blurred_noise_mask(H, W, 200 + i, sigma=3.0, radius=12)
        markers.append(synth.hmax_marker(mask, 16))
        masks.append(mask)
    shapes = [m.shape for m in masks]
    mk = torch.from_numpy(np.concatenate([m.reshape(-1) for m in markers])).to(DEV)
    ms = torch.from_numpy(np.concatenate([m.reshape(-1) for m in masks])).to(DEV)
    out = pm.batch_reconstruction_packed(pm.RecType.dilate, mk, ms, shapes, pm.static_nl_se.square2d)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    off = 0
    for a, b in zip(markers, masks):
        k = b.size
        np.testing.assert_array_equal(got[off:off + k].reshape(b.shape), oracle.reconstruct("dilate", a, b, "SESquare2D"))
        off += k


def test_batch_packed_rejects_short_buffer():
    flat = torch.zeros(50, dtype=torch.uint8, device=DEV)
    with pytest.raises(pm.PoutreError):
        pm.batch_dilate_open_packed(flat, [(10, 10)], [1])


def test_batch_plan_replays_match_eager():
    """A graph-captured plan gives the eager call's bits on every replay, with eager work and
    synchronisations between replays and the inputs changed in place."""
    imgs, radii = _batch(9, 21, 300, 15)
    ts = [torch.from_numpy(a).to(DEV) for a in imgs]
    plan = pm.BatchPlan(ts, radii)
    try:
        for rep in range(4):
            for t in ts:
                t.random_(0, 256)
            dils, opens = plan.run()
            _ = torch.stack([t.float().mean() for t in ts]).sum().item()  # eager work between replays
            torch.cuda.synchronize()
            ed, eo = pm.batch_dilate_open(ts, radii)
            torch.cuda.synchronize()
            for a, b in zip(dils, ed):
                assert torch.equal(a, b), rep
            for a, b in zip(opens, eo):
                assert torch.equal(a, b), rep
    finally:
        plan.close()


def test_batch_plan_packed_matches_oracle():
    imgs, radii = _batch(6, 22, 200, 32)
    flat = torch.from_numpy(np.concatenate([a.reshape(-1) for a in imgs])).to(DEV)
    plan = pm.BatchPlan.packed(flat, [a.shape for a in imgs], radii)
    d, o = plan.run()
    d, o = plan.run()
    torch.cuda.synchronize()
    d, o = d.cpu().numpy(), o.cpu().numpy()
    off = 0
    for a, r in zip(imgs, radii):
        k = a.size
        np.testing.assert_array_equal(d[off:off + k].reshape(a.shape), _square(a, r, True))
        np.testing.assert_array_equal(o[off:off + k].reshape(a.shape), _square(_square(a, r, False), r, True))
        off += k
    plan.close()
    with pytest.raises(pm.PoutreError):
        plan.run()
