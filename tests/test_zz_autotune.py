"""Autotuner: the measured tile-width table is valid, persisted and reloaded, and results stay
bit-exact whatever widths are chosen.  Collected last: it changes the process-wide table."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
pm = pytest.importorskip("poutre2_amd")
import oracle  # noqa: E402

DEV = "cuda:0"


def _sq(img, r, dil):
    return oracle.static_iter(oracle.line_x(img, r, dil), "SESegmentY2D", min(r, img.shape[0]), dil)


def test_tuning_set_get_and_widths_are_bit_exact():
    img = np.random.default_rng(3).integers(0, 256, size=(300, 517), dtype=np.uint8)
    aimg = np.random.default_rng(4).integers(0, 256, size=(300, 512), dtype=np.uint8)
    try:
        for w in (128, 192, 256):
            for r in (1, 5, 9, 20):
                pm._call("pm_tuning_set", 1, r, w)
                pm._call("pm_tuning_set", 0, r, w)
                assert pm._capi.lib().pm_tuning_get(1, r) == w
                for a in (img, aimg):
                    got = pm.dilate(torch.from_numpy(a).to(DEV), pm.static_nl_se.square2d, r).cpu().numpy()
                    np.testing.assert_array_equal(got, _sq(a, r, True))
        with pytest.raises(pm.PoutreError):
            pm._call("pm_tuning_set", 1, 3, 100)
    finally:
        for r in range(1, 33):
            pm._call("pm_tuning_set", 1, r, 0)
            pm._call("pm_tuning_set", 0, r, 0)


def test_autotune_table_and_cache(tmp_path):
    cache = tmp_path / "tune.json"
    t1 = pm.autotune(cache=str(cache))
    assert set(t1) == {"aligned", "any"}
    assert all(len(v) == 32 and set(v) <= {128, 192, 256} for v in t1.values())
    assert cache.exists()
    for r in range(1, 33):  # reset, then reload from the cache
        pm._call("pm_tuning_set", 1, r, 0)
        pm._call("pm_tuning_set", 0, r, 0)
    t2 = pm.autotune(cache=str(cache))
    assert t2 == t1
    img = np.random.default_rng(5).integers(0, 256, size=(200, 333), dtype=np.uint8)
    for r in (2, 17):
        got = pm.erode(torch.from_numpy(img).to(DEV), pm.static_nl_se.square2d, r).cpu().numpy()
        np.testing.assert_array_equal(got, _sq(img, r, False))
