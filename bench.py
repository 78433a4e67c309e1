"""Benchmark of the MI355X morphology path: prints ONE JSON line.

Headline workload (one GPU): flat dilation of an 8192x8192 uint8 image (iid uniform,
seed 2) by square structuring elements of side 3, 7, 15, 31 and 63.  One step = the
five dilations, each on its own rotating input/output pair (3 pairs, 384 MiB in all, so
the 256 MiB MALL cannot hold the working set).  `value` = pixels dilated per second
over the K timed steps (Gpixel/s), inputs resident in HBM before the timed region.

Also reported (not `value`): per-side figures, erosion / opening / closing / gradient,
4096^2 h-maxima reconstruction and leveling, the roofline of the dominant kernel
(HIP events on the launch stream) and a single-thread CPU baseline (the oracle's
restatement of the reference's DilateX + DilateY, timed on a bounded sample).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--sections a,b]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
SIDES = [3, 7, 15, 31, 63]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


_SLEEP_CYCLES_PER_US = []


def _queue_ahead(stream, us):
    """Occupy `stream` for about `us` microseconds with torch's spin kernel, so that launches
    the host issues meanwhile queue up behind it and then run back to back.  The spin rate is
    calibrated once with HIP events.  False when torch has no spin kernel."""
    import torch
    sleep = getattr(torch.cuda, "_sleep", None)
    if sleep is None:
        return False
    with torch.cuda.stream(stream):
        if not _SLEEP_CYCLES_PER_US:
            sleep(1000)  # first launch loads the kernel
            t0 = torch.cuda.Event(enable_timing=True)
            t1 = torch.cuda.Event(enable_timing=True)
            t0.record(stream)
            sleep(2_000_000)
            t1.record(stream)
            t1.synchronize()
            _SLEEP_CYCLES_PER_US.append(2_000_000 / max(t0.elapsed_time(t1) * 1e3, 1.0))
        sleep(int(us * _SLEEP_CYCLES_PER_US[0]))
    return True


def timed(fn, steps, stream, ahead_us=0):
    """Run fn() `steps` times between two HIP events recorded on `stream`; returns ms.
    ahead_us > 0: the stream is first held busy for that long (_queue_ahead), so that the
    timed launches are all queued when the first event fires and the figure is the GPU's
    back-to-back time, not the host's launch rate (for kernels of a few tens of us)."""
    import torch
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    if ahead_us:
        _queue_ahead(stream, ahead_us)
    t0.record(stream)
    for _ in range(steps):
        fn()
    t1.record(stream)
    t1.synchronize()
    return t0.elapsed_time(t1)


def bench_dilate_8k(pm, torch, steps, warmup, dev):
    from poutre2_amd import synth
    H = W = 8192
    npx = H * W
    base = torch.from_numpy(synth.uniform_u8((H, W), 2)).to(dev)
    pairs = [(base.roll(i * 977, dims=1).contiguous(), torch.empty_like(base)) for i in range(3)]
    stream = torch.cuda.current_stream(dev)
    state = {"i": 0}

    def step():
        for s in SIDES:
            a, b = pairs[state["i"] % 3]
            state["i"] += 1
            pm.dilate(a, pm.static_nl_se.square2d, (s - 1) // 2, out=b)

    for _ in range(warmup):
        step()
    ms = timed(step, steps, stream)
    ms_step = ms / steps
    gpx = npx * len(SIDES) * steps / (ms * 1e-3) / 1e9

    per_side = {}
    for s in SIDES:
        def one(s=s):
            a, b = pairs[state["i"] % 3]
            state["i"] += 1
            pm.dilate(a, pm.static_nl_se.square2d, (s - 1) // 2, out=b)
        one()
        t = timed(one, max(3, steps), stream) / max(3, steps)
        per_side[str(s)] = {"ms": round(t, 4), "gpx_s": round(npx / (t * 1e-3) / 1e9, 2),
                            "pct_hbm": round(100 * 2 * npx / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 2)}

    # dominant kernel: the single fused launch of the slowest side (sq_u8_kernel<K, true>)
    dom = max(per_side, key=lambda k: per_side[k]["ms"])
    t_dom = per_side[dom]["ms"]
    r_dom = (int(dom) - 1) // 2
    kk = 32 if r_dom >= 16 else 16 if r_dom >= 8 else 8 if r_dom >= 4 else 4
    achieved = 2 * npx / (t_dom * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": f"sq_u8_kernel<{kk}, true, true, {r_dom}, {pm.tuning()['aligned'][r_dom - 1]}> "
                          f"(side {dom}) on 8192x8192 uint8",
                "kernel_ms": t_dom, "bytes_per_launch": 2 * npx, "side": int(dom)}
    # practical ceiling: a plain elementwise kernel moving the same bytes (64 MiB read + 64 MiB
    # written) over the same rotating pairs, queued back to back
    def copy_one():
        a, b = pairs[state["i"] % 3]
        state["i"] += 1
        torch.bitwise_not(a, out=b)
    copy_one()
    reps = max(10, steps)
    t_copy = timed(copy_one, reps, stream, ahead_us=60 * reps + 300) / reps
    roofline["copy_ceiling"] = {
        "kernel": "torch.bitwise_not on the same 8192x8192 uint8 pairs (64 MiB read + 64 MiB written)",
        "ms": round(t_copy, 4), "achieved_gbs": round(2 * npx / (t_copy * 1e-3) / 1e9, 1),
        "frac_of_peak": round(2 * npx / (t_copy * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "dominant_kernel_frac_of_copy": round(t_copy / t_dom, 4)}
    # op-level fraction of HBM peak over the whole step (2 B/px algorithmic)
    pct = 100 * 2 * npx * len(SIDES) * steps / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS
    del pairs, base
    return gpx, ms_step, per_side, roofline, round(pct, 2)


def bench_ops_8k(pm, torch, dev, reps=9):
    """8192^2 uint8 erosion / opening / closing / gradient by squares of side 3, 15, 63 on three
    rotating input/output pairs (384 MiB, more than the MALL holds), two warm-up calls each."""
    from poutre2_amd import synth
    H = W = 8192
    npx = H * W
    ins = [torch.from_numpy(synth.uniform_u8((H, W), 2 + i)).to(dev) for i in range(3)]
    outs = [torch.empty_like(a) for a in ins]
    stream = torch.cuda.current_stream(dev)
    out = {}
    for name, fn in (("erode", pm.erode), ("opening", pm.opening), ("closing", pm.closing),
                     ("gradient", pm.gradient)):
        per = {}
        for s in (3, 15, 63):
            k = [0]

            def step():
                i = k[0] % 3
                k[0] += 1
                fn(ins[i], pm.static_nl_se.square2d, (s - 1) // 2, out=outs[i])
            step()
            step()
            t = timed(step, reps, stream) / reps
            per[str(s)] = {"ms": round(t, 4), "gpx_s": round(npx / (t * 1e-3) / 1e9, 2),
                           "pct_hbm": round(100 * 2 * npx / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 2)}
        out[name] = per
    return out


def bench_recon_4k(pm, torch, dev, reps=6):
    """4096^2 h-maxima reconstruction (h = 32, blurred-noise mask seed 4) and leveling with the
    15x15-opening marker.  Timed on 6 rotating sets (the mask rolled by a different offset
    each: the same statistics, 288 MiB in all, more than the 256 MiB MALL holds); the
    figure on one set (48 MiB, MALL-hot) is reported beside it."""
    from poutre2_amd import synth
    H = W = 4096
    npx = H * W
    mask0 = synth.blurred_noise_mask(H, W, 4, device=dev)
    sets = []
    for i in range(6):
        m = mask0 if i == 0 else mask0.roll((977 * i, 1931 * i), dims=(0, 1)).contiguous()
        sets.append((synth.hmax_marker(m, 32), m, torch.empty_like(m)))
    stream = torch.cuda.current_stream(dev)
    k = [0]

    def rec(rot=True):
        mk, ms, o = sets[k[0] % 6 if rot else 0]
        k[0] += 1
        pm.reconstruction(pm.RecType.dilate, mk, ms, pm.static_nl_se.square2d, out=o)
    rec()
    t = timed(rec, reps, stream) / reps
    th = timed(lambda: rec(False), reps, stream) / reps

    def fig(t):
        return {"ms": round(t, 3), "gpx_s": round(npx / (t * 1e-3) / 1e9, 3),
                "pct_hbm": round(100 * 3 * npx / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 2)}
    res = {"reconstruct_hmax32": dict(fig(t), rotating_buffers_MiB=288, mall_hot_ms=round(th, 3))}
    # leveling with marker = 15x15 opening of the mask
    lsets = [(ms, pm.opening(ms, pm.static_nl_se.square2d, 7), o) for _, ms, o in sets]

    def lev(rot=True):
        f, g, o = lsets[k[0] % 6 if rot else 0]
        k[0] += 1
        pm.leveling(f, g, pm.static_nl_se.square2d, out=o)
    lev()
    t = timed(lev, reps, stream) / reps
    th = timed(lambda: lev(False), reps, stream) / reps
    res["leveling_open15"] = dict(fig(t), rotating_buffers_MiB=288, mall_hot_ms=round(th, 3))
    return res


def bench_3d(pm, torch, dev, reps=3):
    """512^3 uint8 volume (iid uniform, seed 3): dilation and erosion by SECross3D and SESquare3D
    iterated 1, 2, 3 times (2 B/voxel algorithmic), on two alternating volume pairs."""
    from poutre2_amd import synth
    n = 512 ** 3
    base = torch.from_numpy(synth.uniform_u8((512, 512, 512), 3)).to(dev)
    # two input/output pairs (512 MiB), alternated so that the 256 MiB MALL cannot hold them
    pairs = [(base, torch.empty_like(base)), (base.roll(97, dims=2).contiguous(), torch.empty_like(base))]
    stream = torch.cuda.current_stream(dev)
    out = {}
    state = {"i": 0}
    for se_name, se in (("cross3d", pm.static_nl_se.cross3d), ("square3d", pm.static_nl_se.square3d)):
        for it in (1, 2, 3):
            for op_name, fn in (("dilate", pm.dilate), ("erode", pm.erode)):
                def one(fn=fn, se=se, it=it):
                    a, b = pairs[state["i"] % 2]
                    state["i"] += 1
                    fn(a, se, it, out=b)
                one()
                t = timed(one, 2 * reps, stream) / (2 * reps)
                out[f"{op_name}_{se_name}_iter{it}"] = {"ms": round(t, 4), "gvox_s": round(n / (t * 1e-3) / 1e9, 2),
                                                        "pct_hbm": round(100 * 2 * n / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 2)}
    return out


def bench_types(pm, torch, dev, reps=20):
    """4096x4096 uint16 / int32 / float32 dilation by squares of side 3 and 15; %HBM at
    2 * sizeof(T) B/px algorithmic.  Input/output pairs rotate over more than the 256 MiB MALL
    (3 pairs = 384 MiB for 4-byte types, 5 pairs = 320 MiB for uint16).  The launches (17-50 us
    each) are queued behind a spin kernel before the first event, so the figure is GPU time."""
    import numpy as np
    from poutre2_amd import synth
    g = synth.rng(9)
    out = {}
    for name, npt, npairs in (("uint16", np.uint16, 5), ("int32", np.int32, 3), ("float32", np.float32, 3)):
        if npt == np.float32:
            a_np = g.standard_normal((4096, 4096), dtype=np.float32)
        else:
            a_np = g.integers(0, 60000, size=(4096, 4096)).astype(npt)
        a0 = torch.from_numpy(a_np).to(dev)
        # rolled on the host: torch has no device roll for uint16
        pairs = [(a0 if i == 0 else torch.from_numpy(np.roll(a_np, 613 * i, axis=1)).to(dev), torch.empty_like(a0))
                 for i in range(npairs)]
        stream = torch.cuda.current_stream(dev)
        per = {}
        for side in (3, 15):
            r = (side - 1) // 2
            k = [0]

            def one(r=r):
                a, b = pairs[k[0] % npairs]
                k[0] += 1
                pm.dilate(a, pm.static_nl_se.square2d, r, out=b)
            one()
            t = timed(one, reps, stream, ahead_us=30 * reps + 300) / reps
            per[str(side)] = {"ms": round(t, 4), "gpx_s": round(a0.numel() / (t * 1e-3) / 1e9, 2),
                              "pct_hbm": round(100 * 2 * a0.element_size() * a0.numel() / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 2)}
        per["rotating_buffers_MiB"] = npairs * 2 * a0.numel() * a0.element_size() >> 20
        out[name] = per
        del pairs, a0
    return out


def bench_small_1k(pm, torch, dev, reps=50):
    """The reference's CPU-runnable case: 1024^2 uint8, SESquare2D iter 1, dilation and erosion."""
    from poutre2_amd import synth
    a = torch.from_numpy(synth.uniform_u8((1024, 1024), 1)).to(dev)
    b = torch.empty_like(a)
    stream = torch.cuda.current_stream(dev)
    out = {}
    for name, fn in (("dilate", pm.dilate), ("erode", pm.erode)):
        fn(a, pm.static_nl_se.square2d, 1, out=b)
        t = timed(lambda: fn(a, pm.static_nl_se.square2d, 1, out=b), reps, stream) / reps
        out[name] = {"us": round(t * 1e3, 2), "gpx_s": round(a.numel() / (t * 1e-3) / 1e9, 2)}
    out["note"] = "cache-hot: a 2 MiB working set stays in the MALL; a call is launch-bound at this size"
    return out


def bench_label_4k(pm, torch, dev, reps=5, cpu=True):
    """Connected-component labelling of a 4096^2 uint8 image (blurred-noise mask, seed 4,
    thresholded at 128: a few thousand large winding blobs): label_binary by SESquare2D and
    SECross2D, label_flat_zones of the mask quantised to 8 levels.  Bytes: 1 B read + 8 B
    (int64 label) written per pixel.  cpu: the oracle's one-pass union-find on one core."""
    import time
    import numpy as np
    from poutre2_amd import synth
    H = W = 4096
    npx = H * W
    mask = synth.blurred_noise_mask(H, W, 4, device=dev).cpu().numpy()
    binary = (mask > 127).astype(np.uint8)
    zones = (mask // 32).astype(np.uint8)
    b = torch.from_numpy(binary).to(dev)
    z = torch.from_numpy(zones).to(dev)
    out = torch.empty((H, W), dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream(dev)
    res = {}
    for name, img, fn, se in (("binary_square", b, pm.label_binary, pm.static_nl_se.square2d),
                              ("binary_cross", b, pm.label_binary, pm.static_nl_se.cross2d),
                              ("flat_zones_square", z, pm.label_flat_zones, pm.static_nl_se.square2d)):
        _, n = fn(img, se, out=out)
        t = timed(lambda: fn(img, se, out=out), reps, stream) / reps
        res[name] = {"ms": round(t, 3), "labels": n, "gpx_s": round(npx / (t * 1e-3) / 1e9, 2),
                     "pct_hbm": round(100 * 9 * npx / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 2)}
    if cpu:
        import oracle
        t0 = time.perf_counter()
        oracle.label(binary, "SESquare2D")
        res["binary_square"]["cpu_oracle_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
    return res


def bench_batched(pm, torch, dev, reps=3):
    """512 uint8 images, H, W ~ U{64..2048}, square side ~ U{3,5,..,31}: dilation + opening per
    image in one launch vs the per-image path and vs the single-image 8192^2 kernels' per-pixel
    rate; 256 h-maxima reconstructions (sizes U{128..1024}, blurred-noise masks, h = 16) in one
    round sequence vs one call per image."""
    import numpy as np
    from poutre2_amd import synth
    shapes, radii, imgs_np = synth.batch_dilate_open_workload()
    imgs = [torch.from_numpy(a).to(dev) for a in imgs_np]
    del imgs_np
    dil = [torch.empty_like(a) for a in imgs]
    opn = [torch.empty_like(a) for a in imgs]
    npx = sum(a.numel() for a in imgs)
    stream = torch.cuda.current_stream(dev)
    fb = lambda: pm.batch_dilate_open(imgs, radii, dil, opn)  # noqa: E731
    fb()
    tb = timed(fb, reps, stream) / reps

    def per_image():
        for a, r, d, o in zip(imgs, radii, dil, opn):
            pm.dilate(a, pm.static_nl_se.square2d, r, out=d)
            pm.opening(a, pm.static_nl_se.square2d, r, out=o)
    per_image()
    tp = timed(per_image, 1, stream)
    # the same images packed back to back in one buffer (descriptors computed by numpy)
    flat = torch.cat([a.reshape(-1) for a in imgs])
    shapes = [tuple(a.shape) for a in imgs]
    dflat, oflat = torch.empty_like(flat), torch.empty_like(flat)
    fp = lambda: pm.batch_dilate_open_packed(flat, shapes, radii, dflat, oflat)  # noqa: E731
    fp()
    tpk = timed(fp, reps, stream) / reps
    # the same packed batch captured once into a graph (pm_batch_plan): one launch per run
    plan = pm.BatchPlan.packed(flat, shapes, radii, dflat, oflat)
    plan.run()
    tpl = timed(plan.run, reps, stream) / reps
    plan.close()
    tbest = min(tpk, tpl)
    del imgs, dil, opn, flat, dflat, oflat
    # per-pixel efficiency against the single-image kernels: the time the 8192^2 kernels take
    # for the same pixels at each image's radius (dilation + opening per radius, timed on three
    # rotating 8192^2 input/output pairs, past the MALL)
    base = torch.from_numpy(synth.uniform_u8((8192, 8192), 2)).to(dev)
    pairs8k = [(base.roll(i * 977, dims=1).contiguous(), torch.empty_like(base)) for i in range(3)]
    del base
    per_px = {}
    k8 = [0]
    for r in sorted(set(radii)):
        def both(r=r):
            a, b = pairs8k[k8[0] % 3]
            k8[0] += 1
            pm.dilate(a, pm.static_nl_se.square2d, r, out=b)
            pm.opening(a, pm.static_nl_se.square2d, r, out=b)
        both()
        per_px[r] = timed(both, 3, stream) / 3 / (8192 * 8192)
    del pairs8k
    equiv = sum(h * w * per_px[r] for (h, w), r in zip(shapes, radii))
    res = {"dilate_open_512": {"mpx": round(npx / 1e6, 1), "ms": round(tpl, 3),
                               "gpx_s": round(npx / (tpl * 1e-3) / 1e9, 2),
                               "pct_hbm": round(100 * 3 * npx / (tpl * 1e-3) / 1e9 / HBM_PEAK_GBS, 2),
                               "api": "graph-captured plan over a packed buffer (BatchPlan.packed): one kernel "
                                      "launch for all 512 images and 15 radii",
                               "packed_call_ms": round(tpk, 3), "list_api_ms": round(tb, 3),
                               "single_image_kernels_equiv_ms": round(equiv, 3),
                               "efficiency_vs_single_image_kernels": round(equiv / tpl, 3),
                               "per_image_ms": round(tp, 3), "speedup_vs_per_image": round(tp / tbest, 2)}}
    markers, masks = synth.batch_reconstruction_workload(device=dev)
    outs = [torch.empty_like(m) for m in masks]
    rpx = sum(m.numel() for m in masks)
    fr = lambda: pm.batch_reconstruction(pm.RecType.dilate, markers, masks, pm.static_nl_se.square2d, outs)  # noqa: E731
    fr()
    tr = timed(fr, reps, stream) / reps

    def per_rec():
        for mk, ms, o in zip(markers, masks, outs):
            pm.reconstruction(pm.RecType.dilate, mk, ms, pm.static_nl_se.square2d, out=o)
    per_rec()
    tpr = timed(per_rec, 1, stream)
    mkf, msf = torch.cat([m.reshape(-1) for m in markers]), torch.cat([m.reshape(-1) for m in masks])
    rshapes = [tuple(m.shape) for m in masks]
    of = torch.empty_like(msf)
    frp = lambda: pm.batch_reconstruction_packed(pm.RecType.dilate, mkf, msf, rshapes, pm.static_nl_se.square2d, of)  # noqa: E731
    frp()
    trp = timed(frp, reps, stream) / reps
    res["reconstruct_256"] = {"mpx": round(rpx / 1e6, 1), "ms": round(trp, 3),
                              "gpx_s": round(rpx / (trp * 1e-3) / 1e9, 2),
                              "pct_hbm": round(100 * 3 * rpx / (trp * 1e-3) / 1e9 / HBM_PEAK_GBS, 2),
                              "list_api_ms": round(tr, 3), "api": "packed buffer (batch_reconstruction_packed)",
                              "per_image_ms": round(tpr, 3), "speedup_vs_per_image": round(tpr / trp, 2)}
    return res


def bench_autograd(pm, torch, dev, reps=3):
    """float32 [64, 32, 256, 256] (N(0,1) + 1e-4 jitter, seed 7): flat 3x3 / 7x7 dilation and
    erosion and 5x5 max-plus dilation, forward and forward+backward.  8 B/px forward (x, y);
    when recorded for a backward, 9 B/px forward (x, y, selected tap) + 9 B/px backward (tap,
    gy, gx)."""
    from poutre2_amd import autograd as ag
    from poutre2_amd import synth
    x = torch.from_numpy(synth.normal_f32((64, 32, 256, 256), 7)).to(dev)
    gy = torch.ones_like(x)
    w = torch.from_numpy((0.1 * synth.rng(8).standard_normal((5, 5))).astype("float32")).to(dev)
    n = x.numel()
    stream = torch.cuda.current_stream(dev)
    res = {}
    for name, f, k, ww in (("dilate3", ag.dilation2d, 3, None), ("erode3", ag.erosion2d, 3, None),
                           ("dilate7", ag.dilation2d, 7, None), ("erode7", ag.erosion2d, 7, None),
                           ("maxplus5", ag.dilation2d, 5, w)):
        xf = x.detach()
        fw = lambda: f(xf, k=k, w=ww)  # noqa: E731
        fw()
        tf = timed(fw, reps, stream) / reps
        xr = x.detach().requires_grad_(True)
        wr = ww.detach().requires_grad_(True) if ww is not None else None

        def fb():
            xr.grad = None
            if wr is not None:
                wr.grad = None
            y = f(xr, k=k, w=wr)
            y.backward(gy)
        fb()
        fb()
        tfb = timed(fb, reps, stream) / reps
        res[name] = {"fwd_ms": round(tf, 3), "fwd_gpx_s": round(n / (tf * 1e-3) / 1e9, 2),
                     "fwd_pct_hbm": round(100 * 8 * n / (tf * 1e-3) / 1e9 / HBM_PEAK_GBS, 2),
                     "fwdbwd_ms": round(tfb, 3), "fwdbwd_gpx_s": round(n / (tfb * 1e-3) / 1e9, 2),
                     "fwdbwd_pct_hbm": round(100 * 18 * n / (tfb * 1e-3) / 1e9 / HBM_PEAK_GBS, 2)}
    # the plain per-image path (non-differentiable Dilate per 256x256 image) for the 3x3 forward
    imgs = x.view(-1, 256, 256)
    outp = torch.empty_like(imgs)

    def plain():
        for i in range(imgs.shape[0]):
            pm.dilate(imgs[i], pm.static_nl_se.square2d, 1, out=outp[i])
    plain()
    tp = timed(plain, 1, stream)
    res["dilate3"]["plain_per_image_fwd_ms"] = round(tp, 3)
    res["dilate3"]["fwd_speedup_vs_per_image"] = round(tp / res["dilate3"]["fwd_ms"], 2)
    # the plain float32 square kernel on one image of the same pixel count (8192 x 16384, the
    # same 512 MiB viewed as one image), 3x3 and 7x7: the kernel-level comparison
    big = x.view(8192, 16384)
    outb = torch.empty_like(big)
    for name, it in (("dilate3", 1), ("dilate7", 3)):
        def plain_big(it=it):
            pm.dilate(big, pm.static_nl_se.square2d, it, out=outb)
        plain_big()
        tq = timed(plain_big, reps, stream) / reps
        res[name]["plain_square_kernel_same_px_ms"] = round(tq, 3)
        res[name]["fwd_vs_plain_square_kernel"] = round(tq / res[name]["fwd_ms"], 3)
    return res


def pmc_child(side, tw=0):
    """Workload profiled by measure_traffic: a plain elementwise uint8 kernel of known traffic
    (bitwise_not over 64 MiB: 64 MiB read + 64 MiB written) and the 8K dilation of one side."""
    import torch
    import poutre2_amd as pm
    from poutre2_amd import synth
    dev = torch.device("cuda:0")
    if tw:  # the tile width the parent run uses for this side (no tuning launches here)
        r = (side - 1) // 2
        pm._call("pm_tuning_set", 1, r, int(tw))
    base = torch.from_numpy(synth.uniform_u8((8192, 8192), 2)).to(dev)
    pairs = [(base.roll(i * 977, dims=1).contiguous(), torch.empty_like(base)) for i in range(3)]
    for i in range(6):
        a, b = pairs[i % 3]
        torch.bitwise_not(a, out=b)
    for i in range(6):
        a, b = pairs[i % 3]
        pm.dilate(a, pm.static_nl_se.square2d, (side - 1) // 2, out=b)
    torch.cuda.synchronize()
    print("pmc child done", flush=True)


def _pmc_pass(counter, side, tw, outdir):
    import csv
    import glob
    import subprocess
    env = dict(os.environ, TMPDIR="/tmp")
    cmd = ["rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", outdir, "-o", "run", "--",
           sys.executable, os.path.abspath(__file__), "--pmc-child", str(side), "--pmc-tw", str(tw)]
    subprocess.run(cmd, env=env, cwd="/tmp", stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=120,
                   check=False)
    files = glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise RuntimeError(f"rocprofv3 --pmc {counter}: no counter file")
    vals = {"copy": [], "dil": []}
    with open(files[0]) as f:
        for row in csv.DictReader(f):
            if row["Counter_Name"] != counter:
                continue
            name = row["Kernel_Name"]
            if "bitwise_not" in name:
                vals["copy"].append(float(row["Counter_Value"]))
            elif "sq_u8_kernel" in name:
                vals["dil"].append(float(row["Counter_Value"]))
    if len(vals["copy"]) < 2 or len(vals["dil"]) < 2:
        raise RuntimeError(f"rocprofv3 --pmc {counter}: kernels not found")
    # first dispatch of each kernel is a cold start; average the rest
    return (sum(vals["copy"][1:]) / (len(vals["copy"]) - 1), sum(vals["dil"][1:]) / (len(vals["dil"]) - 1))


def measure_traffic(side, tw):
    """HBM bytes per launch of the dominant kernel from two rocprofv3 --pmc passes (FETCH_SIZE,
    then WRITE_SIZE, KiB per dispatch), each scaled by the factor that turns the same counter on
    a plain elementwise copy of 64 MiB into the bytes that copy moves."""
    import shutil
    import tempfile
    if shutil.which("rocprofv3") is None:
        return None, "rocprofv3 not found"
    npx = 8192 * 8192
    out = {}
    tmp = tempfile.mkdtemp(prefix="pm_pmc_", dir="/tmp")
    try:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            copy_kib, dil_kib = _pmc_pass(counter, side, tw, os.path.join(tmp, counter))
            factor = npx / (copy_kib * 1024.0)
            out[counter] = {"copy_kib": round(copy_kib, 1), "kernel_kib": round(dil_kib, 1),
                            "factor": round(factor, 4), "bytes": dil_kib * 1024.0 * factor}
    except Exception as e:  # noqa: BLE001 - traffic is optional; report why it is missing
        return None, f"{type(e).__name__}: {e}"
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return out["FETCH_SIZE"]["bytes"] + out["WRITE_SIZE"]["bytes"], out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def _cpu_square(oracle, np, img, r, dil):
    """(2r+1)^2 window: the oracle's DilateX along rows, then along columns (numpy transposes,
    as the reference's DilateY transposes; any image shape)"""
    t = oracle.line_x(img, r, dil)
    return np.ascontiguousarray(oracle.line_x(np.ascontiguousarray(t.T), r, dil).T)


def cpu_baseline(torch, dev):
    """Single-thread oracle (the restatement of the reference's CPU algorithms, compiled -O3
    -mavx2 as the reference builds) on bounded samples of each workload, timed on this host:
      headline  one 8192^2 uint8 DilateX+DilateY per side 3..63 (ero_dil_line_se_t.hpp:49-280)
      C1        Dilate / Erode(SESquare2D, 1) on 1024^2 (ero_dil_static_se_t.hpp:375-472), 5 runs each
      C3        Reconstruction (hybrid scans + FIFO, mreconstruct_t.hpp:75-195) and leveling
                (priority queue, leveling_t.hpp:34-202) on the 4096^2 h = 32 workload, once each
      C4        the per-image loop (dilation + opening) over the first 96 of the 512 batch images,
                and Reconstruction over the first 64 of the 256 batch masks"""
    import numpy as np
    import oracle
    from poutre2_amd import synth
    legs = {}
    img = synth.uniform_u8((8192, 8192), 2)
    tot, npx = 0.0, 0
    for s in SIDES:
        r = (s - 1) // 2
        t0 = time.perf_counter()
        tmp = oracle.line_x(img, r, True)
        oracle.line_y(tmp, r, True)
        tot += time.perf_counter() - t0
        npx += img.size
    head = round(npx / tot / 1e9, 4)
    secs = tot
    a = synth.uniform_u8((1024, 1024), 1)
    for name, dil in (("dilate", True), ("erode", False)):
        t0 = time.perf_counter()
        for _ in range(5):
            oracle.static(a, "SESquare2D", dil)
        t = (time.perf_counter() - t0) / 5
        secs += 5 * t
        legs[f"C1_{name}_1024"] = {"ms": round(t * 1e3, 2), "gpx_s": round(a.size / t / 1e9, 4)}
    mask = synth.blurred_noise_mask(4096, 4096, 4, device=dev).cpu().numpy()
    marker = synth.hmax_marker(mask, 32)
    t0 = time.perf_counter()
    oracle.reconstruct("dilate", marker, mask, "SESquare2D")
    t = time.perf_counter() - t0
    legs["C3_reconstruct_hmax32_4096"] = {"ms": round(t * 1e3, 1), "gpx_s": round(mask.size / t / 1e9, 5)}
    g = _cpu_square(oracle, np, _cpu_square(oracle, np, mask, 7, False), 7, True)
    t0 = time.perf_counter()
    oracle.leveling(mask, g, "SESquare2D")
    t2 = time.perf_counter() - t0
    legs["C3_leveling_open15_4096"] = {"ms": round(t2 * 1e3, 1), "gpx_s": round(mask.size / t2 / 1e9, 5)}
    secs += t + t2
    shapes, radii, imgs = synth.batch_dilate_open_workload()
    t0 = time.perf_counter()
    n = 0
    for a, r in list(zip(imgs, radii))[:96]:
        _cpu_square(oracle, np, a, r, True)
        _cpu_square(oracle, np, _cpu_square(oracle, np, a, r, False), r, True)
        n += a.size
    t = time.perf_counter() - t0
    secs += t
    legs["C4_dilate_open_per_image"] = {"images": 96, "mpx": round(n / 1e6, 1), "ms": round(t * 1e3, 1),
                                        "gpx_s": round(n / t / 1e9, 4)}
    del imgs
    g6 = synth.rng(6)
    n, t = 0, 0.0
    for i in range(64):
        H, W = int(g6.integers(128, 1025)), int(g6.integers(128, 1025))
        m = synth.blurred_noise_mask(H, W, 1000 + i, device=dev).cpu().numpy()
        mk = synth.hmax_marker(m, 16)
        t0 = time.perf_counter()
        oracle.reconstruct("dilate", mk, m, "SESquare2D")
        t += time.perf_counter() - t0
        n += m.size
    secs += t
    legs["C4_reconstruct_per_image"] = {"images": 64, "mpx": round(n / 1e6, 1), "ms": round(t * 1e3, 1),
                                        "gpx_s": round(n / t / 1e9, 4)}
    return {"value": head, "unit": "Gpixel/s", "cores": 1, "kind": "port",
            "sample": "one 8192x8192 uint8 DilateX+DilateY per side 3,7,15,31,63 (van Herk rows + transposes), "
                      "oracle compiled -O3 -mavx2, 1 thread; legs: C1 1024^2 x5, C3 4096^2 once, "
                      "C4 first 96 images / 64 masks",
            "seconds": round(secs, 2), "cpu_model": cpu_model(), "threads": 1, "legs": legs}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sections", default="dilate,ops,recon,3d,types,small,batch,autograd,label")
    ap.add_argument("--no-traffic", action="store_true", help="skip the two rocprofv3 --pmc passes")
    ap.add_argument("--pmc-child", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--pmc-tw", type=int, default=0, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.pmc_child:
        pmc_child(args.pmc_child, args.pmc_tw)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the multi-GPU path is a design "
              "(row bands + halo exchange) and is not built in this version.", file=sys.stderr)
        sys.exit(2)
    import torch
    import poutre2_amd as pm
    dev = torch.device("cuda:0")
    sections = set(args.sections.split(","))
    log("bench: autotune (tile width per radius, untimed)")
    tuned = pm.autotune()
    log("bench: 8192x8192 u8 dilation, sides", SIDES)
    gpx, ms_step, per_side, roofline, pct = bench_dilate_8k(pm, torch, args.steps, args.warmup, dev)
    extra = {"dilate_per_side": per_side, "pct_hbm": pct}
    if not args.no_traffic:
        log("bench: rocprofv3 FETCH_SIZE / WRITE_SIZE passes on the dominant kernel")
        traffic, detail = measure_traffic(roofline["side"], pm.tuning()["aligned"][(roofline["side"] - 1) // 2 - 1])
        roofline["traffic"] = None if traffic is None else round(traffic)
        roofline["traffic_detail"] = detail
    if "ops" in sections:
        log("bench: erode/open/close/gradient 8K")
        extra["ops_8k"] = bench_ops_8k(pm, torch, dev)
    if "recon" in sections:
        log("bench: 4K reconstruction / leveling")
        extra["geodesy_4k"] = bench_recon_4k(pm, torch, dev)
    if "3d" in sections:
        log("bench: 512^3 cross3d / square3d")
        extra["volume_512"] = bench_3d(pm, torch, dev)
    if "types" in sections:
        log("bench: 4096^2 uint16/int32/float32 dilation")
        extra["types_4k"] = bench_types(pm, torch, dev)
    if "small" in sections:
        log("bench: 1024^2 square2d iter 1")
        extra["small_1k"] = bench_small_1k(pm, torch, dev)
    if "batch" in sections:
        log("bench: batched 512 dilate+open, 256 reconstructions")
        extra["batched"] = bench_batched(pm, torch, dev)
    if "label" in sections:
        log("bench: 4096^2 connected-component labelling")
        extra["label_4k"] = bench_label_4k(pm, torch, dev, cpu=not args.no_cpu)
    if "autograd" in sections:
        log("bench: autograd [64,32,256,256] fp32")
        extra["autograd"] = bench_autograd(pm, torch, dev)
    cpu = None
    if not args.no_cpu:
        log("bench: cpu baseline (single-thread oracle)")
        cpu = cpu_baseline(torch, dev)
    props = torch.cuda.get_device_properties(dev)
    extra["tile_width_table"] = tuned
    line = {
        "metric": "Gpixel/s & % HBM peak: 8K uint8 dilate, 4K reconstruct; batched fwd+bwd",
        "value": round(gpx, 3), "unit": "Gpixel/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8", "data": "synthetic",
        "config": {"workload": "8192x8192 uint8 flat dilation by square SE sides 3,7,15,31,63 (one step = 5 ops)",
                   "image": [8192, 8192], "se_sides": SIDES, "rotating_buffers_MiB": 384},
        "roofline": roofline, "cpu_baseline": cpu, "extra": extra,
        "device": {"name": props.name, "cus": props.multi_processor_count},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
