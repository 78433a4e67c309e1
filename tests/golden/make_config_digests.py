"""Write tests/golden/config_digests.json: sha256 digests of the oracle's outputs on the five
benchmark workloads at full size (tests/configs.py), and of their inputs.

Run in this container (CPU only; several minutes):
    python tests/golden/make_config_digests.py
The GPU parity tests (tests/test_gpu_configs.py) regenerate the same inputs, check their
digests, run the device path and compare the digests of its outputs.

Oracles (test infrastructure, oracle/oracle.h):
  * squares: oracle.line_x along rows, transposed, line_x again, transposed back -- the
    reference's DilateX + DilateY (ero_dil_line_se_t.hpp:49-280), equal to Dilate(SESquare2D, r)
  * 3D: oracle.static applied iter times (ero_dil_static.cpp:229-305 iteration semantics)
  * reconstruction: oracle.reconstruct (mreconstruct_t.hpp:75-195, hybrid algorithm)
  * leveling: oracle.leveling (leveling.cpp:231-237, priority-queue restatement)
  * autograd: tests/autograd_ref.py (numpy, tap by tap, first extremum in raster order)
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from poutre2_amd import synth  # noqa: E402
from tests import autograd_ref, configs  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "config_digests.json")
D = configs.digest
T0 = time.time()


def log(*a):
    print(f"[{time.time() - T0:7.1f}s]", *a, flush=True)


def square(img, r, dil):
    """(2r+1)^2 window of a 2D image: the oracle's line pass along rows, then along columns
    (numpy transposes: oracle.transpose restates the reference's square-only t_transpose)"""
    t = oracle.line_x(img, r, dil)
    return np.ascontiguousarray(oracle.line_x(np.ascontiguousarray(t.T), r, dil).T)


def ops2d(img, r):
    d, e = square(img, r, True), square(img, r, False)
    return {"dilate": d, "erode": e, "opening": square(e, r, True), "closing": square(d, r, False),
            "gradient": (d.astype(np.int16) - e.astype(np.int16)).astype(np.uint8)}


def c1():
    img = synth.uniform_u8((1024, 1024), 1)
    return {"input": D(img), "dilate": D(oracle.static(img, "SESquare2D", True)),
            "erode": D(oracle.static(img, "SESquare2D", False))}


def c2_2d():
    img = synth.uniform_u8((8192, 8192), 2)
    out = {"input": D(img)}
    for s in configs.SIDES:
        res = ops2d(img, (s - 1) // 2)
        for op in configs.OPS_2D:
            out[f"{op}_{s}"] = D(res[op])
        log("C2 side", s)
    return out


def c2_3d():
    vol = synth.uniform_u8((512, 512, 512), 3)
    out = {"input": D(vol)}
    for se in configs.SES_3D:
        for dil, name in ((True, "dilate"), (False, "erode")):
            cur = vol
            for it in configs.ITERS_3D:
                cur = oracle.static(cur, se, dil)
                out[f"{name}_{se}_{it}"] = D(cur)
            log("C2 3D", se, name)
    return out


def c3():
    mask = synth.blurred_noise_mask(4096, 4096, 4)
    marker = synth.hmax_marker(mask, 32)
    rec = oracle.reconstruct("dilate", marker, mask, "SESquare2D")
    log("C3 reconstruction")
    g = square(square(mask, 7, False), 7, True)
    lev = oracle.leveling(mask, g, "SESquare2D")
    log("C3 leveling")
    return {"mask": D(mask), "marker": D(marker), "reconstruct": D(rec), "opening_7": D(g), "leveling": D(lev)}


def c4():
    shapes, radii, imgs = synth.batch_dilate_open_workload()
    di, dd, do = [], [], []
    for a, r in zip(imgs, radii):
        d = square(a, r, True)
        o = square(square(a, r, False), r, True)
        di.append(D(a))
        dd.append(D(d))
        do.append(D(o))
    log("C4 dilate+open")
    markers, masks = synth.batch_reconstruction_workload()
    rm, rk, ro = [], [], []
    for mk, ms in zip(markers, masks):
        rm.append(D(ms))
        rk.append(D(mk))
        ro.append(D(oracle.reconstruct("dilate", mk, ms, "SESquare2D")))
    log("C4 reconstructions")
    return {"dilate_open": {"n": len(imgs), "mpx": sum(a.size for a in imgs) / 1e6,
                            "inputs": configs.digest_of_digests(di),
                            "dilate": configs.digest_of_digests(dd), "opening": configs.digest_of_digests(do),
                            "dilate_each": [x[:16] for x in dd], "opening_each": [x[:16] for x in do]},
            "reconstruct": {"n": len(masks), "masks": configs.digest_of_digests(rm),
                            "markers": configs.digest_of_digests(rk), "out": configs.digest_of_digests(ro),
                            "out_each": [x[:16] for x in ro]}}


def c5():
    x = configs.c5_input()
    w = configs.c5_w()
    gy = configs.c5_gy()
    out = {"input": D(x), "w": D(w), "gy": D(gy)}
    for name, k, with_w, dil in configs.C5_OPS:
        y, tap = autograd_ref.forward_with_taps(x, k=k, w=w if with_w else None, dilate=dil)
        gx, gw = autograd_ref.backward_from_taps(tap, gy, k)
        e = {"y": D(y), "gx": D(gx.astype(np.float32))}
        if with_w:
            e["gw"] = [[int(v) for v in row] for row in (gw if dil else -gw)]
        out[name] = e
        del y, tap, gx
        log("C5", name)
    return out


def main():
    res = {"about": "sha256 of the oracle's outputs on the benchmark workloads (tests/configs.py); "
                    "written by tests/golden/make_config_digests.py"}
    only = set(sys.argv[1:])
    old = json.load(open(OUT)) if os.path.exists(OUT) and only else {}
    res.update({k: v for k, v in old.items() if k != "about"})
    for name, fn in (("C1", c1), ("C2_2d", c2_2d), ("C2_3d", c2_3d), ("C3", c3), ("C4", c4), ("C5", c5)):
        if only and name not in only:
            continue
        log("start", name)
        res[name] = fn()
        with open(OUT, "w") as f:
            json.dump(res, f, indent=1)
        log("done", name)


if __name__ == "__main__":
    main()
