"""The reference's known-answer tests through the C++ API (include/poutre_amd.hpp), on host
and device images (the compare KATs of test/pixel_processing/compare.cpp and the arith
functions against the oracle included), plus the reference's error behaviour and SE-model KATs."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from tests import kats
from tests.test_oracle import COMPARE_KATS

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "poutre2_amd", "api_kats")


def _arr_str(pt, a):
    return " ".join(["Scalar", pt, str(a.ndim)] + [str(x) for x in a.shape] + [repr(v) for v in a.reshape(-1).tolist()])


def _img_str(d):
    return " ".join(["Scalar", d["ptype"], str(len(d["shape"]))] + [str(x) for x in d["shape"]] +
                    [str(v) for v in d["values"]])


def test_cpp_api_kats(tmp_path):
    if not os.path.exists(BIN):
        pytest.fail("poutre2_amd/api_kats is not built (python -m poutre2_amd.build)")
    lines = []
    for c in kats.cases(lambda c: c["op"] != "t_transpose"):
        op = c["op"]
        if op in ("t_Dilate", "t_Erode") and c["runtime_se"]:
            op = "rt_" + op[2:]
        ins = c["inputs"]
        # the 9th field is the reference's expected string, compared raw as its Catch2 test does
        # (REQUIRE_THAT(ImageToString(out), Equals(expected)))
        lines.append("|".join([op, c.get("se") or "", c.get("rec_type") or "", str(c.get("r") or ""),
                               str(c.get("h") or ""), _img_str(ins[0]), _img_str(ins[1]) if len(ins) > 1 else "",
                               _img_str(c["expected"]), c["expected_text"]]))
    # CompareImage: every whole-image KAT of the reference (test/pixel_processing/compare.cpp)
    for c in COMPARE_KATS:
        vals = ",".join(str(o.get("value", 0)) for o in (c["comp"], c["true"], c["false"]))
        imgs = ";".join(_img_str(o["image"]) for o in (c["comp"], c["true"], c["false"]) if "image" in o)
        lines.append("|".join([f"compare:{c['form']}:{oracle.CMP_OPS[c['op']]}", vals, "", "", "",
                               _img_str(c["input"]), imgs, _img_str(c["expected"])]))
    # the arith functions against the oracle on seeded images of every reference PType
    rng = np.random.default_rng(21)
    for pt, dt in (("GUINT8", np.uint8), ("GINT32", np.int32), ("GINT64", np.int64), ("F32", np.float32),
                   ("D64", np.float64)):
        a = rng.integers(0, 200, size=(7, 9)).astype(dt)
        b = rng.integers(0, 200, size=(7, 9)).astype(dt)
        if dt in (np.float32, np.float64):
            a, b = a + dt(0.5), b - dt(0.25)
        exp = {"ArithInvertImage": oracle.invert(a), "ArithSupImage": oracle.sup(a, b), "ArithInfImage": oracle.inf(a, b),
               "ArithSaturatedAddImage": oracle.sadd_image(a, b), "ArithSaturatedSubImage": oracle.ssub_image(a, b),
               "ArithSaturatedAddConstant": oracle.sadd_const(a, 77), "ArithSaturatedSubConstant": oracle.ssub_const(a, 77)}
        for op, e in exp.items():
            two = op.endswith("Image") and op != "ArithInvertImage"
            lines.append("|".join([op, "", "", "", "77" if "Constant" in op else "", _arr_str(pt, a),
                                   _arr_str(pt, b) if two else "", _arr_str(pt, e)]))
    # ConvertInto between PTypes (a C cast: numpy's astype) and Clone
    c = rng.integers(0, 120, size=(5, 8))
    for pin, din in (("GUINT8", np.uint8), ("GINT32", np.int32), ("D64", np.float64)):
        for pout, dout in (("GINT64", np.int64), ("F32", np.float32), ("GUINT8", np.uint8)):
            x = c.astype(din) + (din(0.75) if din is np.float64 else din(0))
            lines.append("|".join(["ConvertInto", "", "", "", "", _arr_str(pin, x), "", _arr_str(pout, x.astype(dout))]))
        lines.append("|".join(["Clone", "", "", "", "", _arr_str(pin, c.astype(din)), "", _arr_str(pin, c.astype(din))]))
    f = tmp_path / "cases.txt"
    f.write_text("\n".join(lines) + "\n")
    r = subprocess.run([BIN, str(f)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0
    # every KAT but the transpose, once on host images and once on device images
    assert "0 failures" in r.stdout and f"{2 * len(lines)} cases run" in r.stdout
