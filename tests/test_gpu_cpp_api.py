"""The reference's known-answer tests through the C++ API (include/poutre_amd.hpp), on host
and device images, plus the reference's error behaviour and SE-model KATs."""
import os
import subprocess

import pytest

from tests import kats

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "poutre2_amd", "api_kats")


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
        lines.append("|".join([op, c.get("se") or "", c.get("rec_type") or "", str(c.get("r") or ""),
                               str(c.get("h") or ""), _img_str(ins[0]), _img_str(ins[1]) if len(ins) > 1 else "",
                               _img_str(c["expected"])]))
    f = tmp_path / "cases.txt"
    f.write_text("\n".join(lines) + "\n")
    r = subprocess.run([BIN, str(f)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0
    # every KAT but the transpose, once on host images and once on device images
    assert "0 failures" in r.stdout and f"{2 * len(lines)} cases run" in r.stdout
