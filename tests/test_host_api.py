"""CPU: the C++ API's host-side behaviour through tests/cpp/host_kats.cpp (built with the
library): the reference's fixture-format KATs (test/base/image_iostring.cpp:28-242, transcribed
to tests/golden/iostring_kats.json by tests/golden/extract_kats.py) compared raw, the SE-model
KATs (test/structuring_element/neighbor_list_se_t.cpp:20-69), round trips and errors."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "poutre2_amd", "host_kats")
KATS = os.path.join(ROOT, "tests", "golden", "iostring_kats.json")


def _csv(v):
    return ",".join(repr(x) for x in v)


def test_iostring_and_se_kats(tmp_path):
    if not os.path.exists(BIN):
        pytest.fail("poutre2_amd/host_kats is not built (python -m poutre2_amd.build)")
    with open(KATS) as f:
        cases = json.load(f)["cases"]
    assert len(cases) == 8
    lines = []
    for c in cases:
        if c["kind"] == "read":
            lines.append("|".join(["read", c["text"], c["ptype"], c["compound"], str(c["rank"]), _csv(c["shape"]),
                                   _csv(c["values"])]))
        else:
            lines.append("|".join(["write", c["ptype"], c["compound"], _csv(c["shape"]), _csv(c["fill"]),
                                   c["expected"]]))
    path = tmp_path / "cases.txt"
    path.write_text("\n".join(lines) + "\n")
    r = subprocess.run([BIN, str(path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "8 fixture cases" in r.stdout and " 0 failures" in r.stdout
