"""Loader for the reference's known-answer tests, transcribed into tests/golden/kats.json
by tests/golden/extract_kats.py."""
import json
import os

import numpy as np

PTYPES = {"GUINT8": np.uint8, "GINT32": np.int32, "GINT64": np.int64, "F32": np.float32, "D64": np.float64}
_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "kats.json")


def image(d):
    return np.array(d["values"], dtype=PTYPES[d["ptype"]]).reshape(d["shape"])


def cases(pred=None):
    with open(_PATH) as f:
        data = json.load(f)["cases"]
    out = [c for c in data if pred is None or pred(c)]
    return out


def case_id(c):
    return f"{os.path.basename(c['file'])}:{c['line']}:{c['name']}#{c['sub']}"
