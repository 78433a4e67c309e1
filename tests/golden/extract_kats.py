"""Transcribe the reference's inline known-answer tests (KATs) into data fixtures.

The reference keeps its golden vectors inline in Catch2 tests as strings in the
``ImageFromString`` format ``"<Scalar> <GUINT8|GINT32|GINT64|F32|D64> <rank> <dims...> <values...>"``
(src/base/image_interface.cpp:407-425,616-650).  This script reads those test
sources as text, pulls out the input images, the operation called and the expected
output string, and writes them as plain data to ``kats.json`` next to this file.
No reference code is executed or copied: only the numbers and the name of the
operation each case calls are kept.

Sources read (relative to the reference root):
  test/low_level_morpho/ero_dil_static_se_t.cpp   (t_Dilate/t_Erode by Common_NL_SE)
  test/low_level_morpho/ero_dil_runtime_se.cpp    (t_Dilate/t_Erode by runtime neighbor_list_t)
  test/low_level_morpho/ero_dil_line_se.cpp       (t_DilateX/ErodeX/DilateY/ErodeY)
  test/geodesy/mreconstruct.cpp                   (geo::Reconstruction)
  test/geodesy/leveling.cpp                       (low_/high_leveling, leveling)
  test/geodesy/extrema.cpp                        (h_minima/h_maxima/h_concave/h_convex/dynamic_pseudo_*)
  test/pixel_processing/transpose.cpp             (t_transpose)
  test/label/label_binary.cpp                     (label::label_binary)
  test/label/label_flat_zones.cpp                 (label::label_flat_zones)
  test/label/label_extrema.cpp                    (label::label_minima / label_maxima)

and, into ``compare_kats.json``, the whole-image cases of
  test/pixel_processing/compare.cpp                (CompareImage, every operand form)

and, into ``iostring_kats.json``, the fixture-format cases of
  test/base/image_iostring.cpp                     (ImageFromString / ImageToString: the input
                                                    or expected string, the pixel type, compound
                                                    type, shape, and the pixel values or the
                                                    fill value)

Usage: python tests/golden/extract_kats.py [/root/reference]
"""
import json
import os
import re
import sys

FILES = [
    "test/low_level_morpho/ero_dil_static_se_t.cpp",
    "test/low_level_morpho/ero_dil_runtime_se.cpp",
    "test/low_level_morpho/ero_dil_line_se.cpp",
    "test/geodesy/mreconstruct.cpp",
    "test/geodesy/leveling.cpp",
    "test/geodesy/extrema.cpp",
    "test/pixel_processing/transpose.cpp",
    "test/label/label_binary.cpp",
    "test/label/label_flat_zones.cpp",
    "test/label/label_extrema.cpp",
]

OPS = [
    "t_DilateX", "t_ErodeX", "t_DilateY", "t_ErodeY", "t_Dilate", "t_Erode",
    "Reconstruction", "low_leveling", "high_leveling", "leveling",
    "h_minima", "h_maxima", "h_concave", "h_convex",
    "dynamic_pseudo_closing", "dynamic_pseudo_opening", "t_transpose",
    "label_binary", "label_flat_zones", "label_minima", "label_maxima",
]

_STR = r'(?:R"\((?P<raw>.*?)\)"|(?P<cooked>"(?:[^"\\]|\\.|\\\n)*"))'
TOKEN_RE = re.compile(
    r'(?P<tc>TEST_CASE\("(?P<tcname>[^"]*)")'
    r'|(?P<ifs>ImageFromString\(\s*' + _STR.replace("raw", "raw1").replace("cooked", "cooked1") + r')'
    r'|(?P<exp>expected\s*=\s*' + _STR.replace("raw", "raw2").replace("cooked", "cooked2") + r')'
    r'|(?P<const>static_cast<poutre::p\w+>\((?P<cval>-?\d+)\))'
    r'|(?P<call>(?:poutre::)?(?:llm::details::|geo::|details::|label::)?(?P<op>' + "|".join(OPS) + r')\((?P<args>[^;]*)\);)',
    re.S,
)


def _cooked(s):
    """Decode a C string literal that may use backslash-newline continuations."""
    body = s[1:-1]
    body = body.replace("\\\n", "")
    return body.encode().decode("unicode_escape")


def parse_image(text):
    toks = text.split()
    compound, ptype, rank = toks[0], toks[1], int(toks[2])
    shape = [int(t) for t in toks[3:3 + rank]]
    n = 1
    for d in shape:
        n *= d
    vals = toks[3 + rank:]
    conv = float if ptype in ("F32", "D64") else int
    values = [conv(v) for v in vals]
    return {"compound": compound, "ptype": ptype, "shape": shape,
            "values": values[:n], "n_listed": len(values)}


def extract(ref_root):
    cases = []
    for rel in FILES:
        path = os.path.join(ref_root, rel)
        src = open(path).read()
        name = None
        inputs, call, h = [], None, None
        sub = 0
        for m in TOKEN_RE.finditer(src):
            line = src.count("\n", 0, m.start()) + 1
            if m.group("tc"):
                name, inputs, call, h, sub = m.group("tcname"), [], None, None, 0
            elif m.group("ifs"):
                raw = m.group("raw1")
                txt = raw if raw is not None else _cooked(m.group("cooked1"))
                inputs.append(parse_image(txt))
            elif m.group("const"):
                h = int(m.group("cval"))
            elif m.group("call"):
                op = m.group("op")
                args = m.group("args")
                se = re.search(r"\b(SE\w+)", args)
                rec = re.search(r"reconstruction_type::(\w+)", args)
                rr = None
                if op in ("t_DilateX", "t_ErodeX", "t_DilateY", "t_ErodeY"):
                    rr = int(args.split(",")[1].strip())
                call = {"op": op, "se": se.group(1) if se else None,
                        "rec_type": rec.group(1) if rec else None, "r": rr,
                        "runtime_se": "Common_NL_SE" not in args and se is not None,
                        "call_line": line}
            elif m.group("exp"):
                raw = m.group("raw2")
                txt = raw if raw is not None else _cooked(m.group("cooked2"))
                exp = parse_image(txt)
                n = 1
                for d in exp["shape"]:
                    n *= d
                case = {"file": rel, "name": name, "sub": sub, "line": line,
                        "inputs": inputs, "expected": exp, "expected_text": txt,
                        "expected_complete": exp["n_listed"] >= n}
                case.update(call or {})
                if h is not None:
                    case["h"] = h
                cases.append(case)
                sub += 1
                inputs, call, h = [], None, None
    return cases


COMPARE_FILE = "test/pixel_processing/compare.cpp"
_CMP_CALL = re.compile(r'(?:t_CompareImage_|ViewCompare_)(?P<form>[si]{3})\w*\(\s*(?P<args>[^;]*)\);', re.S)
_INT_CONST = re.compile(r'\b(?P<name>\w+)\s*=\s*(?P<val>-?\d+)\s*;')


def extract_compare(ref_root):
    """Cases of compare.cpp on whole images: the input and image operands in the order the test
    creates them (input, comparand, true, false), scalar operands as literals or the named
    constants the test defines, the CompOpType and the expected image.  Cases on sub-views
    (sections) are left out: they compare part of an image."""
    src = open(os.path.join(ref_root, COMPARE_FILE)).read()
    heads = [m for m in re.finditer(r'TEST_CASE\("(?P<name>[^"]*)"', src)]
    cases = []
    for k, hm in enumerate(heads):
        body = src[hm.end():heads[k + 1].start() if k + 1 < len(heads) else len(src)]
        line0 = src.count("\n", 0, hm.start()) + 1
        call = _CMP_CALL.search(body)
        if call is None:
            continue
        args = [a.strip() for a in re.sub(r"//[^\n]*", "", call.group("args")).replace("\n", " ").split(",")]
        if args[0].startswith("s_") or args[-1].startswith("s_"):
            continue  # sections
        images = []
        for m in re.finditer(r'ImageFromString\(\s*' + _STR, body, re.S):
            raw = m.group("raw")
            images.append(parse_image(raw if raw is not None else _cooked(m.group("cooked"))))
        exp = re.search(r'expected\s*=\s*' + _STR, body, re.S)
        if exp is None:
            continue
        consts = {m.group("name"): int(m.group("val")) for m in _INT_CONST.finditer(body[:call.start()])}
        op = re.search(r"CompOpType::CompOp(\w+)", args[1]).group(1)
        operands, nxt = [], 1
        for kind, a in zip(call.group("form"), args[2:5]):
            if kind == "i":
                operands.append({"image": images[nxt]})
                nxt += 1
            else:
                operands.append({"value": int(a) if re.fullmatch(r"-?\d+", a) else consts[a]})
        expected = parse_image(exp.group("raw") if exp.group("raw") is not None else _cooked(exp.group("cooked")))
        cases.append({"file": COMPARE_FILE, "name": hm.group("name"), "line": line0,
                      "form": call.group("form"), "op": op, "input": images[0],
                      "comp": operands[0], "true": operands[1], "false": operands[2], "expected": expected})
    return cases


_PT = {"GrayUINT8": "GUINT8", "GrayINT32": "GINT32", "GrayINT64": "GINT64", "F32": "F32", "D64": "D64"}
_CT_OF_STORAGE = {"pUINT8": "GUINT8", "pINT32": "GINT32", "pINT64": "GINT64", "pFLOAT": "F32", "pDOUBLE": "D64"}
_NUM = re.compile(r"-?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?")


def _nums(text):
    return [float(t) if ("." in t or "e" in t or "E" in t) else int(t) for t in _NUM.findall(text)]


def extract_iostring(ref_root):
    rel = "test/base/image_iostring.cpp"
    src = open(os.path.join(ref_root, rel)).read()
    heads = list(re.finditer(r'TEST_CASE\("([^"]*)"[^)]*\)', src))
    cases = []
    for i, h in enumerate(heads):
        body = src[h.end():heads[i + 1].start() if i + 1 < len(heads) else len(src)]
        line = src[:h.start()].count("\n") + 1
        strings = [_cooked(m.group(0)) for m in re.finditer(r'"(?:[^"\\]|\\.|\\\n)*"', body)]
        ptype = _PT[re.search(r"PType::PType_(\w+)", body).group(1)]
        ctype = re.search(r"CompoundType::CompoundType_(\w+)", body).group(1)
        case = {"file": rel, "line": line, "name": h.group(1), "ptype": ptype, "compound": ctype}
        if "ImageFromString" in body:
            case["kind"] = "read"
            case["text"] = strings[0]
            case["rank"] = int(re.search(r"GetRank\(\)\s*==\s*(\d+)", body).group(1))
            case["shape"] = _nums(re.search(r"expectedcoords\s*=\s*\{([^}]*)\}", body).group(1))
            case["values"] = _nums(re.search(r"expectedvalues\s*=\s*\{(.*?)\};", body, re.S).group(1))
        else:
            case["kind"] = "write"
            m = re.search(r"image_t<(.*?)>\s+img\(\{([^}]*)\}\)", body)
            case["shape"] = _nums(m.group(2))
            fill = re.search(r"img\.fill\((.*?)\);", body).group(1)
            if re.fullmatch(r"\s*\w+\s*", fill) and not _NUM.fullmatch(fill.strip()):
                fill = re.search(r"\b" + fill.strip() + r"\s*=\s*([^;]+);", body).group(1)
            case["fill"] = _nums(fill)
            case["expected"] = strings[0]
        cases.append(case)
    return cases


def main():
    ref_root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    cases = extract(ref_root)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kats.json")
    with open(out, "w") as f:
        json.dump({"source": "ThomasRetornaz/poutre2 inline KATs", "cases": cases}, f, indent=1)
    cmp = extract_compare(ref_root)
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "compare_kats.json"), "w") as f:
        json.dump({"source": "ThomasRetornaz/poutre2 inline KATs (compare.cpp)", "cases": cmp}, f, indent=1)
    ios = extract_iostring(ref_root)
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "iostring_kats.json"), "w") as f:
        json.dump({"source": "ThomasRetornaz/poutre2 inline KATs (image_iostring.cpp)", "cases": ios}, f, indent=1)
    print(f"wrote {len(cases)} cases to {out}, {len(cmp)} compare cases, {len(ios)} iostring cases")


if __name__ == "__main__":
    main()
