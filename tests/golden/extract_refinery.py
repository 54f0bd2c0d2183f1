"""Writes refinery_golden.json: the input meshes and the expected one-step refinements that the
reference's own refinery tests hold as literal arrays.

Sources (read as text, nothing is compiled or run):
  kernel/geometry/test_aux/standard_quad.cpp  create_quad_mesh_2d / validate_refined_quad_mesh_2d
                                              (orientations 0..3)
  kernel/geometry/test_aux/standard_hexa.cpp  create_hexa_mesh_3d / validate_refined_hexa_mesh_3d
                                              (orientations 0..2)
  kernel/geometry/test_aux/tetris_quad.cpp    create_tetris_mesh_2d / validate_refined_tetris_mesh_2d
  kernel/geometry/test_aux/tetris_hexa.cpp    create_tetris_mesh_3d / validate_refined_tetris_mesh_3d
(used by kernel/geometry/standard_refinery-test-conf-{quad,hexa}.cpp).

Usage: python tests/golden/extract_refinery.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
AUX = os.path.join(REF, "kernel", "geometry", "test_aux")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "refinery_golden.json")

ARRAY = re.compile(r"(?:static\s+)?(?:const\s+)?(Real|Index)\s+(\w+)\s*\[[^\]]*\]\s*=\s*\{(.*?)\};", re.S)


def function_body(text, signature):
    """Text of the function whose definition line contains `signature` (brace matched)."""
    at = text.index(signature)
    start = text.index("{", at)
    depth = 0
    for i in range(start, len(text)):
        if text[i] == "{":
            depth += 1
        elif text[i] == "}":
            depth -= 1
            if depth == 0:
                return text[start:i + 1]
    raise ValueError(signature)


def arrays(body):
    body = re.sub(r"//[^\n]*", "", body)
    out = {}
    for kind, name, items in ARRAY.findall(body):
        vals = [v.strip() for v in items.replace("\n", " ").split(",") if v.strip()]
        out[name] = [float(v) for v in vals] if kind == "Real" else [int(v) for v in vals]
    return out


def mesh(arr, dim, suffix=""):
    keys = ["vtx", "v_e", "v_q", "e_q"] + (["v_c", "e_c", "q_c"] if dim == 3 else [])
    # the tetris hexa mesh names its quads-at-hexa array f_c
    return {k: arr[k + suffix] if k + suffix in arr else arr["f_c" + suffix] for k in keys}


def main():
    cases = []
    for fname, dim, create, validate, norient in (
            ("standard_quad.cpp", 2, "create_quad_mesh_2d(", "validate_refined_quad_mesh_2d(", 4),
            ("standard_hexa.cpp", 3, "create_hexa_mesh_3d(", "validate_refined_hexa_mesh_3d(", 3)):
        text = open(os.path.join(AUX, fname)).read()
        ca = arrays(function_body(text, create))
        va = arrays(function_body(text, validate))
        for o in range(norient):
            cases.append({"name": f"{fname}:orientation{o}", "dim": dim,
                          "coarse": mesh(ca, dim, str(o)), "fine": mesh(va, dim, str(o))})
    for fname, dim, create, validate in (
            ("tetris_quad.cpp", 2, "create_tetris_mesh_2d(", "validate_refined_tetris_mesh_2d("),
            ("tetris_hexa.cpp", 3, "create_tetris_mesh_3d(", "validate_refined_tetris_mesh_3d(")):
        text = open(os.path.join(AUX, fname)).read()
        ca = arrays(function_body(text, create))
        va = arrays(function_body(text, validate))
        cases.append({"name": f"{fname}", "dim": dim, "coarse": mesh(ca, dim), "fine": mesh(va, dim)})
    with open(OUT, "w") as f:
        json.dump({"source": "kernel/geometry/test_aux/{standard_quad,standard_hexa,tetris_quad,tetris_hexa}.cpp",
                   "cases": cases}, f, separators=(",", ":"))
    print(f"wrote {OUT}: {len(cases)} cases")


if __name__ == "__main__":
    main()
