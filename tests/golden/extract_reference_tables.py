"""Write ``reference_tables.json``: numeric tables the reference itself holds.

The reference repository has no stored expected outputs; the data it does hold for
the DG diffusion path are literal tables in its sources (quadrature rules, face-point
alignment, P3 nodal coordinates used by its shape-function test scripts, the
P1->P3 interpolation tables, face-node lists, the face-parameter derivatives and the
solver constants).  This script reads those literals *as text* -- it parses the
bracketed list expressions with ``ast`` and evaluates only numeric literals and
+ - * / -- and never imports or runs reference code.

Usage:  python tests/golden/extract_reference_tables.py [/path/to/reference]
The output JSON is committed; tests compare the oracle's tables against it.
"""
import ast
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
Z02 = os.path.join(REF, "z02-implicit", "z06-DGtoCG")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_tables.json")


def _eval(node):
    if isinstance(node, ast.Expression):
        return _eval(node.body)
    if isinstance(node, (ast.List, ast.Tuple)):
        return [_eval(e) for e in node.elts]
    if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
        return node.value
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
        v = _eval(node.operand)
        return -v if isinstance(node.op, ast.USub) else v
    if isinstance(node, ast.BinOp) and isinstance(node.op, (ast.Add, ast.Sub, ast.Mult, ast.Div)):
        a, b = _eval(node.left), _eval(node.right)
        return {ast.Add: a + b, ast.Sub: a - b, ast.Mult: a * b, ast.Div: a / b}[type(node.op)]
    raise ValueError("non-numeric expression in table: %s" % ast.dump(node))


def _lines(fname):
    with open(os.path.join(Z02, fname)) as fh:
        return fh.read().split("\n")


def list_at(fname, lineno, after=None):
    """Evaluate the first bracketed list starting on/after 1-based ``lineno``
    (optionally after the substring ``after`` on that line)."""
    lines = _lines(fname)
    text = "\n".join(lines[lineno - 1:])
    start = text.index(after) + len(after) if after else 0
    i = text.index("[", start)
    depth = 0
    for j in range(i, len(text)):
        if text[j] == "[":
            depth += 1
        elif text[j] == "]":
            depth -= 1
            if depth == 0:
                break
    body = text[i:j + 1].replace("\\\n", "\n")
    body = re.sub(r"#[^\n]*", "", body)
    return _eval(ast.parse(body, mode="eval"))


def scalar_at(fname, first, last, name):
    """Evaluate ``name = <expr>`` found between 1-based lines first..last."""
    for ln in _lines(fname)[first - 1:last]:
        m = re.match(r"\s*%s\s*=\s*([^#]+)" % re.escape(name), ln)
        if m:
            return _eval(ast.parse(m.group(1).strip(), mode="eval"))
    raise KeyError(name)


def flat(x):
    out = []
    for e in x:
        out.extend(flat(e) if isinstance(e, list) else [e])
    return out


def main():
    sf = "shape_function.py"
    tables = {
        "source": "literal tables read from the reference sources (no code executed)",
        # Keast #7 tet rule (shape_function.py:29-64); weights are x 1/6 in the reference
        "tet24_L": list_at(sf, 29),
        "tet24_weight_unscaled": list_at(sf, 55),
        # Strang-Fix #9 triangle rule (shape_function.py:271-308); weights are x 0.5
        "tri12_xy": list_at(sf, 273),
        "tri12_sweight_unscaled": list_at(sf, 287, after="np.asarray("),
        "tri12_alignment_1based": list_at(sf, 304),
        # 4-point line rule (shape_function.py:399-420)
        "line4_a": scalar_at(sf, 392, 420, "a"),
        "line4_b": scalar_at(sf, 392, 420, "b"),
        "line4_w1": scalar_at(sf, 392, 420, "w1"),
        "line4_w2": scalar_at(sf, 392, 420, "w2"),
        "line4_alignment": list_at(sf, 417),
        # P3 nodal coordinates of the reference elements used by the reference's
        # shape-function test scripts
        "tet20_nodes": list_at("test_3d_shape_function.py", 52),
        "tri10_nodes": list_at("test_surface_shape_func.py", 46),
        # P1DG -> P3DG element prolongators (mesh_init.py:224-235, 610-631)
        "tri10_p1_prolong": list_at("mesh_init.py", 224),
        "tet20_p1_prolong": list_at("mesh_init.py", 610),
        # local nodes on each face (mesh_init.py:840-889)
        "tri10_face_iloc": list_at("mesh_init.py", 844),
        "tet20_face_iloc": list_at("mesh_init.py", 865),
        # d(xi)/d(face params) (function_space.py:200-213)
        "tet_drst_duv": list_at("function_space.py", 201),
        "tri_drst_duv": list_at("function_space.py", 208),
        # solver constants (config.py)
        "eta_e": scalar_at("config.py", 245, 245, "eta_e"),
        "jac_wei": scalar_at("config.py", 108, 108, "jac_wei"),
        "pre_smooth_its": scalar_at("config.py", 112, 112, "pre_smooth_its"),
        "post_smooth_its": scalar_at("config.py", 113, 113, "post_smooth_its"),
    }
    # shape checks so that a shifted line number fails loudly
    assert len(tables["tet24_L"]) == 72 and len(tables["tet24_weight_unscaled"]) == 24
    assert len(tables["tri12_xy"]) == 24 and len(tables["tri12_sweight_unscaled"]) == 12
    assert len(tables["tri12_alignment_1based"]) == 36
    assert len(tables["tet20_nodes"]) == 60 and len(tables["tri10_nodes"]) == 20
    assert len(flat(tables["tet20_p1_prolong"])) == 80
    assert len(flat(tables["tri10_p1_prolong"])) == 30
    assert len(tables["tet20_face_iloc"]) == 40 and len(tables["tri10_face_iloc"]) == 12
    with open(OUT, "w") as fh:
        json.dump(tables, fh, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
