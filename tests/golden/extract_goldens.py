"""Extract the golden vectors that folax's own unit/integration tests hold for the FE hot path.

Reads the reference test files *as text* (Python `ast`, nothing is imported or run) and writes
`reference_goldens.json` next to this script: for every test function, the literal arrays it
assigns (inputs such as element coordinates and nodal vectors) and, in source order, every
`np.testing.assert_allclose(actual, expected, rtol, atol)` with its expected literal, the source
text of `actual`, and the tolerances.  The JSON is data (inputs and expected outputs); no
reference source is kept.

Run here (where /root/reference exists):  python tests/golden/extract_goldens.py
"""
import ast
import json
import os
import sys

REF = os.environ.get("FOLAX_REFERENCE", "/root/reference")
FILES = [
    "tests/unit/test_mechanical_loss.py",
    "tests/unit/test_neo_hooke_mechanical_loss.py",
    "tests/unit/test_geometries.py",
    "tests/unit/test_sensitivity_analysis.py",
    "tests/integration/test_mechanical_2D_sa.py",
]


def _literal(node):
    """Literal value of jnp.array([...]) / np.array([...]) / plain list, else None."""
    if isinstance(node, ast.Call) and node.args:
        f = node.func
        if isinstance(f, ast.Attribute) and f.attr == "array":
            node = node.args[0]
        else:
            return None
    try:
        return ast.literal_eval(node)
    except Exception:
        return None


def _kw(call, name, default=None):
    for k in call.keywords:
        if k.arg == name:
            try:
                return ast.literal_eval(k.value)
            except Exception:
                return default
    return default


def extract(path):
    tree = ast.parse(open(path).read())
    out = {}
    for fn in ast.walk(tree):
        if not isinstance(fn, (ast.FunctionDef,)) or not (fn.name.startswith("test") or fn.name == "setUp"):
            continue
        assigns, checks = {}, []
        for node in ast.walk(fn):
            if isinstance(node, ast.Assign) and len(node.targets) == 1:
                t = node.targets[0]
                name = ast.unparse(t)
                v = _literal(node.value)
                if v is not None and isinstance(v, list):
                    assigns[name] = v
            if isinstance(node, ast.Call) and ast.unparse(node.func).endswith("assert_allclose"):
                exp = _literal(node.args[1])
                if exp is None:
                    continue
                checks.append({"line": node.lineno, "actual": ast.unparse(node.args[0]), "expected": exp,
                               "rtol": _kw(node, "rtol", 1e-7), "atol": _kw(node, "atol", 0.0)})
        checks.sort(key=lambda c: c["line"])
        out[fn.name] = {"assigns": assigns, "checks": checks}
    return out


def main():
    res = {}
    for rel in FILES:
        p = os.path.join(REF, rel)
        if not os.path.exists(p):
            print("missing", p, file=sys.stderr)
            sys.exit(1)
        res[rel] = extract(p)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_goldens.json")
    with open(dst, "w") as f:
        json.dump(res, f, indent=1)
    print("wrote", dst, sum(len(t["checks"]) for m in res.values() for t in m.values()), "checks")


if __name__ == "__main__":
    main()
