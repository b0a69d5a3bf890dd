"""Extract the PyElastica golden vectors held by the reference's FullRodTest
sources into JSON fixtures (data only: inputs and expected outputs).

The reference keeps its golden data inline in Catch2 test sources:
    tests/Unit/Systems/CosseratRods/FullRodTest/Test_PerturbedRod.cpp
    tests/Unit/Systems/CosseratRods/FullRodTest/Test_PerturbedRodStep.cpp
    tests/Unit/Systems/CosseratRods/FullRodTest/Test_PerturbedRodStepWithoutDamping.cpp
Each array is declared as
    blaze::Dynamic{Vector,Matrix,Tensor}<elastica::real_t> NAME( dims ); NAME = { ... };
with elements written either as std::strtod("<hex-float>", nullptr) or as
plain decimal literals.  This script reads those files as TEXT (nothing is
compiled, imported or run), converts each element with float.fromhex / float,
and writes one JSON file per source file next to this script.

For every test case it records:
  * the material / straight-rod initializer parameters,
  * whether rods share one block (`Model simple_rod{}`) or get separate
    blocks (RestrictSizeAcrossBlockTypes{8} with 5-element rods),
  * the *_REPLACED arrays and the ordered list of `blocks::get<tag>(rod) = X`
    assignments (with the shrink_using_diagonal flag, which keeps the
    diagonal of a 3x3xn tensor, FullRodTest/Helpers.cpp:19-30),
  * every SECTION with its approx epsilon, its step count and its *_PYEL
    expected arrays; for InitializeStraightRod the single force evaluation.

Run from anywhere:  python tests/golden/extract_golden.py
"""
import ast
import json
import operator
import os
import re
import sys

REF = "/root/reference/tests/Unit/Systems/CosseratRods/FullRodTest"
OUT = os.path.dirname(os.path.abspath(__file__))
FILES = [
    "Test_PerturbedRod.cpp",
    "Test_PerturbedRodStep.cpp",
    "Test_PerturbedRodStepWithoutDamping.cpp",
]

_DECL = re.compile(
    r"blaze::Dynamic(Vector|Matrix|Tensor)<elastica::real_t>\s+(\w+)\(\s*([^)]*)\);\s*"
    r"(\w+)\s*=\s*")
_ASSIGN = re.compile(
    r"blocks::get<elastica::tags::(\w+)>\((\w+)\)\s*=\s*"
    r"(shrink_using_diagonal\()?(\w+)\)?;")
_CASE = re.compile(r'ELASTICA_TEST_CASE\(\s*"([^"]+)"')
_SECTION = re.compile(r'SECTION\("(\w+)"\)')
_STRTOD = re.compile(r'std::strtod\("([^"]+)",\s*nullptr\)')


def _safe_eval(expr):
    """Evaluate a C arithmetic literal expression such as 200.0 / (2. * (1. + 0.50))."""
    expr = expr.replace("UL", "").strip()
    ops = {ast.Add: operator.add, ast.Sub: operator.sub,
           ast.Mult: operator.mul, ast.Div: operator.truediv}

    def ev(node):
        if isinstance(node, ast.Expression):
            return ev(node.body)
        if isinstance(node, ast.Constant):
            return float(node.value)
        if isinstance(node, ast.BinOp) and type(node.op) in ops:
            return ops[type(node.op)](ev(node.left), ev(node.right))
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub):
            return -ev(node.operand)
        raise ValueError("unsupported literal expression: %r" % expr)

    # C literals like "2." and "0." are valid Python floats already
    return ev(ast.parse(expr, mode="eval"))


def _matching_brace(text, start):
    depth = 0
    for i in range(start, len(text)):
        c = text[i]
        if c == "{":
            depth += 1
        elif c == "}":
            depth -= 1
            if depth == 0:
                return i
    raise ValueError("unbalanced braces")


def _parse_braced(body):
    """Parse a brace-nested list whose leaves are hex/decimal float literals."""
    body = _STRTOD.sub(lambda m: repr(float.fromhex(m.group(1))), body)
    body = body.replace("{", "[").replace("}", "]")
    return ast.literal_eval(body)


def _flatten(x):
    if isinstance(x, list):
        out = []
        for e in x:
            out.extend(_flatten(e))
        return out
    return [float(x)]


def _arrays(text):
    """All array declarations in a chunk of text, in order."""
    res = []
    for m in _DECL.finditer(text):
        kind, name, dims, name2 = m.group(1), m.group(2), m.group(3), m.group(4)
        assert name == name2, (name, name2)
        b0 = text.index("{", m.end() - 1)
        b1 = _matching_brace(text, b0)
        nested = _parse_braced(text[b0:b1 + 1])
        shape = [int(d.strip().rstrip("UL")) for d in dims.split(",")]
        flat = _flatten(nested)
        n = 1
        for s in shape:
            n *= s
        assert len(flat) == n, (name, shape, len(flat))
        res.append((name, kind.lower(), shape, flat))
    return res


def _rod_index(var):
    return int(re.search(r"(\d+)$", var).group(1))


def _split_var(name):
    """ROD001_Position_REPLACED -> (1, 'Position', 'REPLACED')."""
    m = re.match(r"ROD(\d+)_(\w+)_(REPLACED|PYEL)$", name)
    return int(m.group(1)), m.group(2), m.group(3)


def _init_params(text):
    p = {}

    def grab(tag):
        m = re.search(r"Init::%s\{([^}]*)\}" % tag, text)
        return m.group(1) if m else None

    p["n_elem"] = int(grab("NElement").rstrip("UL"))
    p["radius"] = _safe_eval(grab("Radius"))
    p["length"] = _safe_eval(grab("Length"))
    for vec in ("Origin", "Direction", "Normal"):
        m = re.search(r"Init::%s\{::elastica::Vec3\{([^}]*)\}\}" % vec, text)
        p[vec.lower()] = [_safe_eval(v) for v in m.group(1).split(",")]
    fd, td = grab("ForceDampingRate"), grab("TorqueDampingRate")
    p["force_damping"] = _safe_eval(fd) if fd else None
    p["torque_damping"] = _safe_eval(td) if td else None
    m = re.search(r"create_material\(([^;]*)\);", text, re.S)
    args = [_safe_eval(a) for a in m.group(1).split(",")]
    p["density"], p["youngs_modulus"], p["shear_modulus"] = args[0], args[2], args[3]
    return p


def extract_file(fname):
    text = open(os.path.join(REF, fname)).read()
    plugin = re.search(r"using Plugin = elastica::cosserat_rod::(\w+);", text).group(1)
    starts = [(m.start(), m.group(1)) for m in _CASE.finditer(text)]
    cases = []
    for ci, (pos, cname) in enumerate(starts):
        end = starts[ci + 1][0] if ci + 1 < len(starts) else len(text)
        body = text[pos:end]
        if "REPLACED" not in body:
            continue  # e.g. SimpleInitialization: no golden data
        case = {"name": cname, "plugin": plugin, "init": _init_params(body)}
        case["separate_blocks"] = "RestrictSizeAcrossBlockTypes" in body
        m = re.search(r"TimeDelta time_delta\{([^}]*)\}", body)
        case["dt"] = _safe_eval(m.group(1)) if m else None
        rods = sorted({_rod_index(v) for v in re.findall(r"slicedrod_(\d+)", body)})
        case["n_rods"] = len(rods)
        sec_pos = [(m.start(), m.group(1)) for m in _SECTION.finditer(body)]
        setup_end = sec_pos[0][0] if sec_pos else len(body)
        setup = body[:setup_end]
        replaced = {}
        for name, kind, shape, flat in _arrays(setup):
            rod, var, what = _split_var(name)
            if what != "REPLACED":
                continue
            replaced.setdefault(str(rod), {})[var] = {"kind": kind, "shape": shape, "data": flat}
        case["replaced"] = replaced
        assigns = []
        for m in _ASSIGN.finditer(setup):
            tag, rodvar, shrink, src = m.group(1), m.group(2), bool(m.group(3)), m.group(4)
            rod, var, what = _split_var(src)
            assert what == "REPLACED" and var == tag, (tag, src)
            assigns.append({"rod": _rod_index(rodvar), "tag": tag, "shrink_diagonal": shrink})
        case["assign"] = assigns
        sections = []
        if sec_pos:
            for si, (spos, sname) in enumerate(sec_pos):
                send = sec_pos[si + 1][0] if si + 1 < len(sec_pos) else len(body)
                sb = body[spos:send]
                eps = float(re.search(r"approx\.epsilon\(([^)]*)\)", sb).group(1))
                st = re.search(r"std::size_t step = (\d+);", sb)
                sections.append({"name": sname, "epsilon": eps,
                                 "steps": int(st.group(1)) if st else 0,
                                 "expected": _expected(sb)})
        else:
            # single evaluation test: compute_internal_forces/torques + update_acceleration
            eps = float(re.search(r"epsilon\(([^)]*)\)", body).group(1))
            calls = re.findall(r"^\s*(compute_internal_forces|compute_internal_torques|update_acceleration)\(",
                               body, re.M)
            sections.append({"name": "Evaluate", "epsilon": eps, "steps": 0,
                             "calls": calls, "expected": _expected(body)})
        case["sections"] = sections
        cases.append(case)
    return {"source": "tests/Unit/Systems/CosseratRods/FullRodTest/" + fname, "cases": cases}


def _expected(text):
    exp = {}
    for name, kind, shape, flat in _arrays(text):
        rod, var, what = _split_var(name)
        if what != "PYEL":
            continue
        exp.setdefault(str(rod), {})[var] = {"kind": kind, "shape": shape, "data": flat}
    return exp


def main():
    if not os.path.isdir(REF):
        print("reference sources not present; fixtures are already committed", file=sys.stderr)
        return 0
    total = 0
    for f in FILES:
        d = extract_file(f)
        n = sum(len(v["data"]) for c in d["cases"] for r in c["replaced"].values() for v in r.values())
        n += sum(len(v["data"]) for c in d["cases"] for s in c["sections"]
                 for r in s["expected"].values() for v in r.values())
        total += n
        out = os.path.join(OUT, f.replace(".cpp", ".json").replace("Test_", "fullrod_"))
        with open(out, "w") as fh:
            json.dump(d, fh, separators=(",", ":"))
        print("%s: %d cases, %d values -> %s" % (f, len(d["cases"]), n, os.path.basename(out)))
    print("total values:", total)
    return 0


if __name__ == "__main__":
    sys.exit(main())
