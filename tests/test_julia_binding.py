"""Static checks of the Julia binding (quantumopticsbase.jl_amd/julia/QobHIP.jl),
which cannot run here (no Julia toolchain).

Julia dispatches on the most specific method.  A binding method whose
operator argument is a Union, or whose bases are not tied the way the
reference ties them, can be neither more nor less specific than the
reference's method for the same call, and the call then fails with an
ambiguity MethodError.  These tests parse every `mul!`, `expect` and
`variance` method of the binding and check that:
  * no operator argument is a Union (or a Union alias);
  * the state and result arguments are narrowed to device storage;
  * the basis type parameters are tied as in the reference method;
  * there is a method for every operator type and state form of the
    reference's `mul!` (src/operators_lazysum.jl:189,202,215,227;
    src/operators_lazytensor.jl:539,559,576,593;
    src/operators_sparse.jl:199-202; src/time_dependent_operator.jl:274-277;
    src/operators_lazyproduct.jl:103,117,131,148; src/operators_dense.jl:394-396);
  * INTEGRATION.md lists each binding method beside the reference method.
"""
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
JL = ROOT / "quantumopticsbase.jl_amd" / "julia" / "QobHIP.jl"
INTEG = ROOT / "INTEGRATION.md"

OPS = ["LazySum", "LazyTensor", "TimeDependentSum", "LazyProduct", "SparseOpPureType", "SparseOpAdjType"]


def _split_args(s):
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "{(":
            depth += 1
        elif ch in "})":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def methods(name):
    """(args, where) of every top-level method definition of `name`."""
    text = JL.read_text()
    text = re.sub(r"#[^\n]*", "", text)
    found = []
    for m in re.finditer(r"(?m)^(?:function\s+)?" + re.escape(name) + r"\(", text):
        i = m.end()
        depth = 1
        j = i
        while depth:
            if text[j] == "(":
                depth += 1
            elif text[j] == ")":
                depth -= 1
            j += 1
        args = _split_args(text[i:j - 1])
        w = re.match(r"\s*where\s*\{", text[j:])
        where = ""
        if w:
            k = j + w.end()
            depth = 1
            e = k
            while depth:
                if text[e] == "{":
                    depth += 1
                elif text[e] == "}":
                    depth -= 1
                e += 1
            where = text[k:e - 1]
        found.append((args, where))
    return found


def _type(arg):
    return arg.split("::", 1)[1].strip() if "::" in arg else ""


def _head(t):
    return t.split("{", 1)[0]


def _params(t):
    return _split_args(t[t.index("{") + 1:t.rindex("}")]) if "{" in t else []


def _form(args):
    """('ket'|'bra'|'left'|'right'|'dense', operator type) of a mul! method."""
    r, a, b = (_type(x) for x in args[:3])
    if _head(r) == "Ket":
        return "ket", a
    if _head(r) == "Bra":
        return "bra", b
    if _head(a) == "Operator" and _head(b) == "Operator":
        return "dense", a
    if _head(a) == "Operator":
        return "right", b
    return "left", a


def test_no_union_typed_operator_argument():
    text = JL.read_text()
    assert "DevOp" not in text
    for name in ("mul!", "expect", "variance"):
        for args, _ in methods(name):
            for a in args:
                t = _type(a)
                assert "Union" not in t, (name, args)
                assert _head(t) not in ("SparseOpType", "DenseOpType", "AbstractOperator"), (name, args)


def test_states_are_device_arrays_and_bases_tied():
    ms = methods("mul!")
    assert len(ms) == 27  # 4 forms x 6 operator types + 3 dense forms
    for args, where in ms:
        assert len(args) == 5, args
        r, a, b = (_type(x) for x in args[:3])
        form, op = _form(args)
        states = {"ket": (r, b), "bra": (r, a), "left": (r, b), "right": (r, a), "dense": (r, a, b)}[form]
        for t in states:
            assert "DevVec" in t or "DevMat" in t, args
        p = {k: _params(_type(x)) for k, x in zip("rab", args[:3])}
        # the reference's basis ties: result/op/state share B1, B2, B3 as in
        # mul!(result::Ket{B1}, a::Op{B1,B2}, b::Ket{B2}) and the Operator forms
        if form == "ket":
            assert p["r"][0] == p["a"][0] and p["a"][1] == p["b"][0], args
        elif form == "bra":
            assert p["a"][0] == p["b"][0] and p["b"][1] == p["r"][0], args
        elif form == "left":
            assert p["r"][0] == p["a"][0] and p["a"][1] == p["b"][0] and p["r"][1] == p["b"][1], args
        elif form == "right":
            assert p["r"][0] == p["a"][0] and p["a"][1] == p["b"][0] and p["r"][1] == p["b"][1], args
        else:
            assert p["r"][0] == p["a"][0] and p["a"][1] == p["b"][0] and p["r"][1] == p["b"][1], args
        for bp in set(x for v in p.values() for x in v[:2] if re.fullmatch(r"B\d", x)):
            assert re.search(r"\b" + bp + r"\b", where), (bp, where, args)


def test_every_reference_form_has_a_method():
    have = {}
    for args, _ in methods("mul!"):
        form, op = _form(args)
        have.setdefault((form, _head(op)), 0)
        have[(form, _head(op))] += 1
    for op in OPS:
        for form in ("ket", "bra", "left", "right"):
            assert have.get((form, op)) == 1, (form, op, have)
    assert have.get(("dense", "Operator")) == 1
    assert have.get(("ket", "Operator")) == 1 and have.get(("bra", "Operator")) == 1
    # LazyTensor methods carry the reference's constraint on the factor tuple
    for args, where in methods("mul!"):
        form, op = _form(args)
        if _head(op) == "LazyTensor":
            assert "Tuple{Vararg{DataOperator}}" in where, where


def test_expect_and_variance_per_operator_type():
    for name in ("expect", "variance"):
        have = set()
        for args, _ in methods(name):
            op, st = _type(args[0]), _type(args[1])
            assert "DevVec" in st or "DevMat" in st, args
            have.add((_head(op), _head(st)))
        for op in OPS:
            assert (op, "Ket") in have and (op, "Operator") in have, (name, op, have)


@pytest.mark.parametrize("ref", [
    "src/operators_lazysum.jl:189", "src/operators_lazysum.jl:202", "src/operators_lazysum.jl:215",
    "src/operators_lazysum.jl:227", "src/operators_lazytensor.jl:539", "src/operators_lazytensor.jl:559",
    "src/operators_lazytensor.jl:576", "src/operators_lazytensor.jl:593", "src/operators_sparse.jl:199",
    "src/operators_sparse.jl:200", "src/operators_sparse.jl:201", "src/operators_sparse.jl:202",
    "src/time_dependent_operator.jl:274", "src/time_dependent_operator.jl:275",
    "src/time_dependent_operator.jl:276", "src/time_dependent_operator.jl:277",
    "src/operators_lazyproduct.jl:103", "src/operators_lazyproduct.jl:117", "src/operators_lazyproduct.jl:131",
    "src/operators_lazyproduct.jl:148", "src/operators_dense.jl:394", "src/operators_dense.jl:395",
    "src/operators_dense.jl:396"])
def test_integration_lists_each_specialised_method(ref):
    assert ref in INTEG.read_text(), ref
