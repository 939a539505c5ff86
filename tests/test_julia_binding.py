"""The Julia module (scale-track_amd/julia/ScaleTrackHIP.jl) checked against the C-ABI
header without Julia (absent here): every `ccall` names a function the header declares,
with as many arguments as the prototype, each of a Julia type that matches the C type
(Int32 for int32_t, Int64 for int64_t, Float64 for double, a pointer type for a
pointer or array); the return types match; and the Julia structs mirror st_mesh,
st_params and st_tune_result field by field, with the same field types, in order."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "scaletrack.h")
JL = os.path.join(ROOT, "scale-track_amd", "julia", "ScaleTrackHIP.jl")


def _strip_c_comments(s):
    return re.sub(r"/\*.*?\*/", " ", s, flags=re.S)


def header_prototypes():
    text = _strip_c_comments(open(HEADER).read())
    out = {}
    for m in re.finditer(r"(\bint|\bvoid|const\s+char\s*\*)\s*\b(st_\w+)\s*\(([^)]*)\)\s*;", text):
        ret, name, args = m.group(1).strip(), m.group(2), m.group(3).strip()
        params = [] if args in ("", "void") else [a.strip() for a in args.split(",")]
        out[name] = (ret, params)
    return out


def header_structs():
    text = _strip_c_comments(open(HEADER).read())
    out = {}
    for m in re.finditer(r"typedef\s+struct\s+(st_\w+)\s*\{(.*?)\}\s*\1\s*;", text, flags=re.S):
        fields = []
        for decl in m.group(2).split(";"):
            decl = " ".join(decl.split())
            if not decl:
                continue
            mm = re.match(r"(\w+)\s+(\w+)(\[(\d+)\])?$", decl)
            assert mm, decl
            fields.append((mm.group(2), mm.group(1), int(mm.group(4)) if mm.group(4) else 1))
        out[m.group(1)] = fields
    return out


def _split_top(s):
    """Splits at commas outside (), {} and []."""
    parts, depth, cur = [], 0, []
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append("".join(cur).strip())
            cur = []
        else:
            cur.append(ch)
    tail = "".join(cur).strip()
    if tail:
        parts.append(tail)
    return parts


def _balanced(s, i):
    """s[i] == '('; returns the index just past its matching ')'."""
    depth = 0
    for j in range(i, len(s)):
        if s[j] == "(":
            depth += 1
        elif s[j] == ")":
            depth -= 1
            if depth == 0:
                return j + 1
    raise AssertionError("unbalanced parentheses")


def julia_ccalls():
    src = re.sub(r"#.*", "", open(JL).read())
    calls = []
    for m in re.finditer(r"ccall\(\(:(\w+),\s*lib\),", src):
        rest = src[m.end():]
        # return type, then the tuple of argument types
        k = rest.index("(")
        ret = rest[:k].strip().rstrip(",").strip()
        end = _balanced(rest, k)
        types = _split_top(rest[k + 1:end - 1])
        calls.append((m.group(1), ret, types))
    return calls


def julia_structs():
    src = re.sub(r"#.*", "", open(JL).read())
    out = {}
    for m in re.finditer(r"^struct (\w+)\n(.*?)^end", src, flags=re.S | re.M):
        fields = []
        for line in m.group(2).splitlines():
            line = line.strip()
            if line:
                name, typ = line.split("::")
                fields.append((name.strip(), typ.strip()))
        out[m.group(1)] = fields
    return out


SCALAR = {"int32_t": "Int32", "int64_t": "Int64", "double": "Float64"}


def _c_param_type(p):
    """(base type, is pointer) of one C parameter declaration."""
    ptr = "*" in p or "[" in p
    base = re.sub(r"\bconst\b", "", p).replace("*", " ").split("[")[0].split()
    return base[0], ptr


def test_every_ccall_matches_the_header():
    protos = header_prototypes()
    calls = julia_ccalls()
    assert len(calls) >= 40, len(calls)
    for name, ret, types in calls:
        assert name in protos, f"{name} is not declared in scaletrack.h"
        cret, params = protos[name]
        assert len(types) == len(params), f"{name}: {len(types)} Julia argument types, {len(params)} C parameters"
        want_ret = {"int": "Cint", "void": "Cvoid"}.get(cret, "Cstring")
        assert ret == want_ret, f"{name}: returns {ret}, header says {cret}"
        for jt, cp in zip(types, params):
            base, ptr = _c_param_type(cp)
            if ptr:
                assert re.match(r"(Ptr|Ref)\{|Cstring$", jt), f"{name}: C pointer parameter '{cp}' bound as {jt}"
                if base == "char":
                    assert jt == "Cstring", f"{name}: string parameter '{cp}' bound as {jt}"
                elif base in SCALAR and jt.startswith("Ref{") and "Ptr{" not in jt:
                    assert jt == f"Ref{{{SCALAR[base]}}}", f"{name}: '{cp}' bound as {jt}"
            else:
                assert base in SCALAR, f"{name}: unexpected C parameter type '{cp}'"
                assert jt == SCALAR[base], f"{name}: '{cp}' bound as {jt}"


def test_every_entry_point_is_bound():
    bound = {name for name, _, _ in julia_ccalls()}
    missing = sorted(set(header_prototypes()) - bound - {"st_diag_copy_bandwidth", "st_diag_kernel_times",
                                                          "st_batch_info"})
    assert not missing, missing


@pytest.mark.parametrize("c_name,jl_name", [("st_mesh", "StMesh"), ("st_params", "StParams"),
                                            ("st_tune_result", "StTuneResult")])
def test_structs_mirror_the_header(c_name, jl_name):
    c = header_structs()[c_name]
    j = julia_structs()[jl_name]
    assert [n for n, _, _ in c] == [n for n, _ in j]
    for (name, ctype, count), (_, jtype) in zip(c, j):
        want = SCALAR[ctype] if count == 1 else f"NTuple{{{count},{SCALAR[ctype]}}}"
        assert jtype == want, f"{c_name}.{name}: {ctype}[{count}] mirrored as {jtype}"
