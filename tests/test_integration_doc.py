"""The Julia ccall shim of the repository's integration notes against the C
ABI of include/canary_dg.h: every ccall names a declared entry point and passes as
many arguments, of matching width, as the declaration takes (Julia cannot
run here, so this keeps the shim's text in step with the header)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_prototypes():
    txt = open(os.path.join(ROOT, "include", "canary_dg.h")).read()
    txt = re.sub(r"/\*.*?\*/", " ", txt, flags=re.S)
    protos = {}
    for m in re.finditer(r"\b(int|double|void|int64_t|canary_mesh\s*\*)\s+(canary_\w+)\s*\(([^;{]*?)\)\s*;",
                         txt, flags=re.S):
        args = [a.strip() for a in m.group(3).split(",") if a.strip() and a.strip() != "void"]
        protos[m.group(2)] = (m.group(1).replace(" ", ""), args)
    return protos


def split_top(s):
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def matching(s, i):
    depth = 0
    for j in range(i, len(s)):
        if s[j] == "(":
            depth += 1
        elif s[j] == ")":
            depth -= 1
            if depth == 0:
                return j
    raise ValueError("unbalanced")


def shim_ccalls():
    txt = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    txt = re.sub(r"#=.*?=#", "", txt, flags=re.S)
    calls = []
    for m in re.finditer(r"ccall\(\(:(canary_\w+),\s*libcanary\)", txt):
        start = txt.index("(", m.start())
        end = matching(txt, start)
        parts = split_top(txt[start + 1:end])
        ret, types = parts[1], parts[2]
        assert types.startswith("(") and types.endswith(")"), (m.group(1), types)
        tlist = split_top(types[1:-1])
        calls.append((m.group(1), ret, tlist, parts[3:]))
    return calls


C2JULIA = {"int": {"Cint"}, "double": {"Cdouble", "Float64"}, "int64_t": {"Int64"}}


def test_shim_calls_match_the_header():
    protos = header_prototypes()
    calls = shim_ccalls()
    assert len(calls) >= 5
    for name, ret, types, args in calls:
        assert name in protos, name
        cret, cargs = protos[name]
        assert ret in C2JULIA.get(cret, {ret}), (name, ret, cret)
        assert len(types) == len(cargs), (name, len(types), len(cargs))
        assert len(args) == len(types), (name, len(args), len(types))
        for t, a in zip(types, cargs):
            ctype = a.rsplit(" ", 1)[0].replace("const ", "").strip()
            if "*" in a:
                assert t.startswith("Ptr") or t.startswith("Ref"), (name, t, a)
            elif ctype in C2JULIA:
                assert t in C2JULIA[ctype], (name, t, a)


@pytest.mark.parametrize("name", ["canary_stage_fused", "canary_lsrk54", "canary_volumerhs"])
def test_header_parser_sees_the_entry_points(name):
    assert name in header_prototypes()
