"""The Julia binding (networkdynamics.jl_amd/julia/NDHIP.jl) cannot run here: no Julia.
These CPU tests read its `ccall`s and its mirrored structs as text and check them
against the C declarations in include/nd_hip.h, so that a changed prototype or struct
in the header fails a test instead of a Julia user's first call: every called symbol
exists in the header and in the built library, the return type and every argument type
map to the C type at that position, and NDKind / NDStat have the fields of nd_kind /
nd_ensemble_stat in the same order and widths."""
import os
import re

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
JL = os.path.join(ROOT, "networkdynamics.jl_amd", "julia", "NDHIP.jl")
HDR = os.path.join(ROOT, "include", "nd_hip.h")


def _strip_c_comments(s):
    return re.sub(r"/\*.*?\*/", " ", s, flags=re.S)


def _header_prototypes():
    """name -> (return type, [argument types]) for every function in the header."""
    s = _strip_c_comments(open(HDR).read())
    s = re.sub(r"#.*", " ", s)
    out = {}
    for m in re.finditer(r"([A-Za-z_][\w \*]*?)\b(nd_\w+)\s*\(([^;{}]*?)\)\s*;", s):
        ret, name, args = " ".join(m.group(1).split()), m.group(2), " ".join(m.group(3).split())
        if ret.startswith("typedef") or "(" in ret:
            continue
        types = []
        if args and args != "void":
            for a in args.split(","):
                a = a.strip()
                # drop the parameter name: the last identifier not part of the type
                t = re.sub(r"\b[A-Za-z_]\w*$", "", a).strip() if not a.endswith("*") else a
                types.append(" ".join(t.replace("*", " * ").split()))
        out[name] = (" ".join(ret.replace("*", " * ").split()), types)
    return out


def _c_to_julia(t):
    """the Julia ccall type a C type maps to"""
    t = t.replace("const ", "").strip()
    base = {"int": "Cint", "void": "Cvoid", "int64_t": "Int64", "int32_t": "Int32", "double": "Float64",
            "size_t": "Csize_t", "uint64_t": "UInt64", "float": "Float32"}
    stars = t.count("*")
    b = t.replace("*", "").strip()
    if b == "char" and stars == 1:
        return "Cstring"
    if stars == 0:
        return base[b]
    inner = {"nd_kind": "NDKind", "nd_ensemble_stat": "NDStat"}.get(b)
    if inner is None:
        inner = base.get(b, "Cvoid")         # opaque handles (nd_plan, nd_network, ...) and void
    out = inner
    for _ in range(stars):
        out = f"Ptr{{{out}}}"
    return out


def _julia_ccalls():
    """(name, return type, [argument types]) of every ccall in the binding"""
    s = open(JL).read()
    out = []
    for m in re.finditer(r"ccall\(\(:(\w+),\s*libnd\),\s*", s):
        i = m.end()
        # return type: up to the next top-level comma
        j = s.index(",", i)
        ret = s[i:j].strip()
        k = j + 1
        while s[k].isspace():
            k += 1
        assert s[k] == "(", (m.group(1), s[k:k + 40])
        depth, e = 0, k
        while True:
            if s[e] in "({":
                depth += 1
            elif s[e] in ")}":
                depth -= 1
                if depth == 0:
                    break
            e += 1
        body = s[k + 1:e]
        args, depth, cur = [], 0, ""
        for ch in body:
            if ch in "({":
                depth += 1
            elif ch in ")}":
                depth -= 1
            if ch == "," and depth == 0:
                args.append(cur.strip())
                cur = ""
            else:
                cur += ch
        if cur.strip():
            args.append(cur.strip())
        out.append((m.group(1), ret, args))
    return out


def test_every_ccall_matches_the_header():
    protos = _header_prototypes()
    calls = _julia_ccalls()
    assert len(calls) >= 15
    for name, ret, args in calls:
        assert name in protos, f"{name} is not declared in include/nd_hip.h"
        cret, cargs = protos[name]
        assert _c_to_julia(cret) == ret, (name, cret, ret)
        assert len(cargs) == len(args), (name, cargs, args)
        for pos, (ct, jt) in enumerate(zip(cargs, args)):
            assert _c_to_julia(ct) == jt, (name, pos, ct, jt)


def test_called_symbols_are_exported():
    import ctypes
    lib = os.path.join(ROOT, "networkdynamics.jl_amd", "libnd_hip.so")
    if not os.path.exists(lib):
        pytest.skip("libnd_hip.so not built")
    L = ctypes.CDLL(lib)
    for name, _, _ in _julia_ccalls():
        assert hasattr(L, name), name


def _c_struct_fields(name):
    s = _strip_c_comments(open(HDR).read())
    m = re.search(r"typedef struct \{([^{}]*)\}\s*" + name + r"\s*;", s)
    assert m, name
    fields = []
    for decl in m.group(1).split(";"):
        decl = " ".join(decl.split())
        if not decl:
            continue
        t, names = decl.split(" ", 1)
        for n in names.split(","):
            n = n.strip()
            arr = re.match(r"(\w+)\[(\w+)\]", n)
            fields.append((arr.group(1), t, arr.group(2)) if arr else (n, t, None))
    return fields


def _julia_struct_fields(name):
    s = open(JL).read()
    m = re.search(r"struct " + name + r"\n(.*?)\nend", s, flags=re.S)
    assert m, name
    out = []
    for line in m.group(1).splitlines():
        line = line.split("#")[0].strip()
        if line:
            n, t = line.split("::")
            out.append((n.strip(), t.strip()))
    return out


@pytest.mark.parametrize("cname,jname", [("nd_kind", "NDKind"), ("nd_ensemble_stat", "NDStat")])
def test_mirrored_structs_match(cname, jname):
    hdr = open(HDR).read()
    max_out = re.search(r"#define ND_MAX_OUT (\d+)", hdr).group(1)
    cf, jf = _c_struct_fields(cname), _julia_struct_fields(jname)
    assert [n for n, _, _ in cf] == [n for n, _ in jf], (cf, jf)
    for (n, ct, arr), (_, jt) in zip(cf, jf):
        want = _c_to_julia(ct)
        if arr is not None:
            want = f"NTuple{{{max_out if arr == 'ND_MAX_OUT' else arr},{want}}}"
        assert jt == want, (cname, n, ct, jt)
