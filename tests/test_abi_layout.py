"""Layout of the C structs in include/nd_hip.h against their Python mirrors (ctypes
Structures and numpy record dtypes): a small C program compiled with gcc prints each
struct's sizeof and every mirrored field's offsetof, and the mirrors must agree byte for
byte and name every field the header declares.  A field added to the header and not to
the mirror (or the reverse) would otherwise shift every later field silently."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "nd_hip.h")

# C field names that the Python mirrors spell differently
ALIAS = {("nd_newton_stat", "pad"): "pad_"}


def _mirrors(ND):
    from networkdynamics_jl_amd import _abi, ensemble, fixpoint
    return {
        "nd_kind": _abi.NDKind,
        "nd_network_info": _abi.NDNetworkInfo,
        "nd_batch": _abi.NDBatch,
        "nd_mass": _abi.NDMass,
        "nd_plan_tuning": _abi.NDPlanTuning,
        "nd_ensemble_opts": ensemble._Opts,
        "nd_ensemble_stat": ensemble.STAT_DTYPE,
        "nd_newton_opts": fixpoint._Opts,
        "nd_newton_stat": fixpoint.STAT_DTYPE,
    }


def _py_layout(m):
    """(size, [(name, offset)]) of a ctypes Structure or a numpy record dtype"""
    import ctypes
    if isinstance(m, type) and issubclass(m, ctypes.Structure):
        return ctypes.sizeof(m), [(n, getattr(m, n).offset) for n, _ in m._fields_]
    return m.itemsize, [(n, m.fields[n][1]) for n in m.names]


def _header_fields(name):
    s = re.sub(r"/\*.*?\*/", " ", open(HDR).read(), flags=re.S)
    m = re.search(r"typedef struct \{([^{}]*)\}\s*" + name + r"\s*;", s)
    assert m, name
    names = []
    for decl in m.group(1).split(";"):
        decl = " ".join(decl.replace("*", " * ").split())
        if not decl:
            continue
        for part in decl.split(","):
            ident = re.findall(r"[A-Za-z_]\w*(?=\s*(?:\[[^\]]*\])?\s*$)", part.strip())
            names.append(ident[-1])
    return names


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_struct_layouts_match_the_header(ND, tmp_path):
    mirrors = _mirrors(ND)
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "nd_hip.h"', "int main(void) {"]
    expect = {}
    for cname, m in mirrors.items():
        size, fields = _py_layout(m)
        cfields = [ALIAS.get((cname, n), n) for n, _ in fields]
        assert cfields == _header_fields(cname), (cname, cfields, _header_fields(cname))
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for n in cfields:
            lines.append(f'printf("{cname}.{n} %zu\\n", offsetof({cname}, {n}));')
        expect[cname] = size
        for (n, off), cn in zip(fields, cfields):
            expect[f"{cname}.{cn}"] = off
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                                text=True).stdout.splitlines())
    for k, v in expect.items():
        assert int(got[k]) == v, (k, int(got[k]), v)


def _ctypes_for(ct):
    """the ctypes types a C parameter type may be bound with"""
    import ctypes as C
    t = ct.replace("const ", "").strip()
    if t.count("*") == 0:
        return {"int": {C.c_int}, "int32_t": {C.c_int32}, "int64_t": {C.c_int64}, "double": {C.c_double},
                "float": {C.c_float}, "size_t": {C.c_size_t}, "uint64_t": {C.c_uint64}, "void": {None}}[t]
    if t == "char *":
        return {C.c_char_p}
    return None                                  # any pointer type


def test_ctypes_prototypes_match_the_header(ND):
    """_abi.PROTOTYPES binds every function the header declares and no other; each has
    the header's argument count, its scalar argument and return types exactly, and a
    pointer type wherever the header has a pointer."""
    import ctypes as C
    from networkdynamics_jl_amd import _abi
    from test_julia_binding import _header_prototypes
    protos = _header_prototypes()
    assert set(protos) == set(_abi.PROTOTYPES), set(protos) ^ set(_abi.PROTOTYPES)
    ptr_types = (C.c_void_p, C.c_char_p, C._Pointer)
    for name, (res, args) in _abi.PROTOTYPES.items():
        assert name in protos, name
        cret, cargs = protos[name]
        assert len(args) == len(cargs), (name, args, cargs)
        want = _ctypes_for(cret)
        if want is None:
            assert res in (C.c_void_p, C.c_char_p) or issubclass(res, C._Pointer), (name, res)
        else:
            assert res in want, (name, cret, res)
        for pos, (a, ct) in enumerate(zip(args, cargs)):
            want = _ctypes_for(ct)
            if want is None:
                assert a is C.c_void_p or issubclass(a, ptr_types), (name, pos, ct, a)
            else:
                assert a in want, (name, pos, ct, a)
