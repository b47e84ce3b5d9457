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
