"""The C-ABI library: it loads without a GPU, exports every function include/frforge_hip.h
declares, and the Python ctypes mirror of the header's structs has the header's layout."""
import ctypes as C
import os
import re
import shutil
import subprocess

import pytest

from frforge_amd import device as dv

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "frforge_hip.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(frf(?:1d)?_[A-Za-z0-9_]+)\s*\(",
                                 text, flags=re.M)))


def test_header_declares_the_entry_points():
    names = _declared_functions()
    assert "frf_residual" in names and "frf_rk3_step" in names and "frf_run" in names
    assert sorted(dv.EXPORTED) == names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(dv.LIB_PATH):
        pytest.fail(f"{dv.LIB_PATH} is not built (python -m frforge_amd.build)")
    lib = C.CDLL(dv.LIB_PATH)
    missing = [n for n in _declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared in frforge_hip.h but not exported: {missing}"


def test_load_library_fails_loudly_when_missing(tmp_path, monkeypatch):
    monkeypatch.setattr(dv, "_lib", None)
    with pytest.raises(dv.FrfError, match="no CPU fallback"):
        dv.load_library(str(tmp_path / "libfrforge_hip.so"))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_ctypes_struct_layout_matches_header(tmp_path):
    structs = {"frf_problem": dv.FrfProblem, "frf_run_result": dv.FrfRunResult,
               "frf_profile": dv.FrfProfile, "frf1d_problem": dv.Frf1dProblem}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "frforge_hip.h"',
             "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in py._fields_:
            lines.append(f'printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in structs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f in py._fields_:
            assert got[(cname, f[0])] == getattr(py, f[0]).offset, (cname, f[0])
