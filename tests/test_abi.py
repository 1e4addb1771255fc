"""The C-ABI libraries load and export every symbol include/*.h declares (no device calls)."""
import ctypes
import os
import re

import pytest

from hydro_rs_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hfv_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,path", [("hfv.h", _lib.HFV_SO), ("hfv_host.h", _lib.HOST_SO)])
def test_library_exports_declared_symbols(header, path):
    assert os.path.exists(path), f"{path} missing: run __graft_entry__.build()"
    lib = ctypes.CDLL(path)
    names = declared(header)
    assert names, header
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_header():
    assert set(declared("hfv.h")) == set(_lib.HFV_SYMBOLS)
    assert set(declared("hfv_host.h")) == set(_lib.HOST_SYMBOLS)


def test_integration_guide_declares_every_entry_point():
    """The integration guide's Rust extern blocks declare exactly the entry points of include/hfv.h."""
    guide = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    rust = set(re.findall(r"pub fn (hfv_[a-z0-9_]+)\(", guide))
    assert rust == set(declared("hfv.h"))


def test_bindings_load():
    assert _lib.hfv() is not None
    assert _lib.host() is not None


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors of the POD structs have the C compiler's sizes and field offsets."""
    import subprocess

    import ctypes as C

    structs = {"hfv_params": _lib.Params, "hfv_timestep_info": _lib.TimestepInfo,
               "hfv_mesh_arrays": _lib.MeshArrays, "hfv_state": _lib.State, "hfv_engine_info": _lib.EngineInfo,
               "hfv_stats": _lib.Stats, "hfv_tune_class": _lib.TuneClass}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "hfv.h"', "int main(void){"]
    for cname, py in structs.items():
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    got = [int(x) for x in subprocess.check_output([str(exe)]).split()]
    want = []
    for py in structs.values():
        want.append(C.sizeof(py))
        want.extend(getattr(py, f).offset for f, _ in py._fields_)
    assert got == want
