"""The C ABI without a GPU: the library loads, exports every function that
include/fvk.h declares, and the ctypes mirrors in finitevolume_amd/abi.py
have the header's struct layouts (sizes and field offsets, from a C program
compiled against the header).  Only host-side functions are called."""
import ctypes as C
import os
import re
import shutil
import subprocess

import pytest

from finitevolume_amd import abi
from finitevolume_amd.build import LIB

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
HEADER = os.path.join(REPO, "include", "fvk.h")

STRUCTS = {
    "fvk_mesh_desc": abi.MeshDescC,
    "fvk_bc": abi.BCC,
    "fvk_sa_constants": abi.SAConstantsC,
    "fvk_sst_constants": abi.SSTConstantsC,
    "fvk_params": abi.ParamsC,
    "fvk_initial_condition": abi.InitialConditionC,
    "fvk_launch_shape": abi.LaunchShapeC,
}


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(fvk_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.skip("libfvk.so is not built (run __graft_entry__.build())")
    return C.CDLL(LIB)


def test_every_declared_function_is_exported(lib):
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_header(tmp_path):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "fvk.h"', "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I" + os.path.join(REPO, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for line in filter(None, out):
        cname, field, value = line.split()
        py = STRUCTS[cname]
        got = C.sizeof(py) if field == "size" else getattr(py, field).offset
        assert got == int(value), (cname, field, got, value)


def test_host_defaults_match_python(lib):
    lib.fvk_params_default.argtypes = [C.POINTER(abi.ParamsC)]
    p = abi.ParamsC()
    lib.fvk_params_default(C.byref(p))
    q = abi.default_params()
    for fname, _ in abi.ParamsC._fields_:
        a, b = getattr(p, fname), getattr(q, fname)
        if isinstance(a, (int, float)):
            assert a == b, fname
    lib.fvk_num_vars.argtypes = [C.c_int32]
    assert [lib.fvk_num_vars(e) for e in range(5)] == [4, 4, 5, 6, 1]
