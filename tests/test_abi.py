"""The C ABI without a GPU: the library loads, exports every function include/gomgpu.h
declares, and the ctypes mirror has the C structs' layout (checked against gcc)."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from gom_release_amd import abi, host

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in sorted(os.listdir(os.path.join(ROOT, "include")))
           if h.endswith(".h")]


def declared_functions():
    names = []
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names += re.findall(r"^[A-Za-z_][\w \*]*?\b(gomgpu_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for need in ("gomgpu_create", "gomgpu_set_state", "gomgpu_set_forcing", "gomgpu_stage",
                 "gomgpu_step", "gomgpu_get_state", "gomgpu_destroy"):
        assert need in names


def test_library_exports_every_declared_function():
    if not os.path.exists(host.LIB_PATH):
        pytest.skip("libgomgpu.so not built")
    L = C.CDLL(host.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(host.EXPORTED) == sorted(declared_functions())


def test_no_torch_or_cxx_types_in_signatures():
    for h in HEADERS:
        text = open(h).read()
        assert "torch" not in text and "at::" not in text and "std::" not in text


def _c_layout():
    """sizeof and offsetof of every struct field, from gcc on the header."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "gomgpu.h"', "int main(void) {"]
    for st, fields in (("gomgpu_mesh", [n for n, _ in abi.MESH_REAL + abi.MESH_INT]),
                       ("gomgpu_state", [n for n, _ in abi.STATE_REAL + abi.STATE_INT]),
                       ("gomgpu_params", abi.PARAMS_INT + abi.PARAMS_REAL),
                       ("gomgpu_forcing", abi.FORCING_SCALARS + [n for n, _ in abi.FORCING_REAL]),
                       ("gomgpu_boundary", [n for n, _ in abi.Boundary._fields_]),
                       ("gomgpu_holland", [n for n, _ in abi.Holland._fields_])):
        lines.append(f'printf("{st} size %zu\\n", sizeof({st}));')
        for f in fields:
            lines.append(f'printf("{st} {f} %zu\\n", offsetof({st}, {f}));')
    lines.append("return 0; }")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(src, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    lay = {}
    for ln in out.splitlines():
        parts = ln.split()
        lay[(parts[0], parts[1])] = int(parts[2])
    return lay


def test_ctypes_structs_match_c_layout():
    lay = _c_layout()
    for st, cls in (("gomgpu_mesh", abi.Mesh), ("gomgpu_state", abi.State),
                    ("gomgpu_params", abi.Params), ("gomgpu_forcing", abi.Forcing),
                    ("gomgpu_boundary", abi.Boundary), ("gomgpu_holland", abi.Holland)):
        assert C.sizeof(cls) == lay[(st, "size")], st
        for name, _ in cls._fields_:
            assert getattr(cls, name).offset == lay[(st, name)], (st, name)


def test_create_without_gpu_fails_cleanly():
    """With no GPU visible the product path reports an error instead of computing on the CPU."""
    if not os.path.exists(host.LIB_PATH):
        pytest.skip("libgomgpu.so not built")
    if host.device_count() > 0:
        pytest.skip("a GPU is visible")
    p = abi.Params()
    p.maxnod, p.maxele, p.maxface, p.maxlayer = 4, 1, 4, 1
    m = abi.mesh_arrays(p)
    ctx = C.c_void_p()
    rc = host.lib().gomgpu_create(C.byref(m.struct()), C.byref(p), C.byref(ctx))
    assert rc != 0
