"""The Fortran side of the boundary: the committed ISO_C_BINDING module is the generated one,
compiles with the image's Fortran compiler, has the C structs' sizes, and links against
libgomgpu.so (a call through it returns the library's argument error without a GPU)."""
from __future__ import annotations

import os
import shutil
import subprocess
import tempfile

import pytest

from gom_release_amd import fortran_binding, host

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
IFACE = os.path.join(ROOT, "include", "gomgpu_iface.f90")
FC = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)

PROGRAM = """
program probe
  use gomgpu_iface
  implicit none
  type(gomgpu_mesh) :: m
  type(gomgpu_state) :: s
  type(gomgpu_params) :: p
  type(gomgpu_forcing) :: f
  type(c_ptr) :: ctx
  integer(c_int) :: rc
  print '(a,1x,i0)', 'gomgpu_mesh', c_sizeof(m)
  print '(a,1x,i0)', 'gomgpu_state', c_sizeof(s)
  print '(a,1x,i0)', 'gomgpu_params', c_sizeof(p)
  print '(a,1x,i0)', 'gomgpu_forcing', c_sizeof(f)
  p%maxnod = 0          ! an empty mesh is rejected before any device call
  rc = gomgpu_create(m, p, ctx)
  print '(a,1x,i0)', 'create_rc', rc
end program probe
"""


def test_committed_module_is_generated():
    assert open(IFACE).read() == fortran_binding.module_text()


@pytest.mark.skipif(FC is None, reason="no Fortran compiler")
def test_module_compiles_links_and_matches_c_sizes():
    if not os.path.exists(host.LIB_PATH):
        pytest.skip("libgomgpu.so not built")
    from test_abi import _c_layout
    lay = _c_layout()
    libdir = os.path.dirname(host.LIB_PATH)
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "probe.f90")
        open(src, "w").write(PROGRAM)
        exe = os.path.join(d, "probe")
        subprocess.run([FC, "-J", d, IFACE, src, "-o", exe, f"-L{libdir}", "-lgomgpu",
                        f"-Wl,-rpath,{libdir}"], check=True, cwd=d, capture_output=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=60).stdout
    got = dict(ln.split() for ln in out.splitlines() if ln.strip())
    for st in ("gomgpu_mesh", "gomgpu_state", "gomgpu_params", "gomgpu_forcing"):
        assert int(got[st]) == lay[(st, "size")], st
    assert int(got["create_rc"]) == 1  # GOMGPU_E_ARG
