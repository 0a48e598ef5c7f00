"""Generate the Fortran ISO_C_BINDING interface module of the C ABI (include/gomgpu.h).

The module is what a GOM maintainer compiles next to the reference sources so that the
Fortran time loop (GOM_Rv1.1.0/source/main.f90:53-304) can call the GPU step: bind(C)
derived types with the same field order as the C structs (generated from the header's
X-macro lists, so they cannot drift) and interfaces of every gomgpu_* function.

    python -m gom_release_amd.fortran_binding > include/gomgpu_iface.f90
"""
from __future__ import annotations

import re
import sys

from . import abi

STAGES = [s.upper() for s in abi.STAGES]


def _enum_values():
    text = open(abi.HEADER).read()
    vals = {}
    for name, val in re.findall(r"\b(GOMGPU_(?:E_|W_|FP64|MIXED)\w*)\s*=\s*(\d+)", text):
        vals[name] = int(val)
    return vals


def _wrap(line: str, indent: str = "  ") -> list:
    """Split a long declaration into free-form continuation lines (<= 100 chars)."""
    out, cur = [], line
    while len(cur) > 100:
        cut = cur.rfind(",", 0, 98) + 1
        out.append(cur[:cut] + " &")
        cur = indent + "  & " + cur[cut:].lstrip()
    out.append(cur)
    return out


def module_text() -> str:
    L = []
    a = L.append
    a("! gomgpu_iface.f90 - ISO_C_BINDING interface of the GPU step (include/gomgpu.h).")
    a("! Generated by `python -m gom_release_amd.fortran_binding`; do not edit by hand.")
    a("! Field order of the bind(C) types is the header's X-macro order.")
    a("module gomgpu_iface")
    a("  use, intrinsic :: iso_c_binding")
    a("  implicit none")
    for i, s in enumerate(STAGES):
        a(f"  integer(c_int), parameter :: GOMGPU_STAGE_{s} = {i}")
    for n, v in sorted(_enum_values().items(), key=lambda t: (t[1], t[0])):
        a(f"  integer(c_int), parameter :: {n} = {v}")
    a("")
    for tname, fields in (("gomgpu_mesh", [n for n, _ in abi.MESH_REAL + abi.MESH_INT]),
                          ("gomgpu_state", [n for n, _ in abi.STATE_REAL + abi.STATE_INT])):
        a(f"  type, bind(C) :: {tname}")
        L.extend(_wrap("    type(c_ptr) :: " + ", ".join(fields) + " = c_null_ptr", "    "))
        a(f"  end type {tname}")
    a("  type, bind(C) :: gomgpu_params")
    L.extend(_wrap("    integer(c_int32_t) :: " + ", ".join(abi.PARAMS_INT) + " = 0", "    "))
    L.extend(_wrap("    real(c_double) :: " + ", ".join(abi.PARAMS_REAL) + " = 0.0d0", "    "))
    a("  end type gomgpu_params")
    a("  type, bind(C) :: gomgpu_forcing")
    L.extend(_wrap("    real(c_double) :: " + ", ".join(abi.FORCING_SCALARS) + " = 0.0d0", "    "))
    L.extend(_wrap("    type(c_ptr) :: " + ", ".join(n for n, _ in abi.FORCING_REAL) + " = c_null_ptr", "    "))
    a("  end type gomgpu_forcing")
    a("")
    a("  interface")
    sigs = [
        ("integer(c_int) function gomgpu_create(mesh, params, ctx) bind(C, name='gomgpu_create')",
         ["import :: c_int, c_ptr, gomgpu_mesh, gomgpu_params",
          "type(gomgpu_mesh), intent(in) :: mesh", "type(gomgpu_params), intent(in) :: params",
          "type(c_ptr), intent(out) :: ctx"]),
        ("integer(c_int) function gomgpu_set_state(ctx, state) bind(C, name='gomgpu_set_state')",
         ["import :: c_int, c_ptr, gomgpu_state", "type(c_ptr), value :: ctx",
          "type(gomgpu_state), intent(in) :: state"]),
        ("integer(c_int) function gomgpu_get_state(ctx, state) bind(C, name='gomgpu_get_state')",
         ["import :: c_int, c_ptr, gomgpu_state", "type(c_ptr), value :: ctx",
          "type(gomgpu_state), intent(in) :: state"]),
        ("integer(c_int) function gomgpu_set_forcing(ctx, forcing) bind(C, name='gomgpu_set_forcing')",
         ["import :: c_int, c_ptr, gomgpu_forcing", "type(c_ptr), value :: ctx",
          "type(gomgpu_forcing), intent(in) :: forcing"]),
        ("integer(c_int) function gomgpu_stage(ctx, stage_id) bind(C, name='gomgpu_stage')",
         ["import :: c_int, c_ptr", "type(c_ptr), value :: ctx", "integer(c_int), value :: stage_id"]),
        ("integer(c_int) function gomgpu_step(ctx) bind(C, name='gomgpu_step')",
         ["import :: c_int, c_ptr", "type(c_ptr), value :: ctx"]),
        ("integer(c_int) function gomgpu_run(ctx, nsteps) bind(C, name='gomgpu_run')",
         ["import :: c_int, c_ptr", "type(c_ptr), value :: ctx", "integer(c_int), value :: nsteps"]),
        ("integer(c_int) function gomgpu_budgets(ctx, volume, salt) bind(C, name='gomgpu_budgets')",
         ["import :: c_int, c_ptr, c_double", "type(c_ptr), value :: ctx",
          "real(c_double), intent(out) :: volume, salt"]),
        ("integer(c_int) function gomgpu_pcg_iterations(ctx) bind(C, name='gomgpu_pcg_iterations')",
         ["import :: c_int, c_ptr", "type(c_ptr), value :: ctx"]),
        ("integer(c_int) function gomgpu_sync(ctx) bind(C, name='gomgpu_sync')",
         ["import :: c_int, c_ptr", "type(c_ptr), value :: ctx"]),
        ("type(c_ptr) function gomgpu_last_error(ctx) bind(C, name='gomgpu_last_error')",
         ["import :: c_ptr", "type(c_ptr), value :: ctx"]),
        ("integer(c_int) function gomgpu_timings(ctx, enable, names, ms, n) bind(C, name='gomgpu_timings')",
         ["import :: c_int, c_ptr, c_double", "type(c_ptr), value :: ctx", "integer(c_int), value :: enable",
          "type(c_ptr), value :: names, ms", "integer(c_int), value :: n"]),
        ("subroutine gomgpu_destroy(ctx) bind(C, name='gomgpu_destroy')",
         ["import :: c_ptr", "type(c_ptr), value :: ctx"]),
    ]
    for head, body in sigs:
        L.extend(_wrap("    " + head, "    "))
        for b in body:
            a("      " + b)
        kind, name = re.search(r"\b(function|subroutine)\s+(\w+)", head).groups()
        a(f"    end {kind} {name}")
    a("  end interface")
    a("end module gomgpu_iface")
    return "\n".join(L) + "\n"


if __name__ == "__main__":
    sys.stdout.write(module_text())
