"""Generate the Fortran ISO_C_BINDING interface module of the C ABI (include/gomgpu.h).

The module is what a GOM maintainer compiles next to the reference sources so that the
Fortran time loop (GOM_Rv1.1.0/source/main.f90:53-304) can call the GPU step: bind(C)
derived types with the same field order as the C structs (generated from the header's
X-macro lists, so they cannot drift) and interfaces of every gomgpu_* function.

    python -m gom_release_amd.fortran_binding > include/gomgpu_iface.f90
    python -m gom_release_amd.fortran_binding bridge > gom_release_amd/fortran/gomgpu_bridge.f90
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


# ---------------------------------------------------------------------------------------
# The bridge module: binds GOM's module arrays (mod_global_variables) to one GPU context.

# gomgpu_params fields that are not module scalars of the same name; set by hand below
_PARAMS_SPECIAL = {"tran_id1", "tran_id2", "precision", "river_on_device", "turbulence_flag",
                   "turbulence_model", "stability_function", "hurricane_flag", "hurricane_read_interval",
                   "diffmax", "xlmin1", "xlmin2", "maxtran2", "trans_sub_iter"}
# arrays with lower bound 0 in the element dimension, passed from element 1
_FROM_ONE = {"top_layer_at_element", "bottom_layer_at_element"}
# mesh fields without a module array of that name (filled from bridge arrays)
_MESH_SPECIAL = {"Qb_cell_id"}
# forcing arrays the host computes in Fortran every step (main.f90:84-161 and the heat model)
_FORCING_HOST = ["Gamma_T", "wind_stress_normal", "wind_stress_tangnt", "air_p1", "eta_at_ob_old",
                 "eta_at_ob_new", "phi_sz", "Q_add"]
_FORCING_GUARD = {"eta_at_ob_old": "num_ob_cell > 0", "eta_at_ob_new": "num_ob_cell > 0",
                  "Q_add": "num_Qb_cell > 0"}
# module names the hand-written part of the bridge uses
_BRIDGE_NAMES = ["dp", "it", "ndt", "dt", "julian_day", "elapsed_time", "spinup_function_tide",
                 "spinup_function_baroclinic", "spinup_function_wind", "tran_id", "maxtran2", "trans_sub_iter",
                 "i_turbulence_model_name", "stability_function", "transport_solver", "transport_flag",
                 "is_temp", "heat_option", "num_ob_cell", "num_Qb_cell", "ob_cell_id", "Q_boundary",
                 "bottom_layer_at_element", "salt_ser_id", "temp_ser_id", "Q_salt_ser_id", "Q_temp_ser_id",
                 "salt_ser_data_num", "salt_ser_time", "salt_ser_salt", "temp_ser_data_num", "temp_ser_time",
                 "temp_ser_temp", "reference_diff_days", "salt_at_obck", "temp_at_obck", "salt_at_Qbc",
                 "temp_at_Qbc"]


def _addr(name: str, kind: str) -> str:
    fn = "gg_i" if kind == "i" else "gg_r"
    return f"{fn}({name}(1))" if name in _FROM_ONE else f"{fn}({name})"


def bridge_text() -> str:
    """gomgpu_bridge.f90: one context bound to GOM's module arrays. The pointer tables are
    generated from the header's field lists; the shims in gomgpu_shims.f90 call it."""
    mesh = [(n, "r") for n, _ in abi.MESH_REAL] + [(n, "i") for n, _ in abi.MESH_INT]
    state = [(n, "r") for n, _ in abi.STATE_REAL] + [(n, "i") for n, _ in abi.STATE_INT]
    pint = [n for n in abi.PARAMS_INT if n not in _PARAMS_SPECIAL]
    preal = [n for n in abi.PARAMS_REAL if n not in _PARAMS_SPECIAL]
    used = sorted(set(pint + preal + [n for n, _ in mesh if n not in _MESH_SPECIAL] + [n for n, _ in state]
                      + _FORCING_HOST + _BRIDGE_NAMES), key=str.lower)
    L = []
    a = L.append
    a("! gomgpu_bridge.f90 - GOM's module arrays bound to one GPU context (include/gomgpu.h).")
    a("! Generated by `python -m gom_release_amd.fortran_binding bridge`; do not edit by hand.")
    a("! Compiled with the reference's modules and linked with gomgpu_iface.o, gomgpu_shims.o and")
    a("! libgomgpu.so in place of the hot-path objects (the shims keep the reference's names).")
    a("!")
    a("! Addresses of the module arrays are taken through assumed-size target dummies (gg_r,")
    a("! gg_i): a contiguous allocatable is passed by address, so the context reads and writes the")
    a("! module arrays themselves and mod_global_variables needs no TARGET attribute. Arrays with")
    a("! lower bound 0 in their layer dimension are passed whole (their first element is level 0);")
    a("! top/bottom_layer_at_element(0:maxele) from element 1. Unallocated arrays pass NULL.")
    a("module gomgpu_bridge")
    a("  use, intrinsic :: iso_c_binding")
    a("  use gomgpu_iface")
    L.extend(_wrap("  use mod_global_variables, only: " + ", ".join(used), "  "))
    a("  use mod_file_definition, only: pw_run_log")
    a("  implicit none")
    a("  private")
    a("  public :: gg_step_begin, gg_stage, gg_startup_stage, gg_after_nodevel, gg_tvd_boundary, &")
    a("    & gg_elm_boundary, gg_step_end, gg_ctx")
    a("  type(c_ptr), save :: gg_ctx = c_null_ptr")
    a("  integer, save :: gg_step = -1          ! the step whose forcing is on the device")
    a("  logical, save :: gg_loaded = .false.   ! module state uploaded since the time loop began")
    a("  integer, save :: gg_sync_every = 1     ! download the state every this many steps")
    a("  integer(c_int32_t), save, allocatable, target :: gg_qb_cells(:)")
    a("  real(c_double), save, allocatable, target :: gg_con_ob(:), gg_con_qb(:)")
    a("contains")
    a("")
    a("  function gg_r(a) result(p)")
    a("    real(c_double), intent(in), target :: a(*)")
    a("    type(c_ptr) :: p")
    a("    p = c_loc(a(1))")
    a("  end function gg_r")
    a("")
    a("  function gg_i(a) result(p)")
    a("    integer(c_int32_t), intent(in), target :: a(*)")
    a("    type(c_ptr) :: p")
    a("    p = c_loc(a(1))")
    a("  end function gg_i")
    a("")
    a("  ! A failing call: the library's message to the run log and to stdout, then stop, as the")
    a("  ! reference does on its own error conditions.")
    a("  subroutine gg_check(rc, what)")
    a("    integer(c_int), intent(in) :: rc")
    a("    character(len=*), intent(in) :: what")
    a("    character(kind=c_char), pointer :: msg(:)")
    a("    character(len=256) :: text")
    a("    integer :: n")
    a("    if (rc == 0) return")
    a("    text = ''")
    a("    call c_f_pointer(gomgpu_last_error(gg_ctx), msg, [256])")
    a("    do n = 1, 256")
    a("      if (msg(n) == c_null_char) exit")
    a("      text(n:n) = msg(n)")
    a("    end do")
    a("    if (rc == GOMGPU_W_PCG_NOT_CONVERGED) then")
    a("      write(pw_run_log, *) 'gomgpu warning at step ', it, ' (', what, '): ', trim(text)")
    a("      return")
    a("    end if")
    a("    write(pw_run_log, *) 'gomgpu error ', rc, ' at step ', it, ' (', what, '): ', trim(text)")
    a("    write(*, *) 'gomgpu error ', rc, ' at step ', it, ' (', what, '): ', trim(text)")
    a("    stop")
    a("  end subroutine gg_check")
    a("")
    a("  subroutine gg_params(p)")
    a("    type(gomgpu_params), intent(out) :: p")
    for n in pint:
        a(f"    p%{n} = {n}")
    for n in preal:
        a(f"    p%{n} = {n}")
    a("    p%maxtran2 = maxtran2")
    a("    p%trans_sub_iter = trans_sub_iter")
    a("    if (maxtran2 >= 1) p%tran_id1 = tran_id(1)")
    a("    if (maxtran2 >= 2) p%tran_id2 = tran_id(2)")
    a("    if (transport_flag == 1 .and. transport_solver == 2) then")
    a("      ! the ELM solver carries salinity and temperature in slots 1 and 2, no sub-cycles")
    a("      p%maxtran2 = 2;  p%tran_id1 = 1;  p%tran_id2 = 2;  p%trans_sub_iter = 0")
    a("    end if")
    a("    p%precision = GOMGPU_FP64")
    a("    ! the river face velocity from Q_add and the device's own face depth")
    a("    ! (compute_boundary_velocity.f90:32-46 reads eta_node, which lives on the device)")
    a("    p%river_on_device = merge(1, 0, num_Qb_cell > 0)")
    a("    ! main.f90 never calls the closure: Av and Kv keep their start-up values")
    a("    p%turbulence_flag = 0")
    a("    select case (i_turbulence_model_name)")
    a("    case ('kl');  p%turbulence_model = 1")
    a("    case ('ke');  p%turbulence_model = 2")
    a("    case ('kw');  p%turbulence_model = 3")
    a("    case ('ub');  p%turbulence_model = 4")
    a("    case default; p%turbulence_model = 0")
    a("    end select")
    a("    p%stability_function = merge(1, 0, stability_function == 'KC')")
    a("    ! the wind stress, Gamma_T and air_p1 are computed on the host (read_hurricane_ser_1,")
    a("    ! calculate_wind_stress) and uploaded every step")
    a("    p%hurricane_flag = 0")
    a("    p%hurricane_read_interval = 0")
    a("    p%diffmax = 0.0_c_double;  p%xlmin1 = 0.0_c_double;  p%xlmin2 = 0.0_c_double")
    a("  end subroutine gg_params")
    a("")
    a("  subroutine gg_mesh(m)")
    a("    type(gomgpu_mesh), intent(out) :: m")
    for n, k in mesh:
        if n in _MESH_SPECIAL:
            continue
        guard = f"allocated({n})" + (" .and. num_ob_cell > 0" if n == "ob_cell_id" else "")
        a(f"    if ({guard}) m%{n} = {_addr(n, k)}")
    a("    if (num_Qb_cell > 0) then")
    a("      if (allocated(gg_qb_cells)) deallocate(gg_qb_cells)")
    a("      allocate(gg_qb_cells(num_Qb_cell))")
    a("      gg_qb_cells = Q_boundary(1:num_Qb_cell, 1)")
    a("      m%Qb_cell_id = c_loc(gg_qb_cells(1))")
    a("    end if")
    a("  end subroutine gg_mesh")
    a("")
    a("  subroutine gg_state(s)")
    a("    type(gomgpu_state), intent(out) :: s")
    for n, k in state:
        a(f"    if (allocated({n})) s%{n} = {_addr(n, k)}")
    a("  end subroutine gg_state")
    a("")
    a("  ! The per-step forcing the host computed this step (main.f90:59-161).")
    a("  subroutine gg_forcing(f)")
    a("    type(gomgpu_forcing), intent(out) :: f")
    a("    f%elapsed_time = elapsed_time")
    a("    f%spinup_function_tide = spinup_function_tide")
    a("    f%spinup_function_baroclinic = spinup_function_baroclinic")
    a("    f%spinup_function_wind = spinup_function_wind")
    a("    f%hurricane_phase = 0.0_c_double")
    for n in _FORCING_HOST:
        g = f"allocated({n})" + (f" .and. {_FORCING_GUARD[n]}" if n in _FORCING_GUARD else "")
        a(f"    if ({g}) f%{n} = gg_r({n})")
    a("  end subroutine gg_forcing")
    a("")
    a("  ! the stage sequencing, boundary concentrations and host synchronisation (hand-written)")
    a("  include 'gomgpu_bridge_procs.inc'")
    a("end module gomgpu_bridge")
    return "\n".join(L) + "\n"


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "bridge":
        sys.stdout.write(bridge_text())
    else:
        sys.stdout.write(module_text())
