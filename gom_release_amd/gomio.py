"""GOM on-disk input formats: reader and writer of a project's ``input/`` directory.

The files follow the formats the reference's readers expect (reference files, relative
to ``GOM_Rv1.1.0/source``):

* ``main.inp``  card deck C00-C46 in the order of ``read_main_inp.f90:75-1217``; lines
  starting with ``!``, ``C`` or ``c`` in column 1 are comments (``read_main_inp.f90:67``).
  The reader skips comment lines only where the reference does, before each card; the
  data lines of one card are read back to back (``read_main_inp.f90:43-73``).
* ``node.inp``  count, then ``id x y depth`` (``read_node_inp.f90:50-87``).
* ``cell.inp``  count, then ``id 3|4 n1 n2 n3 [n4]``, counter-clockwise
  (``read_cell_inp.f90:39-87``).
* ``harmonic_ser.inp``  ``id ncon phase_shift`` then ``period[h] amp phase f V0 name``
  (``read_harmonic_ser.f90:20-26``).
* ``salt_init.inp`` / ``temp_init.inp``  ``id`` and ``maxlayer`` values per node,
  bottom layer first (``read_salt_init.f90:25-27``).
* ``salt_ser1.inp`` / ``temp_ser1.inp`` / ``eta_ser1.inp`` / ``q_ser1.inp``  per series
  ``id n time_conv time_adjust unit_conv adjust [interp]`` then ``n`` lines ``time value``;
  the ``*_ser2.inp`` form has one header and ``n`` lines ``time v1 .. vN``
  (``read_salt_ser.f90:24-52``, ``read_eta_ser.f90:22-50``, ``read_q_ser.f90:24-55``).
* ``hurricane_ser.inp``  snapshots of ``maxnod`` lines ``id u v p[hPa]``
  (``read_hurricane_ser_0.f90:31-43``, ``read_hurricane_ser_1.f90``).
* ``restart.inp``  see ``gom_release_amd.gomout``.

Every value is written with ``repr`` so that the reader recovers the exact binary64, and
the reader converts decimal text to binary64 with correct rounding, as the Fortran
runtime does.
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass, field
from typing import Dict, Iterator, List, Optional, Sequence, Tuple

import numpy as np


@dataclass
class Series:
    """A piecewise-linear time series as it stands in a ``*_ser*.inp`` file.

    Time in seconds is ``(t + time_adjust) * time_conv`` and the value is
    ``(v + adjust) * unit_conv`` (``read_salt_ser.f90:28-29``); with the default factors
    the times are days."""
    times_days: Sequence[float]
    values: Sequence[float]
    time_conv: float = 86400.0
    time_adjust: float = 0.0
    unit_conv: float = 1.0
    adjust: float = 0.0
    interp: int = 1            # river series only: 1 linear, 2 Lagrange (read_q_ser.f90:27)

    def seconds(self) -> List[float]:
        return [(float(t) + self.time_adjust) * self.time_conv for t in self.times_days]

    def scaled(self) -> List[float]:
        return [(float(v) + self.adjust) * self.unit_conv for v in self.values]


@dataclass
class OpenBoundary:
    """One open-boundary cell face given by its two nodes in the cell's CCW order."""
    n1: int
    n2: int
    eta_type: int = 1          # -1 = radiated from the interior, 1 = harmonic, 2 = time series
    harmonic_id: int = 1
    eta_ser_id: int = 0
    salt_ser_id: int = 0
    temp_ser_id: int = 0


@dataclass
class RiverBoundary:
    n1: int
    n2: int
    q_ser_id: int = 1
    portion: float = 1.0
    salt_ser_id: int = 0
    temp_ser_id: int = 0


@dataclass
class DumpSpec:
    """Cards C44 / C45 (``read_main_inp.f90:1172-1185``): model-state dumps."""
    binary: int = 1            # 0 formatted, 1 unformatted sequential
    time: int = 4              # time unit of the record header: 1 s, 2 min, 3 h, 4 days
    file_freq: int = 99999999  # dumps per file = file_freq // frequency
    frequency: int = 1         # every `frequency` steps in [start, end]
    start: int = 1
    end: int = 99999999


@dataclass
class ContourSpec:
    """Cards C42 / C43 (``read_main_inp.f90:1161-1205``): Tecplot and VTK contour output.
    For 3-D output, `surf` and `full` are IS3D_surf_switch and IS3D_full_switch."""
    format: int = 1            # 1 Tecplot, 2 VTK, 3 both
    file_freq: int = 99999999  # outputs per Tecplot file = file_freq // frequency
    time: int = 4              # time unit: 1 s, 2 min, 3 h, 4 days
    frequency: int = 1         # every `frequency` steps in [start, end]
    start: int = 0
    end: int = 99999999
    unit_conv: float = 1.0     # coordinate factor (1 m, 0.001 km)
    variables: Sequence[int] = (1, 1, 1, 1, 1, 1)   # 2-D: eta u v salt temp rho; 3-D: u v w salt temp rho
    surf: int = 1
    full: int = 0
    grid_format: int = 1


@dataclass
class TserSpec:
    """Card C41 (``read_main_inp.f90:1129-1160``): time series at stations."""
    stations: Sequence[Tuple[int, int, str]] = ()   # (cell, node, name)
    hloc: int = 1              # 1 at the cell, 2 at the node
    time: int = 4
    frequency: int = 1
    format: int = 3            # 1 Tecplot, 2 VTK, 3 both
    fields: Sequence[int] = (1, 1, 1, 1, 1, 1, 0)   # eta H u v salt temp airp


@dataclass
class Case:
    """Everything a GOM project needs for the configurations this package runs."""
    name: str
    x: np.ndarray                  # node x [m]
    y: np.ndarray                  # node y [m]
    depth: np.ndarray              # node depth below MSL [m], as in node.inp
    cells: np.ndarray              # (E, 4) 1-based node ids, 0 in column 4 for triangles
    z_level: np.ndarray            # (L+1,) z_level(0..L), increasing
    msl: float = 0.0
    dt: float = 60.0
    ndt: int = 10
    theta: float = 0.6
    advection_flag: int = 1
    adv_onoff_depth: float = 0.01
    elm_backtrace_flag: int = 0
    elm_sub_iter: int = 1
    elm_min_iter: int = 1
    elm_max_iter: int = 20
    bf_flag: int = 1
    manning: float = 0.025
    bf_height: float = 1e-5
    von_karman: float = 0.41
    baroclinic_flag: int = 1
    ana_density: int = 0
    ref_salt: float = 30.0
    ref_temp: float = 20.0
    dry_depth: float = 0.01
    coriolis_option: int = 1
    lat: float = 30.0
    lon: float = -88.0
    smagorinsky: float = 0.1
    ah0: float = 1.0
    kh0: float = 1.0
    av0: float = 1e-2
    kv0: float = 1e-4
    max_iteration_pcg: int = 1000
    error_tolerance: float = 1e-10
    pcg_result_show: int = 0       # C6: 1 logs the PCG iteration count (pre_conj_grad.f90:224-229)
    ishow_frequency: int = 99999999  # C4: ... every ishow_frequency steps
    transport_flag: int = 1
    transport_solver: int = 1      # 1 TVD (solve_transport_equation_v19), 2 ELM (_ELM_v8)
    trans_sub_iter: int = 1
    h_flux_limiter: int = 0
    v_flux_limiter: int = 0
    is_salt: int = 1
    is_temp: int = 0
    salt_init: Optional[np.ndarray] = None     # (P, L) bottom layer first
    temp_init: Optional[np.ndarray] = None
    salt_series: List[Series] = field(default_factory=list)
    temp_series: List[Series] = field(default_factory=list)
    open_boundaries: List[OpenBoundary] = field(default_factory=list)
    # harmonic series: list of (phase_shift, [(period_h, amp, phase_deg, f, V0_deg, name)])
    harmonics: List[Tuple[float, List[Tuple[float, float, float, float, float, str]]]] = \
        field(default_factory=list)
    eta_series: List[Series] = field(default_factory=list)
    tide_spinup: int = 0
    tide_spinup_period: float = 0.5
    baroclinic_spinup: int = 0
    baroclinic_spinup_period: float = 0.5
    rivers: List[RiverBoundary] = field(default_factory=list)
    q_series: List[Series] = field(default_factory=list)
    start: Tuple[int, int, int, int, int, int] = (2020, 1, 1, 0, 0, 0)
    # C3: year the series times count from (None: the start year) and a shift in days
    data_start_year: Optional[int] = None
    data_time_shift: int = 0
    # C01 / C1
    gravity: float = 9.81
    rho_o: float = 1025.0
    rho_a: float = 1.29
    max_no_neighbor_node: int = 10
    h_node_adjust: float = 0.0
    # C4: hot start from input/restart.inp; restart output -1 at the end, 1 every restart_freq
    restart_in: int = 0
    restart_out: int = 0
    restart_freq: int = 99999999
    hot_start: Optional[Dict[str, np.ndarray]] = None    # restart records (gomout.read_restart)
    # C30: drag formula (1 Large-Pond, 2 Wu, 3 ...) and wind spin-up
    wind_formular: int = 2
    wind_spinup: int = 0
    wind_spinup_period: float = 0.0
    # C31 / C31_1: hurricane wind and pressure fields (formatted, linear in time), read every
    # hurricane_dt seconds between the start and end dates; `hurricane` is a list of node
    # snapshots (u [m/s], v [m/s], p [hPa]) or the path of a hurricane_ser.inp
    hurricane_flag: int = 0
    hurricane_dt: int = 21600
    hurricane_start: Tuple[int, int, int] = (1900, 1, 1)
    hurricane_end: Tuple[int, int, int] = (2999, 12, 31)
    hurricane: Optional[object] = None
    # C44 / C45
    dump2d: Optional[DumpSpec] = None
    dump3d: Optional[DumpSpec] = None
    # C41 - C43: time series and contour output (written by gom_release_amd.gomplot)
    tser: Optional[TserSpec] = None
    out2d: Optional[ContourSpec] = None
    out3d: Optional[ContourSpec] = None
    # vertical turbulence closure (turbulence_closure_v1.f90; not read by read_main_inp.f90,
    # so these never reach the GOM input files): 0 constant Av/Kv, 1 Richardson step,
    # 2 Pacanowski-Philander, 3 GLS; model 'my','kl','ke','kw','ub'; stability 'GA','KC'
    turbulence_flag: int = 0
    turbulence_model: str = "kl"
    stability_function: str = "GA"
    q2min: float = 5.0e-6
    eps_min: float = 1.0e-12
    bgdiff: float = 1.0e-6
    diffmax: float = 1.0
    xlmin: float = 1.0e-3
    pp_depths: Tuple[float, float] = (20.0, 50.0)
    pp_vdmax: Tuple[float, float] = (1.0e-2, 1.0e-2)
    pp_vdmin: Tuple[float, float] = (1.0e-5, 1.0e-5)
    pp_tdmin: float = 1.0e-5

    @property
    def maxnod(self) -> int:
        return int(self.x.shape[0])

    @property
    def maxele(self) -> int:
        return int(self.cells.shape[0])

    @property
    def maxlayer(self) -> int:
        return int(self.z_level.shape[0] - 1)

    def reference_diff_days(self) -> int:
        """calculate_reference_diff_days.f90:17-39, leap-year count included as written
        there (364 days for a year divisible by 4)."""
        y1 = self.start[0] if self.data_start_year is None else int(self.data_start_year)
        y2 = self.start[0]
        d = 0
        while y1 != y2:
            d += 365 if y1 % 4 != 0 else 364
            y1 += 1
        return d - int(self.data_time_shift)


# --------------------------------------------------------------------------- writer

def _r(v: float) -> str:
    return repr(float(v))


def _write_series_file(path: str, series: Sequence[Series], with_interp: bool) -> None:
    with open(path, "w") as f:
        f.write("! generated\n")
        for i, s in enumerate(series, 1):
            extra = f"  {int(s.interp)}" if with_interp else ""
            f.write(f"  {i} {len(s.times_days)} {_r(s.time_conv)} {_r(s.time_adjust)} "
                    f"{_r(s.unit_conv)} {_r(s.adjust)}{extra}\n")
            for t, v in zip(s.times_days, s.values):
                f.write(f"  {_r(t)} {_r(v)}\n")


def write_project(case: Case, root: str) -> str:
    """Write ``root/input/*`` and create ``root/output``; returns ``root``."""
    inp = os.path.join(root, "input")
    os.makedirs(inp, exist_ok=True)
    os.makedirs(os.path.join(root, "output"), exist_ok=True)
    P, E, L = case.maxnod, case.maxele, case.maxlayer

    with open(os.path.join(inp, "node.inp"), "w") as f:
        f.write("! node.inp: id x y depth\n")
        f.write(f"  {P}\n")
        for i in range(P):
            f.write(f"  {i + 1} {_r(case.x[i])} {_r(case.y[i])} {_r(case.depth[i])}\n")

    with open(os.path.join(inp, "cell.inp"), "w") as f:
        f.write("! cell.inp: id n nodes (counter-clockwise)\n")
        f.write(f"  {E}\n")
        for i in range(E):
            c = case.cells[i]
            n = 4 if c[3] != 0 else 3
            f.write(f"  {i + 1} {n} " + " ".join(str(int(v)) for v in c[:n]) + "\n")

    if case.is_salt and case.salt_init is not None:
        with open(os.path.join(inp, "salt_init.inp"), "w") as f:
            f.write("! salt_init.inp\n")
            for i in range(P):
                f.write(f"  {i + 1} " + " ".join(_r(v) for v in case.salt_init[i]) + "\n")
    if case.is_temp and case.temp_init is not None:
        with open(os.path.join(inp, "temp_init.inp"), "w") as f:
            f.write("! temp_init.inp\n")
            for i in range(P):
                f.write(f"  {i + 1} " + " ".join(_r(v) for v in case.temp_init[i]) + "\n")
    if case.is_salt and case.salt_series:
        _write_series_file(os.path.join(inp, "salt_ser1.inp"), case.salt_series, False)
    if case.is_temp and case.temp_series:
        _write_series_file(os.path.join(inp, "temp_ser1.inp"), case.temp_series, False)
    if case.rivers:
        _write_series_file(os.path.join(inp, "q_ser1.inp"), case.q_series, True)
    if case.harmonics:
        with open(os.path.join(inp, "harmonic_ser.inp"), "w") as f:
            f.write("! harmonic_ser.inp\n")
            for i, (shift, cons) in enumerate(case.harmonics, 1):
                f.write(f"  {i} {len(cons)} {_r(shift)}\n")
                for (per, amp, ph, fac, v0, nm) in cons:
                    f.write(f"  {_r(per)} {_r(amp)} {_r(ph)} {_r(fac)} {_r(v0)} {nm}\n")
    if case.eta_series and any(b.eta_type == 2 for b in case.open_boundaries):
        _write_series_file(os.path.join(inp, "eta_ser1.inp"), case.eta_series, False)
    if case.hurricane_flag == 1:
        if case.hurricane is None:
            raise ValueError("hurricane_flag = 1 needs the hurricane snapshots (Case.hurricane)")
        write_hurricane_ser(os.path.join(inp, "hurricane_ser.inp"), hurricane_snapshots(case, pascal=False))
    if case.restart_in == 1:
        if case.hot_start is None:
            raise ValueError("restart_in = 1 needs the restart records (Case.hot_start)")
        from .gomout import write_restart
        write_restart(os.path.join(inp, "restart.inp"), case.hot_start)

    with open(os.path.join(inp, "main.inp"), "w") as f:
        f.write(_main_inp(case))
    return root


def _main_inp(c: Case) -> str:
    L = c.maxlayer
    lines: List[str] = []
    card = lines.append

    def data(*vals) -> None:
        lines.append("\t\t" + "\t".join(v if isinstance(v, str) else
                                        (str(v) if isinstance(v, (int, np.integer)) else _r(v))
                                        for v in vals))
    card("!== generated GOM main.inp ==")
    card("C00\tproject_name")
    card("\t\t" + c.name)
    card("C01\tgravity rho_o rho_a max_no_neighbor_node")
    data(float(c.gravity), float(c.rho_o), float(c.rho_a), int(c.max_no_neighbor_node))
    card("C1\tmaxnod maxele h_node_adjust ana_depth coordinate_system lon lat voronoi node_mirr cell_mirr")
    data(c.maxnod, c.maxele, float(c.h_node_adjust), 0, 1, c.lon, c.lat, 0, 0, 0)
    card("C2\tmaxlayer MSL z_level(0)")
    data(L, c.msl, float(c.z_level[0]))
    card("C2_1\tserial_num z_level (surface first)")
    for k in range(L, 0, -1):
        data(k, float(c.z_level[k]))
    card("C3\tdt ndt data_start_year data_time_shift start_year month day hour minute second")
    y, mo, d, h, mi, s = c.start
    dsy = y if c.data_start_year is None else int(c.data_start_year)
    data(c.dt, c.ndt, dsy, int(c.data_time_shift), y, mo, d, h, mi, s)
    card("C4\trestart_in restart_out restart_freq ishow ishow_frequency")
    data(int(c.restart_in), int(c.restart_out), int(c.restart_freq), 0, int(c.ishow_frequency))
    card("C5_01\ttheta advection_flag adv_onoff_depth ELM_backtrace_flag ELM_sub_iter ELM_min_iter ELM_max_iter")
    data(c.theta, c.advection_flag, c.adv_onoff_depth, c.elm_backtrace_flag, c.elm_sub_iter,
         c.elm_min_iter, c.elm_max_iter)
    card("C5_02\tbf_flag bf_varying manning bf_height von_Karman")
    data(c.bf_flag, 0, c.manning, c.bf_height, c.von_karman)
    card("C5_03\tbaroclinic_flag ana_density ref_salt ref_temp dry_depth Coriolis_option")
    data(c.baroclinic_flag, c.ana_density, c.ref_salt, c.ref_temp, c.dry_depth, c.coriolis_option)
    card("C5_04\tSmagorinsky_parameter Ah_0 Kh_0 Av_0 Kv_0")
    data(c.smagorinsky, c.ah0, c.kh0, c.av0, c.kv0)
    card("C6\tmax_iteration_pcg error_tolerance pcg_result_show")
    data(c.max_iteration_pcg, c.error_tolerance, int(c.pcg_result_show))
    card("C7\ttransport_flag transport_solver")
    data(c.transport_flag, int(c.transport_solver))
    if c.transport_flag == 1:
        if c.transport_solver == 1:
            card("C7_1\ttrans_sub_iter h_flux_limiter v_flux_limiter")
            data(c.trans_sub_iter, c.h_flux_limiter, c.v_flux_limiter)
        card("C7_2\tis_tran num_tran_ser tran_ser_shape")
        data(c.is_salt, len(c.salt_series) if c.is_salt else 0, 1)
        data(c.is_temp, len(c.temp_series) if c.is_temp else 0, 1)
        if c.is_temp:
            card("C7_3\theat_option sol_swr fWz_a fWz_b fWz_c wind_height")
            data(1, 2, 9.2, 0.46, 2.0, 2.0)
            card("C7_4\tlight_extinction sol_absorb sed_water_exchange T_sed sed_temp_coeff")
            data(0.45, 0.45, 0.3, 11.5, 0.0)
    card("C8\tterminate_check_freq eta_min eta_max uv salt temp")
    data(99999999, -30.0, 30.0, 50.0, 60.0, 60.0)
    nob = len(c.open_boundaries)
    ntide = sum(1 for b in c.open_boundaries if b.eta_type == 1)
    neta = nob - ntide
    card("C20\topen boundary")
    data(nob, 0)
    data(ntide, len(c.harmonics) if ntide else 0, 0)
    data(neta, len(c.eta_series) if neta else 0, 0, 1)
    if nob:
        card("C20_1\tserial n1 n2 ob_eta_type harmonic_ser_id eta_ser_id salt_ser_id temp_ser_id")
        for i, b in enumerate(c.open_boundaries, 1):
            data(i, b.n1, b.n2, b.eta_type, b.harmonic_id, b.eta_ser_id, b.salt_ser_id, b.temp_ser_id)
    card("C21\ttide_spinup tide_spinup_period / baroclinic_spinup baroclinic_spinup_period")
    data(c.tide_spinup, c.tide_spinup_period)
    data(c.baroclinic_spinup, c.baroclinic_spinup_period)
    card("C22\tnum_Qb_cell Qbc_info_option num_Q_ser check_Q Q_ser_shape")
    data(len(c.rivers), 0, len(c.q_series), 0, 1)
    if c.rivers:
        card("C22_1\tserial n1 n2 Q_ser_id Q_portion Q_salt_ser_id Q_temp_ser_id")
        for i, r in enumerate(c.rivers, 1):
            data(i, r.n1, r.n2, r.q_ser_id, r.portion, r.salt_ser_id, r.temp_ser_id)
    card("C23\tnum_WR_cell")
    data(0)
    card("C24\tnum_SS_cell")
    data(0)
    card("C25\tno_Etide_species Etide_cutoff_depth")
    data(0, 40.0)
    card("C30\twind_flag airp_flag num_windp_ser wind_formular wind_spinup wind_spinup_period")
    data(0, 0, 0, int(c.wind_formular), int(c.wind_spinup), float(c.wind_spinup_period))
    card("C31\thurricane_flag hurricane_data_type hurricane_interp_method hurricane_dt")
    data(int(c.hurricane_flag), 1, 1, int(c.hurricane_dt))
    if c.hurricane_flag == 1:
        card("C31_1\thurricane start year month day / end year month day")
        data(*[int(v) for v in c.hurricane_start])
        data(*[int(v) for v in c.hurricane_end])
    card("C32\tholland_flag holland_start_jday holland_end_jday")
    data(0, 0.0, 0.0)
    card("C40\tcheck_grid")
    data(0, 0, 0, 1, 0, 0.001)
    card("C41\ttser")
    ts = c.tser
    if ts is None or not ts.stations:
        data(0, 0, 1, 4, 24, 1)
    else:
        data(len(ts.stations), 0, ts.hloc, ts.time, ts.frequency, ts.format)
        card("C41_1\tserial cell node name")
        for i, (ce, nd, nm) in enumerate(ts.stations):
            data(i + 1, int(ce), int(nd), nm)
        card("C41_2\teta H u v salt temp airp")
        data(*[int(v) for v in ts.fields])
    o2 = c.out2d
    card("C42\t2D output")
    if o2 is None:
        data(0, 0)
        data(1, 0, 99999999)
        data(4, 99999999, 0, 99999999, 0.001)
    else:
        data(1, 0)
        data(o2.format, 0, o2.file_freq)
        data(o2.time, o2.frequency, o2.start, o2.end, float(o2.unit_conv))
    card("C42_1\tvariables")
    for nm, v in zip(("eta", "u", "v", "salt", "temp", "rho"), o2.variables if o2 else (0,) * 6):
        data(nm, int(v))
    o3 = c.out3d
    card("C43\t3D output")
    if o3 is None:
        data(0, 1, 0)
        data(1, 0, 99999999)
        data(4, 99999999, 0, 99999999, 0.001)
    else:
        data(o3.full, o3.grid_format, o3.surf)
        data(o3.format, 0, o3.file_freq)
        data(o3.time, o3.frequency, o3.start, o3.end, float(o3.unit_conv))
    card("C43_1\tvariables")
    for nm, v in zip(("u", "v", "w", "salt", "temp", "rho"), o3.variables if o3 else (0,) * 6):
        data(nm, int(v))
    for tag, spec in (("C44\t2D dump", c.dump2d), ("C45\t3D dump", c.dump3d)):
        card(tag)
        sp = spec or DumpSpec()
        data(1 if spec is not None else 0, sp.binary, sp.time, sp.file_freq)
        data(sp.frequency, sp.start, sp.end)
    card("C46\tdiagnostics")
    data(0, 0, 0, 0)
    data(0, 0, 0)
    return "\n".join(lines) + "\n"


# --------------------------------------------------------------------------- reader

def is_comment(line: str) -> bool:
    """A ``!``, ``C`` or ``c`` in column 1 (skip_header_lines.f90:36)."""
    return line[:1] in ("!", "C", "c")


_REAL_NOEXP = re.compile(r"^([+-]?(?:\d+\.?\d*|\.\d+))([+-]\d+)$")
_REPEAT = re.compile(r"^(\d+)\*(.*)$", re.S)


class FormatError(ValueError):
    """An input file does not follow the format the reference reads."""


class _EndOfDeck(Exception):
    pass


class ListReader:
    """Fortran list-directed input over the records (lines) of a text file.

    ``read`` starts at the next record, takes values across as many records as it needs
    (blank or comma separated; quoted or bare strings; ``r*v`` repeats; a null value or a
    ``/`` leaves the item unset, returned as None) and drops the rest of its last record,
    which is what a ``read(unit,*)`` statement does."""

    def __init__(self, lines: List[str], name: str = "<text>"):
        self.lines = lines
        self.pos = 0
        self.name = name

    @classmethod
    def from_file(cls, path: str) -> "ListReader":
        with open(path, "r", errors="replace") as f:
            return cls(f.read().splitlines(), os.path.basename(path))

    def at_eof(self) -> bool:
        return self.pos >= len(self.lines)

    def skip_header(self) -> None:
        """skip_header_lines.f90: skip comment records at the current position."""
        while self.pos < len(self.lines) and is_comment(self.lines[self.pos]):
            self.pos += 1

    def card(self) -> None:
        """The top of the card loop of read_main_inp.f90:43-73: skip comment records;
        the end of the file ends the deck."""
        self.skip_header()
        if self.at_eof():
            raise _EndOfDeck()

    def line(self) -> str:
        """``read(unit,'(A)')``: the next record as it stands."""
        if self.at_eof():
            raise FormatError(f"{self.name}: end of file")
        s = self.lines[self.pos]
        self.pos += 1
        return s

    @staticmethod
    def _items(rec: str) -> Iterator[Optional[str]]:
        """Values of one record; None for a null value; the string '/' ends the read."""
        i, n = 0, len(rec)
        expect_value = True
        while True:
            while i < n and rec[i] in " \t\r":
                i += 1
            if i >= n:
                return
            ch = rec[i]
            if ch == ",":
                if expect_value:
                    yield None
                expect_value = True
                i += 1
                continue
            if ch == "/":
                yield "/"
                return
            if ch in "'\"":
                j = i + 1
                buf = []
                while j < n:
                    if rec[j] == ch:
                        if j + 1 < n and rec[j + 1] == ch:
                            buf.append(ch)
                            j += 2
                            continue
                        break
                    buf.append(rec[j])
                    j += 1
                tok = "".join(buf)
                i = j + 1
                yield "\x00" + tok            # quoted: never a repeat or a number
            else:
                j = i
                while j < n and rec[j] not in " \t\r,/":
                    j += 1
                tok = rec[i:j]
                i = j
                m = _REPEAT.match(tok)
                if m:
                    r, v = int(m.group(1)), m.group(2)
                    if v == "" and i < n and rec[i] in "'\"":
                        # r*'quoted'
                        q = rec[i]
                        j = rec.find(q, i + 1)
                        v = "\x00" + rec[i + 1:j if j >= 0 else n]
                        i = (j + 1) if j >= 0 else n
                    for _ in range(r):
                        yield v if v != "" else None
                else:
                    yield tok
            expect_value = False

    def read(self, kinds: str) -> List:
        """List-directed read of len(kinds) items: 'i' integer, 'r' real, 'c' character."""
        out: List = [None] * len(kinds)
        k = 0
        while k < len(kinds):
            if self.at_eof():
                raise FormatError(f"{self.name}: end of file while reading {len(kinds)} values")
            rec = self.lines[self.pos]
            self.pos += 1
            for tok in self._items(rec):
                if tok == "/":
                    return out
                if tok is not None:
                    out[k] = self._convert(tok, kinds[k])
                k += 1
                if k == len(kinds):
                    break
        return out

    def _convert(self, tok: str, kind: str):
        quoted = tok.startswith("\x00")
        if quoted:
            tok = tok[1:]
        if kind == "c":
            return tok
        if quoted:
            raise FormatError(f"{self.name}:{self.pos}: string {tok!r} where a number is expected")
        try:
            if kind == "i":
                return int(tok)
            t = tok.replace("d", "e").replace("D", "e").replace("q", "e").replace("Q", "e")
            m = _REAL_NOEXP.match(t)
            if m:
                t = m.group(1) + "e" + m.group(2)
            return float(t)
        except ValueError:
            raise FormatError(f"{self.name}:{self.pos}: cannot read {tok!r} as "
                              f"{'an integer' if kind == 'i' else 'a real'}") from None


class MainInp(dict):
    """The variables read_main_inp.f90 sets, under the reference's names."""

    def __getattr__(self, k):
        try:
            return self[k]
        except KeyError:
            raise AttributeError(k) from None

    def __setattr__(self, k, v):
        self[k] = v


def _rows(rd: ListReader, n: int, kinds: str) -> List[List]:
    return [rd.read(kinds) for _ in range(n)]


def read_main_inp(path: str) -> MainInp:
    """Parse a ``main.inp`` card deck as read_main_inp.f90:40-1217 does.

    Returns the card variables under the reference's names. The info files that cards
    C20, C22 and C41 may point to (``Obc_info.inp``, ``Qbc_info.inp``,
    ``tser_station_info.inp``) are read from the same directory. Inputs the reference
    stops on raise FormatError with the reference's reason. A deck that ends early (the
    reference then stops reading and keeps its defaults) gives ``complete`` False."""
    rd = ListReader.from_file(path)
    m = MainInp()
    try:
        _main_cards(rd, m, os.path.dirname(os.path.abspath(path)))
        m.complete = True
    except _EndOfDeck:
        m.complete = False
    return m


def _main_cards(rd: ListReader, m: MainInp, inp: str) -> None:
    def fail(card: str, why: str):
        raise FormatError(f"main.inp card C{card}: {why}")

    rd.card()                                                   # C00 (:78-81)
    m.project_name = rd.line().strip()
    rd.card()                                                   # C01 (:82-84)
    m.gravity, m.rho_o, m.rho_a, m.max_no_neighbor_node = rd.read("rrri")
    rd.card()                                                   # C1 (:85-95)
    (m.maxnod, m.maxele, m.h_node_adjust, m.ana_depth, m.coordinate_system, m.lon, m.lat,
     m.voronoi, m.node_mirr, m.cell_mirr) = rd.read("iiriirriii")
    rd.card()                                                   # C2 (:97-99)
    m.maxlayer, m.MSL, z0 = rd.read("irr")
    L = m.maxlayer
    z = [0.0] * (L + 1)
    z[0] = z0
    rd.card()                                                   # C2_1 (:100-147)
    for k in range(L, 0, -1):
        _serial, z[k] = rd.read("ir")
        if abs(z[k] - m.MSL) < 1.0e-5:
            fail("2_1", f"MSL is too close to level line at z layer number : {k}")
    for k in range(1, L + 1):
        if abs(z[k - 1] + (z[k] - z[k - 1]) - z[k]) > 1.0e-6:
            fail("2_1", f"wrong input from vertical grid file at level {k}")
    if m.MSL > z[L]:
        fail("2_1", "Mean Sea Level(MSL) is above the maximum vertical level")
    if m.MSL < z[0]:
        fail("2_1", "Mean Sea Level(MSL) is below the minimum vertical level")
    m.z_level = z
    rd.card()                                                   # C3 (:148-167)
    (m.dt, m.ndt, m.data_start_year, m.data_time_shift, m.start_year, m.start_month,
     m.start_day, m.start_hour, m.start_minute, m.start_second) = rd.read("riiiiiiiii")
    rd.card()                                                   # C4 (:168-171)
    m.restart_in, m.restart_out, m.restart_freq, m.ishow, m.ishow_frequency = rd.read("iiiii")
    rd.card()                                                   # C5_01 (:172-197)
    (m.theta, m.advection_flag, m.adv_onoff_depth, m.ELM_backtrace_flag, m.ELM_sub_iter,
     m.ELM_min_iter, m.ELM_max_iter) = rd.read("ririiii")
    rd.card()                                                   # C5_02 (:198-233)
    m.bf_flag, m.bf_varying, m.manning, m.bf_height, m.von_Karman = rd.read("iirrr")
    if not 0 <= m.bf_flag <= 2:
        fail("5_02", "bf_flag is not in range")
    if not 0 <= m.bf_varying <= 1:
        fail("5_02", "bf_varying is not in range")
    rd.card()                                                   # C5_03 (:234-240)
    (m.baroclinic_flag, m.ana_density, m.ref_salt, m.ref_temp, m.dry_depth,
     m.Coriolis_option) = rd.read("iirrri")
    rd.card()                                                   # C5_04 (:241-258)
    m.smagorinsky_parameter, m.Ah_0, m.Kh_0, m.Av_0, m.Kv_0 = rd.read("rrrrr")
    rd.card()                                                   # C6 (:259-262)
    m.max_iteration_pcg, m.error_tolerance, m.pcg_result_show = rd.read("iri")
    rd.card()                                                   # C7 (:263-274)
    m.transport_flag, m.transport_solver = rd.read("ii")
    m.is_tran, m.num_tran_ser, m.tran_ser_shape = [0, 0], [0, 0], [0, 0]
    if m.transport_flag == 1:
        if m.transport_solver == 1:
            rd.card()                                           # C7_1 (:275-278)
            m.trans_sub_iter, m.h_flux_limiter, m.v_flux_limiter = rd.read("iii")
        rd.card()                                               # C7_2 (:279-341)
        for i in range(2):
            m.is_tran[i], m.num_tran_ser[i], m.tran_ser_shape[i] = rd.read("iii")
        if m.is_tran[1] == 1:
            rd.card()                                           # C7_3 (:343-347)
            (m.heat_option, m.sol_swr, m.fWz_a, m.fWz_b, m.fWz_c,
             m.wind_height) = rd.read("iirrrr")
            rd.card()                                           # C7_4 (:349-352)
            (m.light_extinction, m.sol_absorb, m.sed_water_exchange, m.T_sed,
             m.sed_temp_coeff) = rd.read("rrrrr")
    m.is_salt, m.is_temp = m.is_tran
    rd.card()                                                   # C8 (:354-360)
    (m.terminate_check_freq, m.eta_min_terminate, m.eta_max_terminate, m.uv_terminate,
     m.salt_terminate, m.temp_terminate) = rd.read("irrrrr")
    rd.card()                                                   # C20 (:364-418)
    m.num_ob_cell, m.ob_info_option = rd.read("ii")
    m.num_tide_bc, m.num_harmonic_ser, m.check_tide = rd.read("iii")
    m.num_eta_bc, m.num_eta_ser, m.check_eta, m.eta_ser_shape = rd.read("iiii")
    if m.num_ob_cell != m.num_tide_bc + m.num_eta_bc:
        fail("20", "num_ob_cell is not equal to (num_tide_bc + num_eta_bc)")
    if m.num_tide_bc == 0 and m.num_harmonic_ser != 0:
        fail("20", "If [num_tide_bc] == 0, [num_harmonic_ser] should be 0.")
    if m.num_eta_bc == 0 and m.num_eta_ser != 0:
        fail("20", "If [num_eta_bc] == 0, [num_eta_ser] should be 0.")
    m.ob_rows = []
    if m.num_ob_cell > 0:
        if m.num_tide_bc > 0 and m.num_harmonic_ser <= 0:
            fail("20", "num_tide_bc > 0 needs num_harmonic_ser > 0")
        if m.num_eta_bc > 0 and (m.num_eta_ser <= 0 or m.eta_ser_shape not in (1, 2)):
            fail("20", "num_eta_bc > 0 needs num_eta_ser > 0 and eta_ser_shape 1 or 2")
        if m.ob_info_option == 0:
            rd.card()                                           # C20_1 (:538-650)
            src = rd
        else:
            src = ListReader.from_file(os.path.join(inp, "Obc_info.inp"))
            src.skip_header()
        m.ob_rows = _rows(src, m.num_ob_cell, "iiiiiiii")
        _check_ob_rows(m, fail)
    rd.card()                                                   # C21 (:651-656)
    m.tide_spinup, m.tide_spinup_period = rd.read("ir")
    m.baroclinic_spinup, m.baroclinic_spinup_period = rd.read("ir")
    rd.card()                                                   # C22 (:657-742)
    m.num_Qb_cell, m.Qbc_info_option, m.num_Q_ser, m.check_Q, m.Q_ser_shape = rd.read("iiiii")
    m.q_rows = []
    if m.num_Qb_cell > 0:
        if m.Qbc_info_option == 0:
            rd.card()                                           # C22_1 (:743-771)
            src = rd
        else:
            src = ListReader.from_file(os.path.join(inp, "Qbc_info.inp"))
            src.skip_header()
        m.q_rows = _rows(src, m.num_Qb_cell, "iiiirii")
        if max(r[3] for r in m.q_rows) > m.num_Q_ser:
            fail("22_1", "One of Q_ser_id(i) is geater than num_Q_ser")
    rd.card()                                                   # C23 (:772-781)
    (m.num_WR_cell,) = rd.read("i")
    m.wr_rows = []
    if m.num_WR_cell > 0:
        rd.card()                                               # C23_1 (:782-800)
        m.wr_rows = _rows(rd, m.num_WR_cell, "iiiiirii")
    rd.card()                                                   # C24 (:801-808)
    (m.num_SS_cell,) = rd.read("i")
    m.ss_rows = []
    if m.num_SS_cell > 0:
        rd.card()                                               # C24_1 (:809-816)
        m.ss_rows = _rows(rd, m.num_SS_cell, "iiiirii")
    rd.card()                                                   # C25 (:817-824)
    m.no_Etide_species, m.Etide_cutoff_depth = rd.read("ir")
    m.etide_rows = []
    if m.no_Etide_species > 0:
        rd.card()                                               # C25_1 (:825-860)
        m.etide_rows = _rows(rd, m.no_Etide_species, "irrrr")
    rd.card()                                                   # C30 (:864-878)
    (m.wind_flag, m.airp_flag, m.num_windp_ser, m.wind_formular, m.wind_spinup,
     m.wind_spinup_period) = rd.read("iiiiir")
    rd.card()                                                   # C31 (:879-910)
    m.hurricane_flag, m.hurricane_data_type, m.hurricane_interp_method, m.hurricane_dt = \
        rd.read("iiii")
    if m.hurricane_flag == 1:
        rd.card()                                               # C31_1 (:911-916)
        m.hurricane_start = tuple(rd.read("iii"))
        m.hurricane_end = tuple(rd.read("iii"))
    rd.card()                                                   # C32 (:917-925)
    m.holland_flag, m.holland_start_jday, m.holland_end_jday = rd.read("irr")
    rd.card()                                                   # C40 (:929-933)
    (m.check_grid_2D, m.check_grid_2DO, m.check_grid_3D, m.check_grid_format,
     m.check_grid_info, m.check_grid_unit_conv) = rd.read("iiiiir")
    rd.card()                                                   # C41 (:934-960)
    (m.tser_station_num, m.tser_info_option, m.tser_hloc, m.tser_time, m.tser_frequency,
     m.tser_format) = rd.read("iiiiii")
    m.tser_stations = []
    if m.tser_station_num > 0:
        if m.tser_info_option == 0:
            rd.card()                                           # C41_1 (:961-966)
            src = rd
        else:
            src = ListReader.from_file(os.path.join(inp, "tser_station_info.inp"))
            src.skip_header()
        m.tser_stations = _rows(src, m.tser_station_num, "iiic")
        rd.card()                                               # C41_2 (:967-971)
        (m.tser_eta, m.tser_H, m.tser_u, m.tser_v, m.tser_salt, m.tser_temp,
         m.tser_airp) = rd.read("iiiiiii")
    rd.card()                                                   # C42 (:972-980)
    m.IS2D_switch, m.IS2D_flood_map = rd.read("ii")
    m.IS2D_format, m.IS2D_binary, m.IS2D_File_freq = rd.read("iii")
    m.IS2D_time, m.IS2D_frequency, m.IS2D_start, m.IS2D_end, m.IS2D_unit_conv = rd.read("iiiir")
    rd.card()                                                   # C42_1 (:981-988)
    m.IS2D_variable = [rd.read("ci")[1] for _ in range(6)]
    rd.card()                                                   # C43 (:989-997)
    m.IS3D_full_switch, m.IS3D_grid_format, m.IS3D_surf_switch = rd.read("iii")
    m.IS3D_format, m.IS3D_binary, m.IS3D_File_freq = rd.read("iii")
    m.IS3D_time, m.IS3D_frequency, m.IS3D_start, m.IS3D_end, m.IS3D_unit_conv = rd.read("iiiir")
    rd.card()                                                   # C43_1 (:998-1005)
    m.IS3D_variable = [rd.read("ci")[1] for _ in range(6)]
    for tag in ("2D", "3D"):                                    # C44, C45 (:1006-1019)
        rd.card()
        (m[f"IS{tag}_dump_switch"], m[f"IS{tag}_dump_binary"], m[f"IS{tag}_dump_time"],
         m[f"IS{tag}_dump_File_freq"]) = rd.read("iiii")
        (m[f"IS{tag}_dump_frequency"], m[f"IS{tag}_dump_start"],
         m[f"IS{tag}_dump_end"]) = rd.read("iii")
    rd.card()                                                   # C46 (:1020-1027)
    m.dia_advection, m.dia_momentum, m.dia_freesurface, m.dia_eta_at_ob = rd.read("iiii")
    m.dia_bottom_friction, m.dia_face_velocity, m.dia_node_velocity = rd.read("iii")


def _check_ob_rows(m: MainInp, fail) -> None:
    """The consistency checks of read_main_inp.f90:552-648 on the open-boundary rows."""
    counts = dict(h=0, e=0, s=0, t=0)
    for (_serial, _n1, _n2, typ, hid, eid, sid, tid) in m.ob_rows:
        if typ not in (-1, 1, 2):
            fail("20_1", f"ob_eta_type {typ} is not -1, 1 or 2")
        counts["h"] += hid > 0
        counts["e"] += eid > 0
        counts["s"] += sid > 0
        counts["t"] += tid > 0
        if typ == -1 and (hid > 0 or eid > 0 or sid > 0 or tid > 0):
            fail("20_1", "[ob_eta_type == -1] takes no series ids")
        if typ == 1 and not 1 <= hid <= m.num_harmonic_ser:
            fail("20_1", "[ob_eta_type == 1], but [harmonic_ser_id < 1] or "
                         "[harmonic_ser_id > num_harmonic_ser]")
        if typ == 2 and not 1 <= eid <= m.num_eta_ser:
            fail("20_1", "[ob_eta_type == 2], but [eta_ser_id < 1] or [eta_ser_id > num_eta_ser]")
    if m.num_tide_bc == 0 and counts["h"] > 0:
        fail("20_1", "[num_tide_bc == 0] but harmonic series ids are given")
    if m.num_eta_bc == 0 and counts["e"] > 0:
        fail("20_1", "[num_eta_bc == 0] but eta series ids are given")
    if m.transport_flag == 0 and (counts["s"] > 0 or counts["t"] > 0):
        fail("20_1", "[transport_flag == 0] but salt/temp series ids are given")
    if m.transport_flag > 0:
        if m.is_tran[0] > 0 and max(r[6] for r in m.ob_rows) > m.num_tran_ser[0]:
            fail("20_1", "maxval(salt_ser_id) > num_tran_ser(1)")
        if m.is_tran[1] > 0 and max(r[7] for r in m.ob_rows) > m.num_tran_ser[1]:
            fail("20_1", "maxval(temp_ser_id) > num_tran_ser(2)")


def read_node_inp(path: str, maxnod: int) -> np.ndarray:
    """(maxnod, 3) array of x, y, depth as written (read_node_inp.f90:50-87)."""
    rd = ListReader.from_file(path)
    rd.skip_header()
    (n,) = rd.read("i")
    if n != maxnod:
        raise FormatError("Total node number does not mathch: STOP")
    out = np.empty((maxnod, 3))
    for i in range(maxnod):
        _id, out[i, 0], out[i, 1], out[i, 2] = rd.read("irrr")
    return out


def read_cell_inp(path: str, maxele: int) -> np.ndarray:
    """(maxele, 4) 1-based node ids, 0 in column 4 for triangles (read_cell_inp.f90:39-87)."""
    rd = ListReader.from_file(path)
    rd.skip_header()
    (n,) = rd.read("i")
    if n != maxele:
        raise FormatError("Total cell number does not match")
    cells = np.zeros((maxele, 4), dtype=np.int64)
    for i in range(maxele):
        rec = rd.lines[rd.pos]
        head = ListReader([rec], rd.name).read("ii")
        nv = head[1]
        if nv not in (3, 4):
            raise FormatError(f"{rd.name}: cell {i + 1} has {nv} nodes")
        vals = rd.read("ii" + "i" * nv)
        cells[i, :nv] = vals[2:]
    return cells


def read_init(path: str, maxnod: int, maxlayer: int) -> np.ndarray:
    """(maxnod, maxlayer) tracer values per node, bottom layer first (read_salt_init.f90:25-27)."""
    rd = ListReader.from_file(path)
    rd.skip_header()
    out = np.empty((maxnod, maxlayer))
    kinds = "i" + "r" * maxlayer
    for i in range(maxnod):
        out[i] = rd.read(kinds)[1:]
    return out


def read_series(path: str, count: int, shape: int = 1, with_interp: bool = False) -> List[Series]:
    """`count` series of a ``*_ser1.inp`` (shape 1) or ``*_ser2.inp`` (shape 2) file
    (read_salt_ser.f90:22-52, read_q_ser.f90:24-55, read_eta_ser.f90:20-50)."""
    rd = ListReader.from_file(path)
    rd.skip_header()
    out: List[Series] = []
    if shape == 1:
        for _ in range(count):
            hdr = rd.read("iirrrr" + ("i" if with_interp else ""))
            n = hdr[1]
            rows = [rd.read("rr") for _ in range(n)]
            out.append(Series([r[0] for r in rows], [r[1] for r in rows], hdr[2], hdr[3], hdr[4],
                              hdr[5], hdr[6] if with_interp and hdr[6] is not None else 1))
        return out
    if shape != 2:
        raise FormatError(f"{os.path.basename(path)}: series shape {shape} is not 1 or 2")
    hdr = rd.read("iirrrr")
    n = hdr[1]
    rows = [rd.read("r" * (count + 1)) for _ in range(n)]
    for i in range(count):
        out.append(Series([r[0] for r in rows], [r[i + 1] for r in rows], hdr[2], hdr[3], hdr[4],
                          hdr[5], 1))
    return out


def read_harmonic_ser(path: str, count: int):
    """[(phase_shift, [(period_h, amp, phase, f, V0, name)])] (read_harmonic_ser.f90:18-27)."""
    rd = ListReader.from_file(path)
    rd.skip_header()
    out = []
    for _ in range(count):
        _id, ncon, shift = rd.read("iir")
        cons = [tuple(rd.read("rrrrrc")) for _ in range(ncon)]
        out.append((shift, cons))
    return out


def write_hurricane_ser(path: str, snapshots) -> None:
    """hurricane_ser.inp, formatted (hurricane_data_type 1): per snapshot ``maxnod`` lines
    ``id u v p`` with p in hPa (read_hurricane_ser_0.f90:31-43)."""
    with open(path, "w") as f:
        f.write("! hurricane_ser.inp: id u[m/s] v[m/s] p[hPa], one block per snapshot\n")
        for (u, v, p) in snapshots:
            for i in range(len(u)):
                f.write(f"  {i + 1} {_r(u[i])} {_r(v[i])} {_r(p[i])}\n")


def read_hurricane_ser(path: str, maxnod: int, pascal: bool = True) -> Iterator[Tuple[np.ndarray, np.ndarray, np.ndarray]]:
    """Snapshots (u, v, p) of a formatted hurricane_ser.inp in file order; with `pascal`
    the pressure is converted from hPa with ``* 100`` as read_hurricane_ser_0.f90:34-35
    does."""
    rd = ListReader.from_file(path)
    rd.skip_header()
    while True:
        rd_pos = rd.pos
        while rd_pos < len(rd.lines) and not rd.lines[rd_pos].strip():
            rd_pos += 1
        if rd_pos >= len(rd.lines):
            return
        u = np.empty(maxnod)
        v = np.empty(maxnod)
        p = np.empty(maxnod)
        for i in range(maxnod):
            _id, u[i], v[i], p[i] = rd.read("irrr")
        yield u, v, (p * 100.0 if pascal else p)


def hurricane_snapshots(case: "Case", pascal: bool = True):
    """The node snapshots of a case's hurricane fields in time order, from its list or its
    hurricane_ser.inp; pressures in Pa (``p * 100``, read_hurricane_ser_0.f90:35) when
    `pascal`, else in hPa as stored."""
    src = case.hurricane
    if isinstance(src, str):
        yield from read_hurricane_ser(src, case.maxnod, pascal)
        return
    for (u, v, p) in src:
        p = np.asarray(p, dtype=np.float64)
        yield (np.asarray(u, dtype=np.float64), np.asarray(v, dtype=np.float64),
               p * 100.0 if pascal else p)


# --------------------------------------------------------------------------- project

@dataclass
class Project:
    """The input files of one GOM project, as the reference reads them."""
    root: str
    main: MainInp
    nodes: np.ndarray                       # (P, 3): x, y, depth as in node.inp
    cells: np.ndarray                       # (E, 4)
    salt_init: Optional[np.ndarray] = None
    temp_init: Optional[np.ndarray] = None
    salt_series: List[Series] = field(default_factory=list)
    temp_series: List[Series] = field(default_factory=list)
    eta_series: List[Series] = field(default_factory=list)
    q_series: List[Series] = field(default_factory=list)
    harmonics: list = field(default_factory=list)
    restart: Optional[Dict[str, np.ndarray]] = None
    present: List[str] = field(default_factory=list)    # input files of the project


def read_project(root: str) -> Project:
    """Read ``root/input`` the way the reference's start-up does (read_main_inp.f90 and the
    readers it calls, read_node_inp.f90, read_cell_inp.f90): the series and initial
    fields a card switches on are read, with the file each shape names."""
    inp = os.path.join(root, "input")
    m = read_main_inp(os.path.join(inp, "main.inp"))
    if not m.get("complete"):
        raise FormatError("main.inp ends before card C46")
    pr = Project(root=root, main=m,
                 nodes=read_node_inp(os.path.join(inp, "node.inp"), m.maxnod),
                 cells=read_cell_inp(os.path.join(inp, "cell.inp"), m.maxele),
                 present=sorted(os.listdir(inp)))
    L = m.maxlayer
    if m.transport_flag == 1:
        for i, tr in enumerate(("salt", "temp")):
            if m.is_tran[i] != 1:
                continue
            if m.restart_in == 0:
                setattr(pr, f"{tr}_init", read_init(os.path.join(inp, f"{tr}_init.inp"), m.maxnod, L))
            if m.num_tran_ser[i] > 0:
                shp = m.tran_ser_shape[i]
                setattr(pr, f"{tr}_series",
                        read_series(os.path.join(inp, f"{tr}_ser{shp}.inp"), m.num_tran_ser[i], shp))
    if m.num_ob_cell > 0:
        if m.num_tide_bc > 0:
            pr.harmonics = read_harmonic_ser(os.path.join(inp, "harmonic_ser.inp"), m.num_harmonic_ser)
        if m.num_eta_bc > 0:
            shp = m.eta_ser_shape
            pr.eta_series = read_series(os.path.join(inp, f"eta_ser{shp}.inp"), m.num_eta_ser, shp)
    if m.num_Qb_cell > 0:
        shp = m.Q_ser_shape
        pr.q_series = read_series(os.path.join(inp, f"q_ser{shp}.inp"), m.num_Q_ser, shp,
                                  with_interp=(shp == 1))
    if m.restart_in == 1:
        from .gomout import read_restart
        pr.restart = read_restart(os.path.join(inp, "restart.inp"), m.maxele, m.maxnod, L)
    return pr


def unsupported(pr: Project) -> List[str]:
    """Options of a project that the Case (and so the device path and its host forcing)
    does not carry; each entry names the card and what it switches on."""
    m = pr.main
    out = []
    checks = [
        (m.coordinate_system != 1, "C1 coordinate_system 2 (lon/lat input)"),
        (m.ana_depth != 0, "C1 ana_depth (analytical depth)"),
        (m.voronoi != 0, "C1 voronoi cell centres"),
        (m.bf_varying != 0, "C5_02 bf_varying 1 (bottom_roughness.inp)"),
        (m.transport_flag == 1 and m.transport_solver not in (1, 2), "C7 transport_solver not 1 or 2"),
        (m.transport_flag == 1 and m.is_tran[1] == 1 and m.is_tran[0] == 0,
         "C7_2 temperature without salinity"),
        (m.transport_flag == 1 and m.is_tran[1] == 1,
         "C7_3 surface heat flux (heat_option with air_ser.inp)"),
        (any(r[3] == -1 for r in m.ob_rows), "C20_1 ob_eta_type -1 (radiated boundary)"),
        (m.num_WR_cell > 0, "C23 withdrawal/return cells"),
        (m.num_SS_cell > 0, "C24 source/sink cells"),
        (m.no_Etide_species > 0, "C25 equilibrium tide"),
        (m.wind_flag != 0 or m.airp_flag != 0, "C30 wind/air-pressure series"),
        (m.hurricane_flag != 0 and m.hurricane_data_type != 1,
         "C31 unformatted hurricane fields (hurricane_data_type 2)"),
        (m.hurricane_flag != 0 and m.hurricane_interp_method != 1,
         "C31 hurricane interpolation in space (hurricane_interp_method 2)"),
        (m.hurricane_flag != 0 and "hurricane_ser.inp" not in pr.present,
         "C31 hurricane fields: input/hurricane_ser.inp is missing"),
        (m.holland_flag != 0, "C32 Holland storm surge"),
        # read_q_ser.f90:46 leaves the last series' method in the global q_interp_method,
        # which calculate_Q.f90:21,46,92 branches on; only linear (1) is carried here
        (bool(pr.q_series) and pr.q_series[-1].interp != 1,
         "C22 river discharge interpolation other than linear (q_interp_method 2 or 3)"),
    ]
    for bad, what in checks:
        if bad:
            out.append(what)
    return out


def case_from_project(pr: Project, strict: bool = True) -> Tuple[Case, List[str]]:
    """Build the Case of a project read by read_project.

    Options the Case does not carry (``unsupported``) raise NotImplementedError when
    `strict`; otherwise they are switched off and listed in the second return value."""
    bad = unsupported(pr)
    if bad and strict:
        raise NotImplementedError("project options outside this package: " + "; ".join(bad))
    m = pr.main
    L = m.maxlayer
    c = Case(name=m.project_name, x=pr.nodes[:, 0].copy(), y=pr.nodes[:, 1].copy(),
             depth=pr.nodes[:, 2].copy(), cells=pr.cells.copy(), z_level=np.asarray(m.z_level, float))
    c.msl, c.dt, c.ndt = m.MSL, m.dt, m.ndt
    c.start = (m.start_year, m.start_month, m.start_day, m.start_hour, m.start_minute, m.start_second)
    c.data_start_year = None if m.data_start_year == m.start_year else m.data_start_year
    c.data_time_shift = m.data_time_shift
    c.gravity, c.rho_o, c.rho_a = m.gravity, m.rho_o, m.rho_a
    c.max_no_neighbor_node = m.max_no_neighbor_node
    c.h_node_adjust = m.h_node_adjust
    c.lat, c.lon = m.lat, m.lon
    c.restart_in, c.restart_out, c.restart_freq = m.restart_in, m.restart_out, m.restart_freq
    c.hot_start = pr.restart
    c.theta, c.advection_flag, c.adv_onoff_depth = m.theta, m.advection_flag, m.adv_onoff_depth
    c.elm_backtrace_flag, c.elm_sub_iter = m.ELM_backtrace_flag, m.ELM_sub_iter
    c.elm_min_iter, c.elm_max_iter = m.ELM_min_iter, m.ELM_max_iter
    c.bf_flag, c.manning, c.bf_height, c.von_karman = m.bf_flag, m.manning, m.bf_height, m.von_Karman
    c.baroclinic_flag, c.ana_density = m.baroclinic_flag, m.ana_density
    c.ref_salt, c.ref_temp, c.dry_depth = m.ref_salt, m.ref_temp, m.dry_depth
    c.coriolis_option = m.Coriolis_option
    c.smagorinsky, c.ah0, c.kh0, c.av0, c.kv0 = m.smagorinsky_parameter, m.Ah_0, m.Kh_0, m.Av_0, m.Kv_0
    c.max_iteration_pcg, c.error_tolerance = m.max_iteration_pcg, m.error_tolerance
    c.pcg_result_show, c.ishow_frequency = m.pcg_result_show, m.ishow_frequency
    c.transport_flag = m.transport_flag
    c.transport_solver = m.transport_solver if m.transport_flag == 1 else 1
    if m.transport_flag == 1:
        c.trans_sub_iter = m.get("trans_sub_iter", 1)
        c.h_flux_limiter = m.get("h_flux_limiter", 0)
        c.v_flux_limiter = m.get("v_flux_limiter", 0)
    c.is_salt, c.is_temp = m.is_tran
    c.salt_init, c.temp_init = pr.salt_init, pr.temp_init
    c.salt_series, c.temp_series = list(pr.salt_series), list(pr.temp_series)
    c.open_boundaries = [OpenBoundary(r[1], r[2], r[3], r[4], r[5], r[6], r[7]) for r in m.ob_rows]
    c.harmonics = list(pr.harmonics)
    c.eta_series = list(pr.eta_series)
    c.tide_spinup, c.tide_spinup_period = m.tide_spinup, m.tide_spinup_period
    c.baroclinic_spinup, c.baroclinic_spinup_period = m.baroclinic_spinup, m.baroclinic_spinup_period
    c.rivers = [RiverBoundary(r[1], r[2], r[3], r[4], r[5], r[6]) for r in m.q_rows]
    c.q_series = list(pr.q_series)
    c.wind_formular, c.wind_spinup, c.wind_spinup_period = m.wind_formular, m.wind_spinup, m.wind_spinup_period
    if m.hurricane_flag == 1 and not any(b.startswith("C31") for b in bad):
        c.hurricane_flag, c.hurricane_dt = 1, m.hurricane_dt
        c.hurricane_start, c.hurricane_end = m.hurricane_start, m.hurricane_end
        c.hurricane = os.path.join(pr.root, "input", "hurricane_ser.inp")
    if m.tser_station_num > 0:
        c.tser = TserSpec([(int(r[1]), int(r[2]), str(r[3])) for r in m.tser_stations], m.tser_hloc,
                          m.tser_time, m.tser_frequency, m.tser_format,
                          (m.tser_eta, m.tser_H, m.tser_u, m.tser_v, m.tser_salt, m.tser_temp, m.tser_airp))
    if m.IS2D_switch == 1:
        c.out2d = ContourSpec(m.IS2D_format, m.IS2D_File_freq, m.IS2D_time, m.IS2D_frequency, m.IS2D_start,
                              m.IS2D_end, m.IS2D_unit_conv, tuple(m.IS2D_variable))
    if m.IS3D_surf_switch == 1 or m.IS3D_full_switch == 1:
        c.out3d = ContourSpec(m.IS3D_format, m.IS3D_File_freq, m.IS3D_time, m.IS3D_frequency, m.IS3D_start,
                              m.IS3D_end, m.IS3D_unit_conv, tuple(m.IS3D_variable), m.IS3D_surf_switch,
                              m.IS3D_full_switch, m.IS3D_grid_format)
    for tag, attr in (("2D", "dump2d"), ("3D", "dump3d")):
        if m[f"IS{tag}_dump_switch"] == 1:
            setattr(c, attr, DumpSpec(m[f"IS{tag}_dump_binary"], m[f"IS{tag}_dump_time"],
                                      m[f"IS{tag}_dump_File_freq"], m[f"IS{tag}_dump_frequency"],
                                      m[f"IS{tag}_dump_start"], m[f"IS{tag}_dump_end"]))
    if bad:
        # switched off: temperature (its surface heat flux needs the air series) and
        # radiated boundaries; the other listed options have no Case field at all
        if c.is_temp:
            c.is_temp, c.temp_init, c.temp_series = 0, None, []
            for b in c.open_boundaries:
                b.temp_ser_id = 0
            for r in c.rivers:
                r.temp_ser_id = 0
        c.open_boundaries = [b for b in c.open_boundaries if b.eta_type != -1]
    if L > 0 and c.salt_init is not None and c.salt_init.shape != (c.maxnod, L):
        raise FormatError("salt_init.inp does not hold maxlayer values per node")
    return c, bad


def read_case(root: str, strict: bool = True) -> Case:
    """read_project followed by case_from_project; returns the Case."""
    return case_from_project(read_project(root), strict)[0]
