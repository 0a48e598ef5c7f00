"""GOM on-disk input formats: writer for a project's ``input/`` directory.

The files follow the formats the reference's readers expect (reference files, relative
to ``GOM_Rv1.1.0/source``):

* ``main.inp``  card deck C00-C46 in the order of ``read_main_inp.f90:75-1217``; lines
  starting with ``!``, ``C`` or ``c`` in column 1 are comments (``read_main_inp.f90:67``),
  cards are located by ``seek_card.f90`` on the first five characters of a line.
* ``node.inp``  count, then ``id x y depth`` (``read_node_inp.f90:50-87``).
* ``cell.inp``  count, then ``id 3|4 n1 n2 n3 [n4]``, counter-clockwise
  (``read_cell_inp.f90:39-87``).
* ``harmonic_ser.inp``  ``id ncon phase_shift`` then ``period[h] amp phase f V0 name``
  (``read_harmonic_ser.f90:20-26``).
* ``salt_init.inp`` / ``temp_init.inp``  ``id`` and ``maxlayer`` values per node,
  bottom layer first (``read_salt_init.f90:25-27``).
* ``salt_ser1.inp`` / ``temp_ser1.inp`` / ``q_ser1.inp``  per series
  ``id n time_conv time_adjust unit_conv adjust [interp]`` then ``n`` lines ``time value``
  (``read_salt_ser.f90:24-31``, ``read_q_ser.f90:24-32``).

Every value is written with ``repr`` so that the reader recovers the exact binary64.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np


@dataclass
class Series:
    """A piecewise-linear time series; times in days (written with time_conv = 86400)."""
    times_days: Sequence[float]
    values: Sequence[float]


@dataclass
class OpenBoundary:
    """One open-boundary cell face given by its two nodes in the cell's CCW order."""
    n1: int
    n2: int
    eta_type: int = 1          # 1 = harmonic, 2 = time series (v1.1.0 meaning)
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
class Case:
    """Everything a GOM project needs for the synthetic configurations."""
    name: str
    x: np.ndarray                  # node x [m]
    y: np.ndarray                  # node y [m]
    depth: np.ndarray              # node depth below MSL [m]
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
    transport_flag: int = 1
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
    tide_spinup: int = 0
    tide_spinup_period: float = 0.5
    baroclinic_spinup: int = 0
    baroclinic_spinup_period: float = 0.5
    rivers: List[RiverBoundary] = field(default_factory=list)
    q_series: List[Series] = field(default_factory=list)
    start: Tuple[int, int, int, int, int, int] = (2020, 1, 1, 0, 0, 0)
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


def _r(v: float) -> str:
    return repr(float(v))


def _write_series_file(path: str, series: Sequence[Series], with_interp: bool) -> None:
    with open(path, "w") as f:
        f.write("! generated\n")
        for i, s in enumerate(series, 1):
            extra = "  1" if with_interp else ""
            f.write(f"  {i} {len(s.times_days)} 86400.0 0.0 1.0 0.0{extra}\n")
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
    data(9.81, 1025.0, 1.29, 10)
    card("C1\tmaxnod maxele h_node_adjust ana_depth coordinate_system lon lat voronoi node_mirr cell_mirr")
    data(c.maxnod, c.maxele, 0.0, 0, 1, c.lon, c.lat, 0, 0, 0)
    card("C2\tmaxlayer MSL z_level(0)")
    data(L, c.msl, float(c.z_level[0]))
    card("C2_1\tserial_num z_level (surface first)")
    for k in range(L, 0, -1):
        data(k, float(c.z_level[k]))
    card("C3\tdt ndt data_start_year data_time_shift start_year month day hour minute second")
    y, mo, d, h, mi, s = c.start
    data(c.dt, c.ndt, y, 0, y, mo, d, h, mi, s)
    card("C4\trestart_in restart_out restart_freq ishow ishow_frequency")
    data(0, 0, 99999999, 0, 99999999)
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
    data(c.max_iteration_pcg, c.error_tolerance, 0)
    card("C7\ttransport_flag transport_solver")
    data(c.transport_flag, 1)
    if c.transport_flag == 1:
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
    data(neta, 0, 0, 1)
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
    data(0, 0, 0, 2, 0, 0.0)
    card("C31\thurricane_flag hurricane_data_type hurricane_interp_method hurricane_dt")
    data(0, 1, 1, 21600)
    card("C32\tholland_flag holland_start_jday holland_end_jday")
    data(0, 0.0, 0.0)
    card("C40\tcheck_grid")
    data(0, 0, 0, 1, 0, 0.001)
    card("C41\ttser")
    data(0, 0, 1, 4, 24, 1)
    card("C42\t2D output")
    data(0, 0)
    data(1, 0, 99999999)
    data(4, 99999999, 0, 99999999, 0.001)
    card("C42_1\tvariables")
    for nm in ("eta", "u", "v", "salt", "temp", "rho"):
        data(nm, 0)
    card("C43\t3D output")
    data(0, 1, 0)
    data(1, 0, 99999999)
    data(4, 99999999, 0, 99999999, 0.001)
    card("C43_1\tvariables")
    for nm in ("u", "v", "w", "salt", "temp", "rho"):
        data(nm, 0)
    card("C44\t2D dump")
    data(0, 0, 4, 99999999)
    data(99999999, 1, 99999999)
    card("C45\t3D dump")
    data(0, 0, 4, 99999999)
    data(99999999, 1, 99999999)
    card("C46\tdiagnostics")
    data(0, 0, 0, 0)
    data(0, 0, 0)
    return "\n".join(lines) + "\n"
