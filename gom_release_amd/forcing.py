"""Host-side forcing of one step: the 2-D boundary values the reference computes on the
host before the hot path (GOM_Rv1.1.0/source/main.f90:84-161) and the per-sub-cycle
boundary concentrations of the TVD transport.

These stay on the host by design (O(boundary) work per step); their results are handed to
the device through ``gomgpu_set_forcing``.
"""
from __future__ import annotations

import math
from typing import Optional, Sequence

import numpy as np

from .gomio import Case, Series

DEG2RAD = math.pi / 180.0


def series_value(s: Series, u2: float, ref_diff_days: int = 0) -> Optional[float]:
    """Linear interpolation of a series at time u2 [s], as the reference does it
    (solve_transport_equation_v19.f90:245-258, compute_boundary_eta.f90:96-108): the first
    interval whose ends bracket u2; None when no interval does."""
    t = s.seconds()
    v = s.scaled()
    shift = ref_diff_days * 86400.0
    for i in range(1, len(t)):
        u1 = t[i - 1] - shift
        u3 = t[i] - shift
        if u1 <= u2 <= u3:
            v1, v3 = v[i - 1], v[i]
            return (u2 - u1) * (v3 - v1) / (u3 - u1) + v1
    return None


def spinup(flag: int, period_days: float, elapsed: float) -> float:
    """setup_spinup.f90:10-54."""
    if flag == 1:
        per = period_days * 86400.0
        if elapsed < per:
            return math.tanh(2.0 * elapsed / per)
    return 1.0


def julian(y: int, mo: int, d: int, h: int = 0, mi: int = 0, s: int = 0) -> float:
    """Day of the year, 0-based and fractional (julian.f90:15-57)."""
    cum = [0, 31, 59, 90, 120, 151, 181, 212, 243, 273, 304, 334][mo - 1]
    jd = float(cum)
    if y % 4 == 0 and y != 1900 and mo >= 3:
        jd = jd + 1
    return jd + (d - 1) + h / 24.0 + mi / 1440.0 + s / 86400.0


def julian_day0(case: Case) -> float:
    """Day-of-year (0-based, fractional) of the start time, julian.f90."""
    return julian(*case.start)


def _year_days(y1: int, y2: int) -> int:
    """Days from the start of year y1 to that of y2 as find_hurricane_start_end_jday.f90:
    17-30 counts them (364 for a year divisible by 4)."""
    d = 0
    while y1 != y2:
        d += 365 if y1 % 4 != 0 else 364
        y1 += 1
    return d


def hurricane_window(case: Case):
    """(hurricane_start_jday, hurricane_end_jday) of find_hurricane_start_end_jday.f90,
    on the model's julian-day axis."""
    y0 = case.start[0]
    a = case.hurricane_start
    b = case.hurricane_end
    return (_year_days(y0, a[0]) + julian(a[0], a[1], a[2]),
            _year_days(y0, b[0]) + julian(b[0], b[1], b[2]))


def hurricane_phase(it: int, julian_day: float, window, read_interval: int) -> int:
    """What main.f90:116-123 and read_hurricane_ser_1.f90:16-17 do at step `it`: 0 outside
    the hurricane window, 2 when the step reads the next snapshot (mod(it, interval) = 0),
    else 1 (interpolate)."""
    if not (window[0] <= julian_day <= window[1]):
        return 0
    return 2 if it % read_interval == 0 else 1


def boundary_eta(case: Case, it: int, sp_tide: float, julian_day: Optional[float] = None):
    """eta_at_ob_old/new for step `it` (compute_boundary_eta.f90:72-124): harmonic
    boundaries (type 1) and elevation series (type 2, at julian_day*86400 - dt and
    julian_day*86400). A series that does not cover the time keeps the value of the
    boundary before it in the same call, as the reference's local variable does."""
    dt = case.dt
    elapsed = it * dt
    nob = len(case.open_boundaries)
    old = np.zeros(max(nob, 1))
    new = np.zeros(max(nob, 1))
    rdd = case.reference_diff_days()
    v2_old = v2_new = 0.0
    memo = {}
    tides = {}  # boundaries share harmonic sets: one sum per set
    for i, b in enumerate(case.open_boundaries):
        if b.eta_type == 1:
            if b.harmonic_id not in tides:
                shift, cons = case.harmonics[b.harmonic_id - 1]
                o = 0.0
                n = 0.0
                for (per, amp, ph, fac, v0, _nm) in cons:
                    w = 2.0 * math.pi / (per * 3600.0)
                    o = o + sp_tide * amp * fac * math.cos(w * (elapsed - dt) + v0 * DEG2RAD - ph * DEG2RAD
                                                           - shift * DEG2RAD)
                    n = n + sp_tide * amp * fac * math.cos(w * elapsed + v0 * DEG2RAD - ph * DEG2RAD - shift * DEG2RAD)
                tides[b.harmonic_id] = (o, n)
            old[i], new[i] = tides[b.harmonic_id]
        elif b.eta_type == 2:
            if julian_day is None:
                raise ValueError("elevation-series boundaries need the julian day")
            if b.eta_ser_id not in memo:
                ser = case.eta_series[b.eta_ser_id - 1]
                memo[b.eta_ser_id] = (series_value(ser, julian_day * 86400.0 - dt, rdd),
                                      series_value(ser, julian_day * 86400.0, rdd))
            vo, vn = memo[b.eta_ser_id]
            v2_old = v2_old if vo is None else vo
            v2_new = v2_new if vn is None else vn
            old[i] = sp_tide * v2_old
            new[i] = sp_tide * v2_new
        else:
            raise NotImplementedError("ob_eta_type -1 (elevation radiated from interior cells) "
                                      "needs the device state on the host each step")
    return old, new


def tvd_boundary_concentrations(case: Case, julian_day: float, sp_baroclinic: float,
                                ntran: int, tran_ids: Sequence[int]):
    """con_at_ob (Nt, ntran, OB) and con_at_Qb (Nt, ntran, QB), Fortran order, for the
    sub-cycles of one step (solve_transport_equation_v19.f90:229-397).

    Series times are shifted by reference_diff_days (:249-250, :270-271, and the river
    block). The interpolated value is the routine's local v2, which carries over between
    boundaries, tracers and sub-cycles in the reference's loop order (sub-cycle, then open
    boundary, then tracer; then river cell, then tracer): a series that does not bracket
    the time leaves the previous v2 in place, 0 at the start of the call (the reference's
    is undefined there). A boundary without a series keeps the routine's automatic array
    value, which is never set in the call: 0."""
    nt = max(case.trans_sub_iter, 1)
    dt2 = case.dt / nt
    nob = max(len(case.open_boundaries), 1)
    nqb = max(len(case.rivers), 1)
    ob = np.zeros((nt, max(ntran, 1), nob))
    qb = np.zeros((nt, max(ntran, 1), nqb))
    rdd = case.reference_diff_days()
    memo = {}   # boundaries share series: one interpolation per (tracer, series, sub-cycle)

    def interp(tid, sid, u2):
        key = (tid, sid, u2)
        if key not in memo:
            ser = case.salt_series if tid == 1 else case.temp_series
            memo[key] = series_value(ser[sid - 1], u2, rdd)
        return memo[key]

    v2 = 0.0
    for sub in range(nt):
        u2 = (julian_day * 86400.0 - case.dt) + (sub + 1) * dt2
        for i, b in enumerate(case.open_boundaries):
            for t in range(ntran):
                tid = tran_ids[t]
                sid = b.salt_ser_id if tid == 1 else (b.temp_ser_id if tid == 2 else 0)
                if sid > 0:
                    v = interp(tid, sid, u2)
                    v2 = v2 if v is None else v
                    ob[sub, t, i] = sp_baroclinic * v2
        for i, r in enumerate(case.rivers):
            for t in range(ntran):
                tid = tran_ids[t]
                sid = r.salt_ser_id if tid == 1 else (r.temp_ser_id if tid == 2 else 0)
                if sid > 0:
                    v = interp(tid, sid, u2)
                    v2 = v2 if v is None else v
                    qb[sub, t, i] = sp_baroclinic * v2
    return ob.ravel(order="F"), qb.ravel(order="F")


def elm_boundary_concentrations(case: Case, julian_day: float, sp_baroclinic: float, qb_prev=None):
    """The boundary values of the ELM transport (find_openboundary_salt_temp_v3.f90:17-97),
    as con_at_ob (1, 2, OB) and con_at_Qb (1, 2, QB) in Fortran order: salt in tracer slot 1, temperature in slot 2.

    An open boundary takes its series value (reference_diff_days applied) or ref_salt /
    ref_temp when it has none; a river cell takes its series value only when it has one and
    otherwise keeps its previous value (`qb_prev`, (2, QB), zero at the start). A series that
    does not cover the time keeps the value of the boundary before it in the same call."""
    nt = 1      # the ELM solver reads no sub-cycle count: the arrays keep extent 1
    nob = max(len(case.open_boundaries), 1)
    nqb = max(len(case.rivers), 1)
    ob = np.zeros((nt, 2, nob))
    qb = np.zeros((nt, 2, nqb))
    u2 = julian_day * 86400.0
    rdd = case.reference_diff_days()
    v2 = 0.0
    for i, b in enumerate(case.open_boundaries):
        for t, (on, sid, ser, ref) in enumerate(((case.is_salt, b.salt_ser_id, case.salt_series, case.ref_salt),
                                                 (case.is_temp, b.temp_ser_id, case.temp_series, case.ref_temp))):
            if on == 1 and sid > 0:
                v = series_value(ser[sid - 1], u2, rdd)
                v2 = v2 if v is None else v
            else:
                v2 = ref
            ob[0, t, i] = sp_baroclinic * v2
    if qb_prev is not None:
        qb[0] = np.asarray(qb_prev, dtype=np.float64).reshape(2, nqb)
    for i, r in enumerate(case.rivers):
        for t, (on, sid, ser) in enumerate(((case.is_salt, r.salt_ser_id, case.salt_series),
                                            (case.is_temp, r.temp_ser_id, case.temp_series))):
            if on == 1 and sid > 0:
                v = series_value(ser[sid - 1], u2, rdd)
                v2 = v2 if v is None else v
                qb[0, t, i] = sp_baroclinic * v2
    return ob.ravel(order="F"), qb.ravel(order="F")


def wind_stress(mesh, u: float, v: float, rho_a: float = 1.29, rho_o: float = 1025.0,
                formula: int = 2, spinup_wind: float = 1.0):
    """Face wind stress for a uniform wind (calculate_wind_stress.f90:19-66, wind_flag 1).

    Returns (Gamma_T, wind_stress_normal, wind_stress_tangnt) per face.
    """
    f32 = np.float32
    rho_ratio = rho_a / rho_o
    coef = {1: (0.75, 0.067, 4.0), 2: (0.61, 0.063, 6.0), 3: (0.80, 0.065, 1.0)}[formula]
    # Cda_min/max: the inner sums are default-real (single precision) in the reference
    cmin = 0.001 * float(f32(coef[0]) + f32(coef[1]) * f32(coef[2]))
    cmax = 0.001 * float(f32(coef[0]) + f32(coef[1]) * f32(50.0))
    ct = mesh["cos_theta"]
    st = mesh["sin_theta"]
    ws = math.sqrt(u * u + v * v)
    wn = u * ct + v * st
    wt = -u * st + v * ct
    cda = 0.001 * (coef[0] + coef[1] * ws)
    cda = min(max(cda, cmin), cmax)
    gt = cda * rho_ratio * ws
    n = ct.shape[0]
    return np.full(n, gt), gt * wn * spinup_wind, gt * wt * spinup_wind


def river_discharge(case: Case, julian_day: float, prev=None):
    """Q_add per river cell (calculate_Q.f90:21-45, q_interp_method 1): the series at
    julian_day shifted by reference_diff_days, times Q_portion. Q_add is a module array:
    a cell whose series does not bracket the time keeps its previous value (`prev`, zero
    before the first call, as allocate_variables leaves it)."""
    q = np.zeros(max(len(case.rivers), 1))
    if prev is not None:
        q[:] = np.asarray(prev, dtype=np.float64)[:q.size]
    u2 = julian_day * 86400.0
    rdd = case.reference_diff_days()
    for i, r in enumerate(case.rivers):
        v = series_value(case.q_series[r.q_ser_id - 1], u2, rdd)
        if v is not None:
            q[i] = v * r.portion
    return q


def boundary_tables(case: Case, params=None) -> dict:
    """The tables of the device boundary forcing (gomgpu_boundary of include/gomgpu.h): the
    case's harmonic sets, every series (elevation, salinity, temperature, discharge, in that
    order, in one table, as the times and scaled values series_value reads), and which set or
    series each open boundary and river cell reads. With them the device computes what
    boundary_eta, tvd_boundary_concentrations / elm_boundary_concentrations and
    river_discharge compute here, from the step's scalars alone."""
    first, shift, cons = [0], [], []
    for sh, cn in case.harmonics:
        shift.append(sh)
        for (per, amp, ph, fac, v0, _nm) in cn:
            cons.extend([per, amp, fac, v0, ph])
        first.append(len(cons) // 5)
    groups = (case.eta_series, case.salt_series, case.temp_series, case.q_series)
    base = np.cumsum([0] + [len(g) for g in groups])
    sfirst, stime, sval = [0], [], []
    for g in groups:
        for s in g:
            stime.extend(s.seconds())
            sval.extend(s.scaled())
            sfirst.append(len(stime))

    def sid(group, i):
        return int(base[group] + i) if i > 0 else 0

    obs, rivs = case.open_boundaries, case.rivers
    for b in obs:
        if b.eta_type not in (1, 2):
            raise NotImplementedError("ob_eta_type -1 (elevation radiated from interior cells) needs the device "
                                      "state on the host each step")
    return dict(
        num_harmonic_sets=len(case.harmonics), harmonic_first=np.array(first, np.int32),
        tidal_phase_shift=np.array(shift, np.float64), constituents=np.array(cons, np.float64),
        num_series=len(sfirst) - 1, series_first=np.array(sfirst, np.int32),
        series_time=np.array(stime, np.float64), series_value=np.array(sval, np.float64),
        num_ob=len(obs),
        ob_eta_type=np.array([b.eta_type for b in obs], np.int32),
        ob_eta_id=np.array([b.harmonic_id if b.eta_type == 1 else sid(0, b.eta_ser_id) for b in obs], np.int32),
        ob_tracer_series=np.array([[sid(1, b.salt_ser_id), sid(2, b.temp_ser_id)] for b in obs], np.int32).ravel(),
        num_qb=len(rivs),
        qb_q_series=np.array([sid(3, r.q_ser_id) for r in rivs], np.int32),
        qb_portion=np.array([r.portion for r in rivs], np.float64),
        qb_tracer_series=np.array([[sid(1, r.salt_ser_id), sid(2, r.temp_ser_id)] for r in rivs], np.int32).ravel(),
        is_salt=int(case.is_salt), is_temp=int(case.is_temp), ref_salt=float(case.ref_salt),
        ref_temp=float(case.ref_temp), reference_diff_days=int(case.reference_diff_days()))
