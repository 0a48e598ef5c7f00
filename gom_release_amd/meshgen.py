"""Synthetic meshes and cases for the benchmark configurations.

* ``channel_case``  rectangular channel of counter-clockwise quads, uniform depth, M2 tide
  on the west open boundary, other sides closed (the 64x32, 200 m, 10-layer configuration
  and smaller variants of it).
* ``estuary_case``  mixed orthogonal mesh: rows of 100x100 m squares stitched to rows of
  equilateral 100 m triangles, Gaussian thalweg bathymetry with seeded noise, salinity
  ramp from river to sea, river inflow at the east end, M2+K1 tide at the west end.

Cells are written counter-clockwise; node and cell ids are 1-based as in the reference.
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

from .gomio import Case, OpenBoundary, RiverBoundary, Series

M2_HOURS = 12.42
K1_HOURS = 23.93


def uniform_levels(depth: float, nlayer: int, top: float, surf_gap: float) -> np.ndarray:
    """z_level(0..L): L-1 equal layers from -depth up to -surf_gap, then a top layer to `top`."""
    z = np.empty(nlayer + 1)
    z[0] = -depth
    lower = np.linspace(-depth, -surf_gap, nlayer)
    z[:nlayer] = lower
    z[nlayer] = top
    return z


def channel_case(nx: int = 64, ny: int = 32, dx: float = 200.0, depth: float = 5.0,
                 nlayer: int = 10, amp: float = 0.5, dt: float = 60.0, ndt: int = 1440,
                 salt: float = 30.0, name: str = "channel", **kw) -> Case:
    """Rectangular channel of nx x ny quads; the west side (x = 0) is an open boundary."""
    ii, jj = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing="xy")
    x = (ii * dx).astype(np.float64).ravel()
    y = (jj * dx).astype(np.float64).ravel()
    nid = lambda i, j: j * (nx + 1) + i + 1          # noqa: E731
    cells = np.zeros((nx * ny, 4), dtype=np.int64)
    c = 0
    for j in range(ny):
        for i in range(nx):
            cells[c] = (nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1))
            c += 1
    P = x.shape[0]
    z = uniform_levels(depth, nlayer, top=max(3.0, 4 * amp), surf_gap=max(0.4, 2.0 * amp + 0.4))
    obs = [OpenBoundary(nid(0, j + 1), nid(0, j), eta_type=1, harmonic_id=1, salt_ser_id=1)
           for j in range(ny)]
    case = Case(name=name, x=x, y=y, depth=np.full(P, depth), cells=cells, z_level=z,
                dt=dt, ndt=ndt,
                salt_init=np.full((P, nlayer), salt),
                salt_series=[Series([-1000.0, 100000.0], [salt, salt])],
                open_boundaries=obs,
                harmonics=[(0.0, [(M2_HOURS, amp, 0.0, 1.0, 0.0, "M2")])])
    for k, v in kw.items():
        setattr(case, k, v)
    return case


def estuary_mesh(ncol: int, nrep: int, pattern: str = "STT", dx: float = 100.0):
    """Rows stacked in y following `pattern` repeated `nrep` times.

    'S' is a row of ncol squares (dx x dx); 'T' a row of equilateral triangles of side dx.
    Triangle rows alternate the x offset of their node rows; rows touching the x = 0 and
    x = ncol*dx sides are closed by right triangles so the domain stays rectangular.
    Returns (x, y, cells) with 1-based node ids.
    """
    h_tri = dx * math.sqrt(3.0) / 2.0
    rows = pattern * nrep
    xs: List[float] = []
    ys: List[float] = []
    cells: List[Tuple[int, int, int, int]] = []

    def add_row(yv: float, offset: bool) -> List[int]:
        ids = []
        if offset:
            pts = [0.0] + [dx * (i + 0.5) for i in range(ncol)] + [dx * ncol]
        else:
            pts = [dx * i for i in range(ncol + 1)]
        for px in pts:
            xs.append(px)
            ys.append(yv)
            ids.append(len(xs))
        return ids

    yv = 0.0
    offset = False
    cur = add_row(yv, offset)
    for r in rows:
        if r == "S":
            if offset:
                raise ValueError("square row needs an aligned node row below it")
            yv += dx
            nxt = add_row(yv, False)
            for i in range(ncol):
                cells.append((cur[i], cur[i + 1], nxt[i + 1], nxt[i]))
            cur = nxt
        else:
            yv += h_tri
            nxt = add_row(yv, not offset)
            if not offset:
                # lower row aligned (ncol+1 pts), upper offset (ncol+2 pts incl. side pts)
                up = nxt
                cells.append((cur[0], up[1], up[0], 0))
                for i in range(ncol):
                    cells.append((cur[i], cur[i + 1], up[i + 1], 0))
                    if i + 1 < ncol:
                        cells.append((cur[i + 1], up[i + 2], up[i + 1], 0))
                cells.append((cur[ncol], up[ncol + 1], up[ncol], 0))
            else:
                lo = cur
                up = nxt
                cells.append((lo[0], lo[1], up[0], 0))
                for i in range(ncol):
                    cells.append((lo[i + 1], up[i + 1], up[i], 0))
                    cells.append((lo[i + 1], lo[i + 2], up[i + 1], 0))
            cur = nxt
            offset = not offset
    return np.array(xs), np.array(ys), np.array(cells, dtype=np.int64)


def estuary_case(ncol: int = 40, nrep: int = 4, nlayer: int = 5, dx: float = 100.0,
                 depth_axis: float = 20.0, depth_bank: float = 2.0, noise: float = 0.5,
                 seed: int = 1, river_q: float = 500.0, wind: float = 0.0,
                 dt: float = 120.0, ndt: int = 10, dry_land: bool = False,
                 tide_phase: float = 0.0,
                 name: str = "estuary", **kw) -> Case:
    """Synthetic estuary on the stitched square/triangle mesh (axis along x)."""
    x, y, cells = estuary_mesh(ncol, nrep, dx=dx)
    rng = np.random.default_rng(seed)
    ymid = 0.5 * (y.min() + y.max())
    width = y.max() - y.min()
    sig = width / 5.0
    depth = depth_bank + (depth_axis - depth_bank) * np.exp(-((y - ymid) / sig) ** 2)
    depth = depth + rng.uniform(-noise, noise, size=depth.shape)
    if dry_land:
        # intertidal banks: beyond 0.3 of the width from the axis the bed rises linearly
        # from 1 m depth to 2 cm depth at the sides; with a falling tide they dry out
        rel = np.abs(y - ymid) / (0.5 * width)
        bank = rel > 0.6
        depth[bank] = 1.0 - (rel[bank] - 0.6) / 0.4 * 0.98
    depth = np.round(depth, 6)
    zb = float(np.ceil(depth.max() + 0.5))
    z = uniform_levels(zb, nlayer, top=4.0, surf_gap=1.5)
    P = x.shape[0]
    L = nlayer
    xr = (x - x.min()) / (x.max() - x.min())
    s_node = 35.0 * (1.0 - xr)
    salt_init = np.repeat(np.round(s_node, 6)[:, None], L, axis=1)
    # open boundary at x = 0 (west): faces on the x=0 side in their cell's CCW order
    obs: List[OpenBoundary] = []
    rivers: List[RiverBoundary] = []
    xmax = x.max()
    west: List[Tuple[int, int]] = []
    east: List[Tuple[int, int]] = []
    for c in cells:
        n = 4 if c[3] else 3
        for l in range(n):
            a, b = int(c[l]), int(c[(l + 1) % n])
            if x[a - 1] == 0.0 and x[b - 1] == 0.0:
                west.append((a, b))
            if x[a - 1] == xmax and x[b - 1] == xmax:
                east.append((a, b))
    for (a, b) in west:
        obs.append(OpenBoundary(a, b, eta_type=1, harmonic_id=1, salt_ser_id=1))
    # river: the east-side faces nearest to the axis (3 faces)
    east.sort(key=lambda ab: abs(0.5 * (y[ab[0] - 1] + y[ab[1] - 1]) - ymid))
    nriv = min(3, len(east))
    for (a, b) in east[:nriv]:
        rivers.append(RiverBoundary(a, b, q_ser_id=1, portion=1.0 / nriv, salt_ser_id=2))
    case = Case(name=name, x=x, y=y, depth=depth, cells=cells, z_level=z, dt=dt, ndt=ndt,
                salt_init=salt_init,
                salt_series=[Series([-1000.0, 100000.0], [35.0, 35.0]),
                             Series([-1000.0, 100000.0], [0.0, 0.0])],
                open_boundaries=obs,
                harmonics=[(0.0, [(M2_HOURS, 0.5, tide_phase, 1.0, 0.0, "M2"),
                                  (K1_HOURS, 0.2, tide_phase + 30.0, 1.0, 0.0, "K1")])],
                rivers=rivers,
                q_series=[Series([-1000.0, 100000.0], [river_q, river_q])] if rivers else [],
                trans_sub_iter=2, h_flux_limiter=1, v_flux_limiter=1)
    for k, v in kw.items():
        setattr(case, k, v)
    return case


class HollandTrack:
    """Node snapshots (u [m/s], v [m/s], p [hPa]) of a translating Holland (1980) cyclone,
    one per time in `times` [s], generated as they are iterated (hurricane_ser.inp order).

    Holland, G. J. (1980), Mon. Wea. Rev. 108, 1212-1218: with dp = p_env - p_min and
    s = (r_max / r)^b, the surface pressure is p = p_min + dp exp(-s) and the gradient wind
    V = sqrt(b dp s exp(-s) / rho_air + (r f / 2)^2) - r f / 2. The surface wind is
    `surface_factor` V, counter-clockwise about the centre (northern hemisphere), plus the
    translation velocity of the centre. The centre starts at `centre0` [m] and moves at
    `velocity` [m/s]."""

    def __init__(self, x, y, times: Sequence[float], centre0: Tuple[float, float],
                 velocity: Tuple[float, float], p_min: float = 950.0, p_env: float = 1013.0,
                 r_max: float = 40.0e3, b: float = 1.5, lat: float = 30.0, rho_air: float = 1.15,
                 surface_factor: float = 0.8):
        self.x = np.asarray(x, dtype=np.float64)
        self.y = np.asarray(y, dtype=np.float64)
        self.times = [float(t) for t in times]
        self.centre0, self.velocity = tuple(centre0), tuple(velocity)
        self.p_min, self.p_env, self.r_max, self.b = p_min, p_env, r_max, b
        self.lat, self.rho_air, self.surface_factor = lat, rho_air, surface_factor

    def __len__(self) -> int:
        return len(self.times)

    def coriolis(self) -> float:
        return float(2.0 * 7.292e-5 * np.sin(np.radians(self.lat)))

    def centre(self, t: float):
        return self.centre0[0] + self.velocity[0] * t, self.centre0[1] + self.velocity[1] * t

    def at(self, t: float):
        f = self.coriolis()
        dp = (self.p_env - self.p_min) * 100.0               # Pa
        cx, cy = self.centre(t)
        dx, dy = self.x - cx, self.y - cy
        r = np.hypot(dx, dy)
        rs = np.maximum(r, 1.0)
        s = (self.r_max / rs) ** self.b
        es = np.exp(-s)
        p = self.p_min + (self.p_env - self.p_min) * es      # hPa
        vg = np.sqrt(self.b * dp * s * es / self.rho_air + (rs * f / 2.0) ** 2) - rs * f / 2.0
        vs = np.where(r > 1.0, self.surface_factor * vg, 0.0)
        u = -vs * dy / rs + self.velocity[0]
        v = vs * dx / rs + self.velocity[1]
        return u, v, p

    def __iter__(self):
        for t in self.times:
            yield self.at(t)


def holland_snapshots(x, y, times: Sequence[float], centre0: Tuple[float, float],
                      velocity: Tuple[float, float], p_min: float = 950.0, p_env: float = 1013.0,
                      r_max: float = 40.0e3, b: float = 1.5, lat: float = 30.0, rho_air: float = 1.15,
                      surface_factor: float = 0.8) -> HollandTrack:
    """The HollandTrack of these parameters (snapshots generated when iterated)."""
    return HollandTrack(x, y, times, centre0, velocity, p_min, p_env, r_max, b, lat, rho_air, surface_factor)


def add_cyclone(case: Case, steps: int, seed: int = 0, p_min: float = 950.0, r_max: float = 40.0e3,
                b: float = 1.5, speed: float = 5.0, hurricane_dt: Optional[int] = None) -> Case:
    """Give a case Holland-cyclone forcing for `steps` steps: hurricane fields every
    `hurricane_dt` seconds (default 10 steps) from the model start, a track entering from a
    seeded side and crossing the domain at `speed` m/s; the uniform wind is replaced."""
    rng = np.random.default_rng(seed)
    cx, cy = 0.5 * (case.x.min() + case.x.max()), 0.5 * (case.y.min() + case.y.max())
    ang = rng.uniform(0.0, 2.0 * np.pi)
    heading = ang + np.pi + rng.uniform(-np.pi / 6, np.pi / 6)
    start = (cx + 1.5 * r_max * np.cos(ang), cy + 1.5 * r_max * np.sin(ang))
    vel = (speed * np.cos(heading), speed * np.sin(heading))
    hdt = int(hurricane_dt if hurricane_dt is not None else 10 * case.dt)
    times = [k * float(hdt) for k in range(int(steps * case.dt // hdt) + 3)]
    case.hurricane_flag, case.hurricane_dt = 1, hdt
    y, mo, d = case.start[:3]
    case.hurricane_start, case.hurricane_end = (y, mo, d), (y + 1, mo, d)
    case.hurricane = holland_snapshots(case.x, case.y, times, start, vel, p_min=p_min, r_max=r_max, b=b,
                                       lat=case.lat)
    return case


def fractal_field(x, y, seed: int, beta: float = 3.4, n: int = 512) -> np.ndarray:
    """A seeded fractal surface sampled at the points (x, y), scaled to [0, 1]: spectral
    synthesis on an n x n periodic grid spanning the points' box (amplitude |k|^(-beta/2)
    with uniform random phases, i.e. a power spectrum |k|^-beta; beta = 3.4 is a fractional
    Brownian surface of Hurst exponent 0.7), interpolated bilinearly."""
    rng = np.random.default_rng(seed)
    kx = np.fft.fftfreq(n)[:, None]
    ky = np.fft.rfftfreq(n)[None, :]
    kk = np.hypot(kx, ky)
    kk[0, 0] = 1.0
    amp = kk ** (-beta / 2.0)
    amp[0, 0] = 0.0
    spec = amp * np.exp(2j * np.pi * rng.uniform(size=amp.shape))
    g = np.fft.irfft2(spec, s=(n, n))
    g = (g - g.min()) / (g.max() - g.min())
    x0, x1, y0, y1 = float(np.min(x)), float(np.max(x)), float(np.min(y)), float(np.max(y))
    fx = (np.asarray(x) - x0) / max(x1 - x0, 1e-30) * (n - 1)
    fy = (np.asarray(y) - y0) / max(y1 - y0, 1e-30) * (n - 1)
    i0 = np.clip(np.floor(fx).astype(np.int64), 0, n - 2)
    j0 = np.clip(np.floor(fy).astype(np.int64), 0, n - 2)
    tx, ty = fx - i0, fy - j0
    v = (g[i0, j0] * (1 - tx) * (1 - ty) + g[i0 + 1, j0] * tx * (1 - ty) + g[i0, j0 + 1] * (1 - tx) * ty
         + g[i0 + 1, j0 + 1] * tx * ty)
    return (v - v.min()) / (v.max() - v.min())


def storm_surge_case(ncol: int = 980, nrep: int = 245, nlayer: int = 30, dx: float = 100.0, dt: float = 60.0,
                     ndt: int = 2880, seed: int = 11, p_min: float = 950.0, r_max: float = 40.0e3,
                     b: float = 1.5, speed: float = 5.0, heading_deg: float = 30.0, d_temp: float = 10.0,
                     d_salt: float = 3.0, coast_weight: float = 0.55, track_offset: float = 0.0,
                     surface_gap: float = 4.0, name: str = "storm_surge", **kw) -> Case:
    """The baroclinic storm-surge configuration: the stitched square/triangle mesh (1,200,990
    cells at the default size), `nlayer` z-layers, a seeded fractal bathymetry of 1-40 m that
    deepens towards the open (west) boundary, M2+K1 tide there, wet/dry active, stratified
    temperature and salinity (`d_temp` C and `d_salt` psu from surface to bottom: 25 -> 15 C,
    32 -> 35 psu), and a Holland cyclone (`p_min` hPa, `r_max`, B = `b`) moving at `speed`
    m/s on `heading_deg` (from east, counter-clockwise) whose centre passes `track_offset` m to
    the left of the middle of the domain half way through the run; its fields every 10 steps.
    `coast_weight` is the share of the landward ramp in the bathymetry (the rest is the fractal
    field). The surface layer spans `surface_gap` m below to 6 m above mean sea level, more
    than the surge range, so the surface never comes near a level interface: a z-level top
    layer only centimetres thick would take the whole storm stress (currents of tens of m/s,
    and tracers driven to the transport's clamps). No river."""
    kw.setdefault("tide_phase", -90.0)     # the tide starts from zero elevation
    # four transport sub-cycles: the explicit TVD step stays below one cell volume per
    # sub-cycle on the 100 m triangles up to ~3 m/s
    kw.setdefault("trans_sub_iter", 4)
    case = estuary_case(ncol=ncol, nrep=nrep, nlayer=nlayer, dx=dx, dt=dt, ndt=ndt, seed=seed,
                        river_q=0.0, name=name, **kw)
    x, y = case.x, case.y
    coast = (x - x.min()) / (x.max() - x.min())                 # 0 at the open boundary, 1 inland
    g = coast_weight * (1.0 - coast) + (1.0 - coast_weight) * fractal_field(x, y, seed)
    g = (g - g.min()) / (g.max() - g.min())
    case.depth = np.round(1.0 + 39.0 * g, 6)
    zb = float(np.ceil(case.depth.max() + 0.5))
    case.z_level = uniform_levels(zb, nlayer, top=6.0, surf_gap=surface_gap)
    case.rivers, case.q_series = [], []
    L = nlayer
    frac = np.arange(L) / max(L - 1, 1)                           # 0 bottom .. 1 top
    case.salt_init = np.repeat((35.0 - d_salt * frac)[None, :], case.maxnod, axis=0)
    case.salt_series = [Series([-1000.0, 100000.0], [35.0, 35.0])]
    case.is_temp = 1
    t_bottom = 25.0 - d_temp
    case.temp_init = np.repeat((t_bottom + d_temp * frac)[None, :], case.maxnod, axis=0)
    case.temp_series = [Series([-1000.0, 100000.0], [25.0, 25.0])]
    for bnd in case.open_boundaries:
        bnd.salt_ser_id, bnd.temp_ser_id = 1, 1
    # the track: through the domain centre at the middle of the run
    hd = math.radians(heading_deg)
    vel = (speed * math.cos(hd), speed * math.sin(hd))
    half = 0.5 * ndt * dt
    cx, cy = 0.5 * (x.min() + x.max()), 0.5 * (y.min() + y.max())
    cx, cy = cx - track_offset * math.sin(hd), cy + track_offset * math.cos(hd)   # left of the heading
    start = (cx - vel[0] * half, cy - vel[1] * half)
    hdt = int(10 * dt)
    times = [k * float(hdt) for k in range(int(ndt * dt // hdt) + 3)]
    case.hurricane_flag, case.hurricane_dt = 1, hdt
    y0, mo, d = case.start[:3]
    case.hurricane_start, case.hurricane_end = (y0, mo, d), (y0 + 1, mo, d)
    case.hurricane = holland_snapshots(x, y, times, start, vel, p_min=p_min, r_max=r_max, b=b, lat=case.lat)
    return case


def stratify(case: Case, d_temp: float = 10.0, d_salt: float = 3.0, t_bottom: float = 15.0) -> Case:
    """Stratified temperature and salinity: temperature from t_bottom at the bottom layer
    to t_bottom + d_temp at the top; salinity raised towards the bottom by up to d_salt,
    in proportion to the along-axis salinity. Turns on temperature transport (without a
    surface heat flux)."""
    L = case.maxlayer
    frac = np.arange(L) / max(L - 1, 1)                          # 0 bottom .. 1 top
    case.is_temp = 1
    case.temp_init = np.repeat((t_bottom + d_temp * frac)[None, :], case.maxnod, axis=0)
    case.temp_series = [Series([-1000.0, 100000.0], [t_bottom + d_temp, t_bottom + d_temp])]
    for bnd in case.open_boundaries:
        bnd.temp_ser_id = 1
    for r in case.rivers:
        r.temp_ser_id = 1
    if case.salt_init is not None:
        s = np.asarray(case.salt_init, dtype=np.float64)
        case.salt_init = s + d_salt * (1.0 - frac)[None, :] * (s / max(float(s.max()), 1e-30))
    return case


def elm_tracers(case: Case, t: float = 20.0, s_river: float = 5.0, temp_input: bool = True) -> Case:
    """Tracers of an estuary_case that the ELM transport scheme (transport_solver 2) keeps in
    range: temperature t everywhere (at the start, on every open boundary and river) and
    salinity from 35 psu at sea to s_river in the river. That scheme adds temperature differences to salinity in its
    horizontal diffusion and has no clamp, so a cold or fresh river drives salinity below
    zero, where the UNESCO equation of state is undefined. With temp_input False temperature
    is no input (is_temp 0: no temperature files, and no air series for the heat model,
    which the reference then reads) and t must be 0, the value a river without a
    temperature series keeps."""
    if temp_input:
        case.is_temp = 1
        case.temp_init = np.full((case.maxnod, case.maxlayer), t)
        case.temp_series = [Series([-1000.0, 100000.0], [t, t])]
        for b in list(case.open_boundaries) + list(case.rivers):
            b.temp_ser_id = 1
    else:
        if t != 0.0:
            raise ValueError("without temperature input the rivers stay at 0 C")
        case.is_temp, case.ref_temp = 0, 0.0
    case.salt_init = s_river + np.asarray(case.salt_init) * ((35.0 - s_river) / 35.0)
    case.salt_series[1] = Series([-1000.0, 100000.0], [s_river, s_river])
    return case
