"""Boundary forcing on the device (GOMGPU_STAGE_BOUNDARY): the open-boundary elevations of
compute_boundary_eta.f90:72-124, the river discharge of calculate_Q.f90:21-45 and the
transport's boundary concentrations (solve_transport_equation_v19.f90:229-397, or
find_openboundary_salt_temp_v3.f90:17-97 for the ELM solver), against the host forcing
(gom_release_amd.forcing), which the host-driver runs pin to the reference's fixtures.

The cases exercise what the reference's loops carry: harmonic and elevation-series open
boundaries side by side, series that stop covering the time part way through (the value of
the boundary before carries over, and a river's Q_add keeps its last value), boundaries
without a series, two tracers over TVD sub-cycles, a nonzero reference_diff_days and
spin-up of the tide and the baroclinic forcing.

The device's harmonic terms use gom_cos, which is correctly rounded; the C library's cos
(which the reference and the host forcing call) rounds a few arguments in a thousand the
other way. So the device is held bit for bit to the host forcing evaluated with gom_cos,
and within a few ulps to the host forcing as it stands.
"""
from __future__ import annotations

import ctypes as C
import math
import types

import numpy as np
import pytest

from gom_release_amd import abi, forcing as fz, host, meshgen
from gom_release_amd.driver import HostForcing, Scenario
from gom_release_amd.gomio import Series

pytestmark = pytest.mark.gpu

ARRAYS = ["eta_at_ob_old", "eta_at_ob_new", "con_at_ob", "con_at_Qb", "Q_add"]


def _gom_cos_module():
    L = C.CDLL(host.LIB_PATH)
    L.gom_cos_array.argtypes = [C.c_void_p, C.c_void_p, C.c_long]

    def gcos(x):
        a = np.array([x], dtype=np.float64)
        y = np.empty_like(a)
        L.gom_cos_array(a.ctypes.data, y.ctypes.data, 1)
        return float(y[0])

    m = types.ModuleType("math_gom_cos")
    m.__dict__.update({k: getattr(math, k) for k in dir(math) if not k.startswith("_")})
    m.cos = gcos
    return m


def _series(days, vals):
    return Series(list(days), list(vals))


def _case(transport: str):
    kw = dict(ncol=16, nrep=3, nlayer=5, dt=120.0, ndt=40, tide_phase=10.0, tide_spinup=1,
              tide_spinup_period=0.02, baroclinic_spinup=1, baroclinic_spinup_period=0.02)
    if transport == "elm":
        kw["transport_solver"] = 2
    c = meshgen.estuary_case(**kw)
    c.baroclinic_flag = 1
    if transport == "elm":
        meshgen.elm_tracers(c)
    else:
        meshgen.stratify(c)
    # the series count from two days after the start (reference_diff_days = -2): they cover
    # day -2 of their own axis on, and the "gap" ones only up to -1.99, i.e. the first ~7 steps
    c.data_time_shift = 2
    full = [-3.0, -2.03, -1.97, 5.0]
    gap = [-3.0, -2.004, -1.99]
    c.salt_series = [_series(full, [34.0, 35.0, 34.5, 35.0]), _series(full, [1.0, 0.5, 2.0, 0.0]),
                     _series(gap, [30.0, 33.0, 31.0])]
    c.temp_series = [_series(full, [18.0, 20.0, 19.0, 21.0]), _series(gap, [10.0, 12.0, 11.0])]
    c.q_series = [_series(full, [400.0, 500.0, 450.0, 600.0]), _series(gap, [50.0, 80.0, 60.0])]
    c.eta_series = [_series(full, [0.1, 0.3, -0.2, 0.0]), _series(gap, [0.0, 0.4, 0.1])]
    for i, b in enumerate(c.open_boundaries):
        if i % 3 == 0:
            b.eta_type, b.eta_ser_id = 2, 2
        elif i % 3 == 1 and i % 2 == 1:
            b.eta_type, b.eta_ser_id = 2, 1
        b.salt_ser_id = (1, 3, 0)[i % 3]
        b.temp_ser_id = (1, 2, 0)[(i + 1) % 3]
    for i, r in enumerate(c.rivers):
        r.q_ser_id = (1, 2, 1)[i % 3]
        r.salt_ser_id = (2, 3, 0)[i % 3]
        r.temp_ser_id = (1, 0, 2)[i % 3]
    assert c.reference_diff_days() == -2
    return c


@pytest.mark.parametrize("transport", ["tvd", "elm"])
def test_boundary_stage_equals_host_forcing(transport, monkeypatch):
    case = _case(transport)
    sc = Scenario(case, device_boundary=False)   # the host path, for its state and arrays
    p, mesh = sc.params, sc.mesh
    model = sc.model
    model.set_boundary(fz.boundary_tables(case))
    monkeypatch.setattr(fz, "math", _gom_cos_module())
    hf_cr = HostForcing(case, p, mesh)            # host forcing with the correctly rounded cos
    monkeypatch.undo()
    hf_libm = HostForcing(case, p, mesh)          # host forcing as it stands
    out = abi.ForcingArrays(p)
    eta_ulps = []
    carried = False
    for it in range(1, 31):
        monkeypatch.setattr(fz, "math", _gom_cos_module())
        want = hf_cr.advance(it)
        monkeypatch.undo()
        libm = hf_libm.advance(it)
        model.set_forcing(want, only=[])          # the step's scalars only
        model.stage("boundary")
        model.get_forcing(out, only=ARRAYS)
        for name in ARRAYS:
            assert np.array_equal(out[name], want[name]), (transport, it, name)
        for name in ("eta_at_ob_old", "eta_at_ob_new"):
            d = np.abs(out[name].view(np.int64) - libm[name].view(np.int64))
            eta_ulps.append(int(d.max()))
        for name in ("con_at_ob", "con_at_Qb", "Q_add"):
            assert np.array_equal(libm[name], want[name])
        carried = carried or it > 8
    assert carried and max(eta_ulps) <= 4
    # the gap series did stop covering the time (the carried values are in play)
    assert fz.series_value(case.salt_series[2], hf_cr.julian_day * 86400.0, case.reference_diff_days()) is None
    model.close()


def test_scenario_with_device_boundary_follows_the_host_path():
    case = _case("tvd")
    a = Scenario(case)                              # boundary forcing on the device
    b = Scenario(case, device_boundary=False)
    assert a.host_forcing.device_boundary and not b.host_forcing.device_boundary
    names = ["eta_cell", "un_face", "salt_cell", "temp_cell", "top_layer_at_element"]
    for _ in range(12):
        a.step()
        b.step()
    fa, fb = a.fields(names), b.fields(names)
    for n in names:
        x, y = fa[n], fb[n]
        err = np.max(np.abs(x - y)) / max(np.max(np.abs(y)), 1e-30)
        assert err < 1e-9, (n, err)
    assert np.array_equal(fa["top_layer_at_element"], fb["top_layer_at_element"])
    a.model.close()
    b.model.close()
