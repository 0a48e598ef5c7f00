"""The Holland cyclone generated on the device (gomgpu_set_holland, k_holland_next) against
the host generator (meshgen.HollandTrack): the snapshot each read step takes, and whole runs
through several snapshots against the same runs fed the host's snapshots. The node fields
agree to a few ulps (the device's pow, exp and hypot are not the host library's); the runs to
the whole-step tolerances."""
from __future__ import annotations

import numpy as np
import pytest

from gom_release_amd import abi, meshgen
from gom_release_amd.driver import Scenario

pytestmark = pytest.mark.gpu


def _case():
    c = meshgen.estuary_case(ncol=24, nrep=4, nlayer=5, dt=120.0, ndt=40)
    # fields every 3 steps: several snapshots within the run, the centre crossing the domain
    meshgen.add_cyclone(c, 40, seed=2, p_min=955.0, r_max=2.0e3, b=1.5, speed=8.0, hurricane_dt=360)
    return c


def test_device_snapshots_equal_the_host_generator():
    case = _case()
    sc = Scenario(case)
    assert sc.host_forcing.device_holland
    tr = case.hurricane
    f = abi.ForcingArrays(sc.params)
    checked = 0
    for it in range(1, 16):
        sc.step()
        if sc.forcing.scalars["hurricane_phase"] != 2.0:
            continue
        sc.model.get_forcing(f, only=["wind_u2", "wind_v2", "air_p2"])
        u, v, p = tr.at(tr.times[sc.host_forcing._snap_index - 1])
        for got, want in ((f["wind_u2"], u), (f["wind_v2"], v), (f["air_p2"], np.asarray(p) * 100.0)):
            err = np.max(np.abs(got - want)) / np.max(np.abs(want))
            assert err < 1e-13, (it, err)
        checked += 1
    assert checked >= 3
    sc.model.close()


def test_runs_with_device_and_host_snapshots_agree():
    case = _case()
    a = Scenario(case)
    b = Scenario(case, device_holland=False)
    assert a.host_forcing.device_holland and not b.host_forcing.device_holland
    names = ["eta_cell", "un_face", "vn_face", "salt_cell", "top_layer_at_element"]
    for _ in range(20):
        a.step()
        b.step()
    fa, fb = a.fields(names), b.fields(names)
    for n in names[:-1]:
        err = np.max(np.abs(fa[n] - fb[n])) / max(np.max(np.abs(fb[n])), 1e-30)
        assert err < 1e-9, (n, err)
    assert np.array_equal(fa["top_layer_at_element"], fb["top_layer_at_element"])
    a.model.close()
    b.model.close()
