"""The launch-shape autotuner on the GPU: tuning changes no result (a tuned scenario goes on
bit for bit like an untuned one) and the launch table rejects shapes outside its limits."""
from __future__ import annotations

import numpy as np
import pytest

from gom_release_amd import abi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu_available():
    from gom_release_amd import host
    if host.device_count() < 1:
        pytest.fail("no GPU visible: the HIP library cannot run")
    return True


def test_tuned_scenario_continues_identically(gpu_available):
    from gom_release_amd import autotune, meshgen
    from gom_release_amd.driver import Scenario
    print("[gpu-tune] tune + compare", flush=True)
    case = meshgen.estuary_case(ncol=60, nrep=20, nlayer=8, river_q=100.0, trans_sub_iter=2,
                                h_flux_limiter=1, v_flux_limiter=1, turbulence_flag=3)
    # cyclone fields read every third step: tuning must leave the wind stage's device
    # fields and the host's snapshot stream where they were
    meshgen.add_cyclone(case, 40, seed=4, r_max=20e3, p_min=970.0, hurricane_dt=int(3 * case.dt))
    a = Scenario(case, wind=(6.0, 2.0))
    a.run(2)
    cfg = autotune.tune(a, steps=1, candidates=(64, 512))
    assert set(cfg) == set(__import__("gom_release_amd.host", fromlist=["x"]).LAUNCH)
    a.run(3)
    a.model.sync()
    b = Scenario(case, wind=(6.0, 2.0))
    b.run(5)
    b.model.sync()
    sa, sb = abi.state_arrays(a.params), abi.state_arrays(b.params)
    a.model.get_state(sa)
    b.model.get_state(sb)
    for f in sa.arr:
        assert np.array_equal(sa[f], sb[f]), f
    a.model.close()
    b.model.close()


def test_launch_limits(gpu_available):
    from gom_release_amd import host, meshgen
    from gom_release_amd.driver import Scenario
    sc = Scenario(meshgen.channel_case(nx=8, ny=4, nlayer=3))
    for bad in (32, 100, 2048):
        with pytest.raises(host.GomGpuError):
            sc.model.set_launch("elm", bad)
    for which in ("turb_solve", "mom_rhs", "elm"):   # kernels compiled for at most 256 threads
        with pytest.raises(host.GomGpuError):
            sc.model.set_launch(which, 512)
    sc.model.set_launch("elm", 256)
    assert sc.model.get_launch("elm") == 256
    sc.model.set_launch("tvd_rhs", 512)
    assert sc.model.get_launch("tvd_rhs") == 512
    sc.model.close()
