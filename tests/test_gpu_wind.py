"""The device's wind stage and hurricane-forced steps against the reference.

On tests/golden/hurricane_estuary.npz (Holland-cyclone fields read every third step, the
window opening at step 4): the wind stage run on the device from the reference's fields
before each step must give, bit for bit, the node fields, face winds, wind stress and air
pressure the reference had after its forcing routines; and eight device steps from a cold
start through the host driver (snapshots, phases, uploads) must follow the reference's
states within the multi-step tolerances (the PCG sums in another order).
"""
from __future__ import annotations

import numpy as np
import pytest

from golden_io import WIND_OUTPUTS, Golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu_available():
    from gom_release_amd import host
    if host.device_count() < 1:
        pytest.fail("no GPU visible: the HIP library cannot run")
    return True


def test_gpu_wind_stage_bitwise(gpu_available):
    from gom_release_amd import abi, host
    g = Golden("hurricane_estuary")
    case = g.case()
    model = host.GpuModel(g.mesh, g.params)
    for it in g.steps:
        print(f"[gpu-wind] step {it}", flush=True)
        f = g.step_forcing(it, case)
        model.set_state(g.state_in(it))
        model.set_forcing(f)
        model.stage("wind")
        out = abi.ForcingArrays(g.params)
        model.get_forcing(out, only=WIND_OUTPUTS)
        pre = g.group(f"s{it}_pre")
        for n in WIND_OUTPUTS:
            ref = np.asarray(pre[n]).ravel()
            bad = np.flatnonzero(out[n] != ref)
            assert bad.size == 0, f"step {it} {n}: {bad.size} differ, first {bad[:5]}"
    model.close()


def test_gpu_hurricane_run_matches_reference(gpu_available):
    from gom_release_amd.driver import Scenario
    g = Golden("hurricane_estuary")
    sc = Scenario(g.case())
    for it in g.steps:
        print(f"[gpu-wind] driver step {it}", flush=True)
        sc.step()
        exp = g.expected(it, "post")
        st = sc.fields(list(exp))
        for k, ref in exp.items():
            got = st[k]
            if got.dtype != np.float64:
                assert np.array_equal(got, ref), (it, k)
            else:
                np.testing.assert_allclose(got, ref, rtol=1e-7, atol=1e-10, err_msg=f"step {it} {k}")
    sc.model.close()


def test_gpu_storm_case_matches_oracle(gpu_available):
    """A storm-surge case of the benchmark family on a small mesh: Holland cyclone fields
    read every third step, stratified temperature and salinity transported, GLS closure,
    wet/dry banks: eight device steps through the host driver against the oracle driven
    the same way."""
    from oracle_model import OracleModel
    from gom_release_amd import meshgen
    from gom_release_amd.driver import Scenario
    case = meshgen.estuary_case(ncol=24, nrep=4, nlayer=6, ndt=8, dt=60.0, river_q=20.0, dry_land=True,
                                trans_sub_iter=2, h_flux_limiter=1, v_flux_limiter=1, turbulence_flag=3)
    meshgen.add_cyclone(case, 8, seed=5, r_max=20e3, p_min=960.0, hurricane_dt=int(3 * case.dt))
    meshgen.stratify(case)
    gpu = Scenario(case)
    cpu = Scenario(case, model_factory=OracleModel)
    names = ["eta_cell", "un_face", "vn_face", "wn_cell", "salt_cell", "temp_cell", "rho_cell",
             "top_layer_at_element", "Av", "q2"]
    for it in range(1, case.ndt + 1):
        print(f"[gpu-wind] storm step {it}", flush=True)
        gpu.step()
        cpu.step()
        a, b = gpu.fields(names), cpu.fields(names)
        for k in names:
            if a[k].dtype != np.float64:
                assert np.array_equal(a[k], b[k]), (it, k)
            else:
                np.testing.assert_allclose(a[k], b[k], rtol=1e-7, atol=1e-10, err_msg=f"step {it} {k}")
    gpu.model.close()
