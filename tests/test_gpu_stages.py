"""HIP stages against the reference fixtures and the oracle, stage by stage.

Each stage runs on the device from the reference's own input state. Outside the free
surface every stage must reproduce the reference bit for bit (FP64, same evaluation order,
no contraction). The free-surface solve differs only in the association of the PCG dot
products (deterministic tree reductions on the device, a serial sum in the reference), so
eta is compared to a relative tolerance of 1e-9 (absolute floor 1e-12 m) and the wet/dry
node flags must be exact.
"""
import numpy as np
import pytest

from golden_io import FIXTURES, STAGES, Golden, apply

pytestmark = pytest.mark.gpu

FS_RTOL = 1e-9
FS_ATOL = 1e-12


@pytest.fixture(scope="module")
def gpu_available():
    from gom_release_amd import host
    if host.device_count() < 1:
        pytest.fail("no GPU visible: the HIP library cannot run")
    return True


def _model(g):
    from gom_release_amd import host
    return host.GpuModel(g.mesh, g.params)


@pytest.mark.parametrize("name", FIXTURES)
def test_gpu_stages_match_reference(gpu_available, name):
    g = Golden(name)
    case = g.case()
    model = _model(g)
    from gom_release_amd import abi
    for it in g.steps:
        print(f"[gpu-stage] {name} step {it}", flush=True)
        state = g.state_in(it)
        forcing = g.forcing(it, case)
        model.set_forcing(forcing)
        for stage, tag in STAGES:
            model.set_state(state)
            model.stage(stage)
            model.sync()
            if stage == "freesurf" and g.pcg_iterations(it) is not None:
                assert model.pcg_iterations() == g.pcg_iterations(it), (name, it, model.pcg_iterations())
            out = abi.state_arrays(g.params)
            model.get_state(out)
            for field, ref in g.expected(it, tag).items():
                if field not in out.arr:
                    continue
                got = out[field]
                if stage == "freesurf" and field in ("eta_cell_new", "eta_node"):
                    np.testing.assert_allclose(got, ref, rtol=FS_RTOL, atol=FS_ATOL,
                                               err_msg=f"{name} step {it} {field}")
                else:
                    bad = np.flatnonzero(got != ref)
                    assert bad.size == 0, f"{name} step {it} {stage} {field}: {bad.size} differ, first {bad[:5]}"
            apply(state, g.expected(it, tag))
    model.close()


@pytest.mark.parametrize("name", FIXTURES)
def test_gpu_whole_step_matches_reference(gpu_available, name):
    """One full device step from the reference state: eta within 1e-9 relative, wet/dry
    masks and ELM-derived indices exact, everything else within 1e-8 relative."""
    g = Golden(name)
    case = g.case()
    from gom_release_amd import abi
    model = _model(g)
    for it in g.steps:
        model.set_state(g.state_in(it))
        model.set_forcing(g.step_forcing(it, case))
        model.step()
        model.sync()
        out = abi.state_arrays(g.params)
        model.get_state(out)
        for tag in ("lay", "post"):
            for field, ref in g.expected(it, tag).items():
                got = out[field]
                if got.dtype != np.float64:
                    assert np.array_equal(got, ref), (name, it, field)
                else:
                    np.testing.assert_allclose(got, ref, rtol=1e-8, atol=1e-11, err_msg=f"{name} {it} {field}")
    model.close()


@pytest.mark.parametrize("name", FIXTURES)
def test_gpu_freesurf_memory_resident_pcg(gpu_available, name, monkeypatch):
    """The memory-resident solve k_pcg_cg1m (the one meshes above 524,288 cells and large
    batch members run) forced on every fixture: eta within the free-surface tolerance, the
    wet/dry node flags exact, and the reference's iteration count when the fixture holds it."""
    monkeypatch.setenv("GOMGPU_PCG_MODE", "4")
    g = Golden(name)
    case = g.case()
    model = _model(g)
    from gom_release_amd import abi
    for it in g.steps:
        print(f"[gpu-cg1m] {name} step {it}", flush=True)
        state = g.state_in(it)
        for stage, tag in STAGES[:3]:
            apply(state, g.expected(it, tag))
        model.set_forcing(g.forcing(it, case))
        model.set_state(state)
        model.stage("freesurf")
        model.sync()
        out = abi.state_arrays(g.params)
        model.get_state(out)
        exp = g.expected(it, "fs")
        for field in ("eta_cell_new", "eta_node"):
            np.testing.assert_allclose(out[field], exp[field], rtol=FS_RTOL, atol=FS_ATOL,
                                       err_msg=f"{name} step {it} {field}")
        if "wetdry_node" in exp:
            assert np.array_equal(out["wetdry_node"], exp["wetdry_node"]), (name, it)
        ref_it = g.pcg_iterations(it)
        if ref_it is not None:
            assert model.pcg_iterations() == ref_it, (name, it, model.pcg_iterations(), ref_it)
    model.close()
