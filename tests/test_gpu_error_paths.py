"""Failure reporting of the device step.

* A free-surface solve cut off by max_iteration_pcg is reported as
  GOMGPU_W_PCG_NOT_CONVERGED by gomgpu_sync (the reference's pre_conj_grad.f90:224-229 only
  prints cgit2 = 0 and carries on with the last iterate); the iteration count reads 0, the
  warning is reported once, and eta is the last iterate, as the oracle's is with the same
  limit.
* An ELM backtrace starting in a cell whose surface lies below its bed while the cell
  still has layers (search_straight_line.f90:37-46, the reference's `stop`) latches
  GOMGPU_ERR_ELM_DRY_START with a face that starts there, and the oracle stops on the same
  condition.
"""
from __future__ import annotations

import re

import numpy as np
import pytest

from golden_io import Golden, apply
from oracle import Oracle

pytestmark = pytest.mark.gpu

W_PCG_NOT_CONVERGED = 5
E_MODEL = 4
ERR_ELM_DRY_START = 8


@pytest.fixture(scope="module")
def gpu_available():
    from gom_release_amd import host
    if host.device_count() < 1:
        pytest.fail("no GPU visible: the HIP library cannot run")
    return True


@pytest.mark.parametrize("name", ["estuary_wetdry", "hurricane_estuary"])
def test_pcg_not_converged_is_reported(gpu_available, name):
    from gom_release_amd import abi, host
    g = Golden(name)
    case = g.case()
    it = g.steps[-1]
    g.params.max_iteration_pcg = 1
    state = g.state_in(it)
    for tag in ("pre", "elm", "mom"):
        apply(state, g.expected(it, tag))
    forcing = g.forcing(it, case)
    model = host.GpuModel(g.mesh, g.params)
    model.set_forcing(forcing)
    model.set_state(state)
    model.stage("freesurf")
    rc = model.sync(allow_pcg_warning=True)
    print(f"[error-path] {name} step {it}: sync -> {rc}, iterations {model.pcg_iterations()}", flush=True)
    assert rc == W_PCG_NOT_CONVERGED
    assert model.pcg_iterations() == 0
    assert model.sync() == 0                 # reported once
    out = abi.state_arrays(g.params)
    model.get_state(out)
    o = Oracle(g.mesh, g.params, state, forcing)
    assert o.run("freesurf") == 0
    assert Oracle.pcg_iterations() == 0
    np.testing.assert_allclose(out["eta_cell_new"], state["eta_cell_new"], rtol=1e-9, atol=1e-12)
    conv = g.expected(it, "fs")["eta_cell_new"]
    assert not np.allclose(out["eta_cell_new"], conv, rtol=1e-9, atol=1e-12)   # a real cut-off
    model.close()


def test_elm_dry_start_is_latched(gpu_available):
    from gom_release_amd import host
    g = Golden("channel_16x8x5")
    case = g.case()
    it = g.steps[-1]
    state = g.state_in(it)
    apply(state, g.expected(it, "pre"))
    forcing = g.forcing(it, case)
    # a cell that starts backtraces: the first adjacent cell of a face whose top layer moves
    probe = Oracle(g.mesh, g.params, g.state_in(it), forcing)
    apply(probe.state, g.expected(it, "pre"))
    rc, nnel, _ = probe.elm_departures()
    assert rc == 0
    adj = np.asarray(g.mesh["adj_cellnum_at_face"]).reshape(-1, 2)
    j0 = int(np.flatnonzero((nnel != 0).any(axis=0))[0])
    cell = int(adj[j0, 0])
    eta = np.array(state["eta_cell"])
    eta[cell - 1] = -float(g.mesh["h_cell"][cell - 1]) - 0.5     # surface below the bed, layers kept
    state["eta_cell"] = eta
    starts = {j + 1 for j in range(g.params.maxface) if adj[j, 0] == cell or adj[j, 1] == cell}

    o = Oracle(g.mesh, g.params, state, forcing)
    assert o.run("elm") == ERR_ELM_DRY_START
    assert Oracle.last_error_entity() in starts

    model = host.GpuModel(g.mesh, g.params)
    model.set_forcing(forcing)
    model.set_state(state)
    model.stage("elm")
    with pytest.raises(host.GomGpuError) as ei:
        model.sync()
    assert ei.value.code == E_MODEL
    m = re.search(r"model error (\d+) at entity (\d+)", str(ei.value))
    print(f"[error-path] latched: {ei.value}", flush=True)
    assert m and int(m.group(1)) == ERR_ELM_DRY_START
    assert int(m.group(2)) in starts
    model.close()
