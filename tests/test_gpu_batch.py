"""Batches (gomgpu_batch_*): several scenarios advanced together must give, bit for bit,
what each gives alone -- with the step captured as a graph or submitted eagerly, and with
members joining and leaving between steps (continuous batching)."""
from __future__ import annotations

import os

import numpy as np
import pytest

from gom_release_amd import abi

pytestmark = pytest.mark.gpu

WIND = (6.0, 2.0)


@pytest.fixture(scope="module")
def gpu_available():
    from gom_release_amd import host
    if host.device_count() < 1:
        pytest.fail("no GPU visible: the HIP library cannot run")
    return True


def _cases():
    from gom_release_amd import meshgen
    # the second member has cyclone forcing (its wind stage runs in the batch), the third
    # transports a stratified temperature as well
    storm = meshgen.estuary_case(ncol=20, nrep=3, nlayer=6, ndt=20, elm_backtrace_flag=1, elm_max_iter=10,
                                 dry_land=True, river_q=50.0, tide_phase=-90.0, turbulence_flag=3)
    meshgen.add_cyclone(storm, 40, seed=1, r_max=2000.0, hurricane_dt=int(3 * storm.dt))
    return [
        meshgen.channel_case(nx=16, ny=8, nlayer=5, ndt=20),
        storm,
        meshgen.stratify(meshgen.estuary_case(ncol=60, nrep=20, nlayer=8, ndt=20, river_q=100.0, trans_sub_iter=2,
                                              h_flux_limiter=1, v_flux_limiter=1)),
    ]


def _state(sc):
    s = abi.state_arrays(sc.params)
    sc.model.get_state(s)
    return s


def _alone(case, nsteps, rows=0):
    from gom_release_amd.driver import Scenario
    sc = Scenario(case, wind=WIND)
    sc.model.set_pcg_rows(rows)
    sc.run(nsteps)
    sc.model.sync()
    out = _state(sc)
    sc.model.close()
    return out


def _assert_same(got, ref, tag):
    for f in ref.arr:
        bad = np.flatnonzero(got[f] != ref[f])
        assert bad.size == 0, f"{tag} {f}: {bad.size} values differ, first {bad[:5]}"


@pytest.mark.parametrize("graph,rows", [(1, 0), (0, 0), (1, 4)])
def test_batch_equals_alone(gpu_available, graph, rows):
    """rows: the PCG partition (0 automatic; 4, the batch-packing one), the same alone."""
    from gom_release_amd import host
    from gom_release_amd.driver import Scenario
    print(f"[gpu-batch] graph={graph} rows={rows}", flush=True)
    cases = _cases()
    n = 4
    refs = [_alone(c, n, rows) for c in cases]
    os.environ["GOMGPU_BATCH_GRAPH"] = str(graph)
    try:
        b = host.GpuBatch()
    finally:
        os.environ.pop("GOMGPU_BATCH_GRAPH", None)
    scs = [Scenario(c, wind=WIND) for c in cases]
    for sc in scs:
        sc.model.set_pcg_rows(rows)
        b.add(sc.model)
    for _ in range(n):
        for sc in scs:
            sc.stage_forcing()
        b.step()
    b.sync()
    for i, sc in enumerate(scs):
        _assert_same(_state(sc), refs[i], f"member {i}")
    b.close()
    for sc in scs:
        sc.model.close()


def test_continuous_batching_members_join_and_leave(gpu_available):
    """Member 2 joins after 2 steps of members 0 and 1; member 0 leaves after 3 steps. Each
    result equals its own run alone for the steps it took."""
    from gom_release_amd import host
    from gom_release_amd.driver import Scenario
    print("[gpu-batch] continuous", flush=True)
    cases = _cases()
    plan = {0: (0, 3), 1: (0, 5), 2: (2, 5)}   # member: (join step, leave step)
    refs = {i: _alone(cases[i], plan[i][1] - plan[i][0]) for i in plan}
    b = host.GpuBatch()
    scs = {}
    for t in range(5):
        for i, (t0, t1) in plan.items():
            if t == t0:
                scs[i] = Scenario(cases[i], wind=WIND)
                b.add(scs[i].model)
            if t == t1 and i in scs and scs[i].model in b.members:
                b.remove(scs[i].model)
        for i, sc in scs.items():
            if sc.model in b.members:
                sc.stage_forcing()
        b.step()
    b.sync()
    for i, sc in scs.items():
        sc.model.sync()
        _assert_same(_state(sc), refs[i], f"member {i}")
    b.close()
    for sc in scs.values():
        sc.model.close()


def test_fused_groups_and_a_lone_solve(gpu_available):
    """Members of uneven size: two of about 300k cells (each half of the co-resident PCG
    grid, so they land in different fused launches), a small one packed beside the second,
    and one of 600k cells (more blocks than the grid holds: solved by its own
    memory-resident launch in the batch step). Each equals its run alone bit for bit."""
    from gom_release_amd import host, meshgen, serving
    from gom_release_amd.driver import Scenario
    cases = [meshgen.channel_case(nx=600, ny=500, nlayer=4, ndt=4),
             serving.estuary_of_size(300000, 4, seed=5, river_q=80.0, ndt=4),
             meshgen.channel_case(nx=16, ny=8, nlayer=5, ndt=4),
             serving.estuary_of_size(600000, 4, seed=6, river_q=80.0, ndt=4)]
    n = 3
    refs = [_alone(c, n) for c in cases]
    print("[gpu-batch] uneven members alone done", flush=True)
    b = host.GpuBatch()
    scs = [Scenario(c, wind=WIND) for c in cases]
    for sc in scs:
        b.add(sc.model)
    for _ in range(n):
        for sc in scs:
            sc.stage_forcing()
        b.step()
    b.sync()
    groups, solo = b.layout()
    print(f"[gpu-batch] {groups} fused PCG launches, {solo} lone", flush=True)
    assert (groups, solo) == (2, 1)
    for i, sc in enumerate(scs):
        _assert_same(_state(sc), refs[i], f"member {i} ({sc.params.maxele} cells)")
    b.close()
    for sc in scs:
        sc.model.close()
