"""The serving loop on the GPU: admission by footprint, continuous batching, and results
that are bit for bit each scenario's result alone."""
from __future__ import annotations

import numpy as np
import pytest

from gom_release_amd import abi

pytestmark = pytest.mark.gpu

KEEP = ("eta_cell", "un_face", "salt_cell", "top_layer_at_element")


@pytest.fixture(scope="module")
def gpu_available():
    from gom_release_amd import host
    if host.device_count() < 1:
        pytest.fail("no GPU visible: the HIP library cannot run")
    return True


def test_footprint_is_what_create_allocates(gpu_available):
    from gom_release_amd import host, meshgen, prep
    for case in (meshgen.channel_case(nx=16, ny=8, nlayer=5),
                 meshgen.estuary_case(ncol=20, nrep=3, nlayer=6, turbulence_flag=3, trans_sub_iter=2)):
        p, m, _ = prep.build(case)
        model = host.GpuModel(m, p)
        assert model.device_bytes() == host.footprint(p)
        model.close()


def _requests(n_steps=4):
    from gom_release_amd import serving
    sizes = [3000, 12000, 6000, 20000]
    arr = [0.0, 0.05, 0.1, 0.15]
    return [serving.Request(i, serving.estuary_of_size(sz, 6, seed=10 + i, river_q=50.0), n_steps + i, arr[i],
                            wind=(5.0, 1.0), keep_state=KEEP) for i, sz in enumerate(sizes)]


def _alone(req):
    from gom_release_amd import serving
    from gom_release_amd.driver import Scenario
    sc = Scenario(req.case, wind=req.wind)
    sc.model.set_pcg_rows(serving.BATCH_PCG_ROWS)   # the served scenarios' partition
    sc.run(req.nsteps)
    sc.model.sync()
    s = abi.state_arrays(sc.params)
    sc.model.get_state(s)
    sc.model.close()
    return s


@pytest.mark.parametrize("budget_mib", [1 << 16, None])
def test_served_results_equal_alone(gpu_available, budget_mib):
    """A large budget admits everything at once; a budget of the largest footprint makes
    requests wait for memory. Either way every result equals the scenario alone and the
    active footprints never exceed the budget."""
    from gom_release_amd import host, prep, serving
    reqs = _requests()
    fps = [host.footprint(prep.build(r.case)[0]) for r in reqs]
    budget = (budget_mib << 20) if budget_mib else max(fps)
    print(f"[gpu-serve] budget {budget >> 20} MiB", flush=True)
    results, summary = serving.ServingLoop(budget, workers=2).run(reqs, log=lambda m: print(m, flush=True))
    assert summary["requests"] == len(reqs)
    assert summary["p99_latency_s"] >= summary["p50_latency_s"] > 0
    for r, req in zip(results, reqs):
        ref = _alone(req)
        for f in KEEP:
            assert np.array_equal(r.state[f], ref[f]), (req.rid, f)
    for r in results:  # footprint of everything active when r was admitted
        live = [x.footprint for x in results if x.admitted <= r.admitted < x.completed]
        assert sum(live) <= budget


@pytest.mark.parametrize("rate", [20.0, 1.5])
def test_large_request_served_alone_in_the_batch(gpu_available, rate):
    """Requests of 8k-40k cells and one of 600k cells (more PCG rows than one fused launch
    holds: it is solved by its own memory-resident launch inside the batch step), arriving
    as Poisson processes of 20/s and 1.5/s. Every result equals the scenario alone."""
    from gom_release_amd import serving
    sizes = [8000, 600000, 40000, 15000]
    arr = serving.poisson_trace(len(sizes), rate, seed=7)
    reqs = [serving.Request(i, serving.estuary_of_size(sz, 4, seed=20 + i, river_q=50.0), 3, arr[i],
                            wind=(5.0, 1.0), keep_state=KEEP) for i, sz in enumerate(sizes)]
    print(f"[gpu-serve] rate {rate}/s, arrivals {[round(a, 3) for a in arr]}", flush=True)
    results, summary = serving.ServingLoop(1 << 36, workers=2).run(reqs, log=lambda m: print(m, flush=True))
    print(f"[gpu-serve] {summary}", flush=True)
    assert summary["requests"] == len(reqs)
    assert summary["most_lone_pcg_solves"] >= 1
    assert max(r.cells for r in results) > 524288
    for r, req in zip(results, reqs):
        ref = _alone(req)
        for f in KEEP:
            assert np.array_equal(r.state[f], ref[f]), (req.rid, f)
