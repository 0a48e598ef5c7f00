"""Serving loop: scenario solves admitted by HBM footprint and stepped with continuous batching.

Requests (a case, a step count, an arrival time) arrive over time. On arrival a request is
prepared on the host (mesh set-up and start-up state, as GOM's start-up does) in a pool of
worker processes, so preparation overlaps the device work. A prepared request is admitted
as soon as its device footprint (gomgpu_footprint) fits in the HBM budget next to the
active scenarios; it then joins the device batch (gomgpu_batch_*), whose every step
advances all active scenarios together. A scenario that has taken its steps leaves the
batch, its result (final state and budgets) is read back and its device memory is freed,
which may admit the next request. Latency is completion minus arrival.

Each scenario's result is bit for bit what it gives alone (the batch guarantees it), so
the loop changes when a result is ready, never what it is.
"""
from __future__ import annotations

import multiprocessing
import time
from concurrent.futures import ProcessPoolExecutor
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import abi, host
from .driver import Scenario, pack_prepared
from .gomio import Case


@dataclass
class Request:
    rid: int
    case: Case
    nsteps: int
    arrival: float                     # seconds after the start of the run
    wind: Tuple[float, float] = (0.0, 0.0)
    keep_state: Tuple[str, ...] = ()   # state fields to return with the result


@dataclass
class Result:
    rid: int
    cells: int
    layers: int
    nsteps: int
    arrival: float
    admitted: float
    completed: float
    footprint: int
    active_cell_layer_steps: int
    volume: float
    salt: float
    state: Dict[str, np.ndarray] = field(default_factory=dict)

    @property
    def latency(self) -> float:
        return self.completed - self.arrival


class _Active:
    def __init__(self, req: Request, sc: Scenario, admitted: float, fp: int):
        self.req, self.sc, self.admitted, self.fp = req, sc, admitted, fp
        self.steps = 0
        self.work = 0
        bot = sc.mesh["bottom_layer_at_element"].astype(np.int64)
        self._bot = bot


def _active_cell_layers(sc: Scenario) -> int:
    s = abi.state_arrays(sc.params)
    sc.model.get_state(s, only=["top_layer_at_element"])
    top = s["top_layer_at_element"].astype(np.int64)
    bot = sc.mesh["bottom_layer_at_element"].astype(np.int64)
    return int(np.sum(np.where(top > 0, top - bot + 1, 0)))


# Rows per thread of a batch member's PCG partition: 4 packs the most members into one
# fused solve launch. A scenario gives the same results alone with the same setting.
BATCH_PCG_ROWS = 4


class ServingLoop:
    """Continuous-batching scenario server on one GPU."""

    def __init__(self, hbm_budget: int, workers: int = 2, max_active: int = 64, pcg_rows: int = BATCH_PCG_ROWS):
        self.budget = int(hbm_budget)
        self.workers = workers
        self.max_active = max_active
        self.pcg_rows = pcg_rows   # PCG partition of the served scenarios (see GpuModel.set_pcg_rows)

    def run(self, requests: List[Request], log=None) -> Tuple[List[Result], Dict[str, float]]:
        reqs = sorted(requests, key=lambda r: r.arrival)
        batch = host.GpuBatch()
        results: List[Result] = []
        active: List[_Active] = []
        prepared: List[Tuple[Request, object, int]] = []   # ready to admit, arrival order
        futures = {}
        used = 0
        nxt = 0
        most_groups = most_solo = 0   # fused PCG launches and lone solves in one step
        # fresh interpreters (never a fork of this process, which holds the GPU); the
        # workers do host work only
        ctx = multiprocessing.get_context("spawn")
        with ProcessPoolExecutor(max_workers=self.workers, mp_context=ctx) as pool:
            # the server is up before the first request: start and warm every worker
            for f in [pool.submit(_warm) for _ in range(self.workers)]:
                f.result()
            t0 = time.perf_counter()
            now = lambda: time.perf_counter() - t0   # noqa: E731
            while len(results) < len(reqs):
                t = now()
                # arrivals go to the preparation pool
                while nxt < len(reqs) and reqs[nxt].arrival <= t:
                    futures[reqs[nxt].rid] = (reqs[nxt], pool.submit(pack_prepared, reqs[nxt].case))
                    nxt += 1
                for rid in [k for k, (_, f) in futures.items() if f.done()]:
                    req, f = futures.pop(rid)
                    packed = f.result()
                    fp = host.footprint(abi.params_from_dict(packed[0]))
                    if fp > self.budget:
                        raise ValueError(f"request {rid} needs {fp} bytes, more than the budget")
                    prepared.append((req, packed, fp))
                prepared.sort(key=lambda x: x[0].arrival)
                # admission by footprint, in arrival order
                while prepared and len(active) < self.max_active and used + prepared[0][2] <= self.budget:
                    req, packed, fp = prepared.pop(0)
                    sc = Scenario(req.case, wind=req.wind, prepared=packed)
                    sc.model.set_pcg_rows(self.pcg_rows)
                    batch.add(sc.model)
                    used += fp
                    active.append(_Active(req, sc, now(), fp))
                    if log:
                        log(f"[serve] t={now():.3f}s admit {req.rid} ({sc.params.maxele} cells, {fp >> 20} MiB)")
                if not active:
                    time.sleep(0.0005)
                    continue
                # one step of every active scenario
                for a in active:
                    a.sc.stage_forcing()
                batch.step()
                batch.sync()
                g, so = batch.layout()
                most_groups, most_solo = max(most_groups, g), max(most_solo, so)
                for a in active:
                    a.steps += 1
                done = [a for a in active if a.steps >= a.req.nsteps]
                for a in done:
                    batch.remove(a.sc.model)
                    active.remove(a)
                    vol, salt = a.sc.model.budgets()
                    state = {}
                    if a.req.keep_state:
                        s = abi.state_arrays(a.sc.params)
                        a.sc.model.get_state(s, only=list(a.req.keep_state))
                        state = {k: s[k].copy() for k in a.req.keep_state}
                    cl = _active_cell_layers(a.sc)
                    fin = now()
                    results.append(Result(a.req.rid, a.sc.params.maxele, a.sc.params.maxlayer, a.req.nsteps,
                                          a.req.arrival, a.admitted, fin, a.fp, cl * a.req.nsteps, vol, salt, state))
                    a.sc.model.close()
                    used -= a.fp
                    if log:
                        log(f"[serve] t={fin:.3f}s done {a.req.rid} latency {fin - a.req.arrival:.3f}s")
        batch.close()
        makespan = max(r.completed for r in results) - min(r.arrival for r in results)
        lat = np.array([r.latency for r in results])
        work = sum(r.active_cell_layer_steps for r in results)
        summary = {"requests": len(results), "p50_latency_s": float(np.percentile(lat, 50)),
                   "p99_latency_s": float(np.percentile(lat, 99)), "makespan_s": makespan,
                   "cell_layer_steps_per_s": work / makespan if makespan > 0 else 0.0,
                   "most_fused_pcg_launches": most_groups, "most_lone_pcg_solves": most_solo}
        return sorted(results, key=lambda r: r.rid), summary


def _warm() -> int:
    """Imports the preparation code in a worker (run once per worker before serving)."""
    from . import driver, meshgen, prep  # noqa: F401
    return 0


def poisson_trace(n: int, rate: float, seed: int = 0) -> List[float]:
    """Arrival times of n requests of a Poisson process with the given rate [1/s]."""
    rng = np.random.default_rng(seed)
    return list(np.cumsum(rng.exponential(1.0 / rate, size=n)) - 0.0)


def estuary_of_size(cells: float, nlayer: int, seed: int, **kw) -> Case:
    """A synthetic estuary of about `cells` cells (the benchmark estuary family)."""
    from . import meshgen
    nrep = max(3, int(round(np.sqrt(cells / 20.0))))
    ncol = max(10, int(round(cells / (5.0 * nrep))))
    return meshgen.estuary_case(ncol=ncol, nrep=nrep, nlayer=nlayer, seed=seed, **kw)
