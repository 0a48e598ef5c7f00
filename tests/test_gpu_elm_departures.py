"""Departure cells and levels of the device's momentum ELM against the oracle's.

The device records, per face-layer, the departure cell (negated for a point within 1e-12 of
an edge of its cell) and level of its backtrace (gomgpu_elm_record / gomgpu_elm_departures),
and counts the points within 1e-12 of an edge. Off the counted points the cells and levels
must equal the oracle's exactly; the counted points are listed (their cell may differ only
through the last bit of a coordinate, which FP64 arithmetic in the same order does not
change, so they are expected to agree as well) and the two counts must be equal. Cases:
every reference fixture from its reference input state, the uniform-flow channel whose
counts are known (tests/test_elm_near_edge.py), and a 20k-cell storm-surge case after 20
device steps.
"""
from __future__ import annotations

import numpy as np
import pytest

from golden_io import FIXTURES, Golden, apply
from oracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu_available():
    from gom_release_amd import host
    if host.device_count() < 1:
        pytest.fail("no GPU visible: the HIP library cannot run")
    return True


def _compare(tag, dev, ora, dev_count, ora_count):
    dn, dj = dev
    on, oj = ora
    near = (dn < 0) | (on < 0)
    off = ~near
    bad_n = np.flatnonzero((dn != on) & off)
    bad_j = np.flatnonzero((dj != oj) & off)
    diff_near = int(np.count_nonzero((np.abs(dn) != np.abs(on))[near] | (dj != oj)[near]))
    print(f"[elm-dep] {tag}: {int(np.count_nonzero(on))} departures, {int(near.sum())} within 1e-12 of an edge "
          f"(device count {dev_count}, oracle {ora_count}), {diff_near} of them differ", flush=True)
    assert bad_n.size == 0, (tag, "cells differ", bad_n[:5])
    assert bad_j.size == 0, (tag, "levels differ", bad_j[:5])
    assert np.array_equal(dn < 0, on < 0), tag
    assert dev_count == ora_count, (tag, dev_count, ora_count)


@pytest.mark.parametrize("name", FIXTURES)
def test_departures_match_oracle_on_fixtures(gpu_available, name):
    from gom_release_amd import host
    g = Golden(name)
    case = g.case()
    model = host.GpuModel(g.mesh, g.params)
    model.elm_record(True)
    for it in g.steps:
        state = g.state_in(it)
        apply(state, g.expected(it, "pre"))
        forcing = g.forcing(it, case)
        model.set_forcing(forcing)
        model.set_state(state)
        model.elm_near_edge(reset=True)
        model.stage("elm")
        model.sync()
        dev = model.elm_departures()
        dev_count = model.elm_near_edge()[0]
        Oracle.elm_near_edge(reset=True)
        rc, on, oj = Oracle(g.mesh, g.params, state, forcing).elm_departures()
        assert rc == 0
        _compare(f"{name} step {it}", dev, (on, oj), dev_count, Oracle.elm_near_edge()[0])
    model.close()


@pytest.mark.parametrize("rel,expect", [(0.0, 120), (1e-13, 120), (1e-11, 0)])
def test_uniform_flow_counts(gpu_available, rel, expect):
    from gom_release_amd import host
    from test_elm_near_edge import DT, DX, uniform_channel
    o, p, m, s = uniform_channel(rel * DX / DT)
    model = host.GpuModel(m, p)
    model.set_forcing(o.forcing)
    model.set_state(s)
    model.elm_record(True)
    model.stage("elm")
    model.sync()
    dev = model.elm_departures()
    count = model.elm_near_edge()[0]
    Oracle.elm_near_edge(reset=True)
    rc, on, oj = o.elm_departures()
    assert rc == 0
    _compare(f"uniform channel, cross-flow {rel:g}", dev, (on, oj), count, Oracle.elm_near_edge()[0])
    assert count == expect
    model.close()


def test_departures_after_device_steps(gpu_available):
    from gom_release_amd import abi, meshgen
    from gom_release_amd.driver import Scenario
    case = meshgen.storm_surge_case(ncol=127, nrep=32, dx=773.6, ndt=20)
    sc = Scenario(case)
    sc.run(20)
    state = abi.state_arrays(sc.params)
    sc.model.get_state(state)
    sc.model.elm_record(True)
    sc.model.elm_near_edge(reset=True)
    sc.model.stage("elm")
    sc.model.sync()
    dev = sc.model.elm_departures()
    count = sc.model.elm_near_edge()[0]
    Oracle.elm_near_edge(reset=True)
    rc, on, oj = Oracle(sc.mesh, sc.params, state, sc.forcing).elm_departures()
    assert rc == 0
    assert np.count_nonzero(on) > 1000
    _compare("storm surge 20k cells after 20 steps", dev, (on, oj), count, Oracle.elm_near_edge()[0])
    sc.model.close()


def test_tracer_elm_counts(gpu_available):
    """The tracer ELM (transport_solver 2) counts its departures near an edge like the oracle."""
    from gom_release_amd import host
    g = Golden("elm_estuary")
    case = g.case()
    model = host.GpuModel(g.mesh, g.params)
    for it in g.steps:
        state = g.state_in(it)
        for tag in ("pre", "elm", "mom", "fs", "vel", "lay", "nodev"):
            apply(state, g.expected(it, tag))
        forcing = g.forcing(it, case)
        model.set_forcing(forcing)
        model.set_state(state)
        model.elm_near_edge(reset=True)
        model.stage("transport")
        model.sync()
        dev_count = model.elm_near_edge()[1]
        Oracle.elm_near_edge(reset=True)
        assert Oracle(g.mesh, g.params, state, forcing).run("transport") == 0
        ora_count = Oracle.elm_near_edge()[1]
        print(f"[elm-dep] elm_estuary step {it} tracer ELM: device {dev_count}, oracle {ora_count}", flush=True)
        assert dev_count == ora_count
    model.close()
