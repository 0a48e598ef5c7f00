"""Host side of batching and serving, without a GPU: the Poisson trace, the synthetic size
family, and the picklable preparation round trip used by the preparation workers."""
from __future__ import annotations

import pickle

import numpy as np

from gom_release_amd import driver, prep, serving


def test_poisson_trace_rate():
    t = serving.poisson_trace(20000, rate=4.0, seed=3)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert np.all(gaps > 0)
    assert abs(gaps.mean() - 0.25) < 0.01          # mean inter-arrival 1/rate
    assert abs(gaps.std() - 0.25) < 0.01           # exponential: std = mean


def test_estuary_of_size():
    for n in (10e3, 40e3, 120e3):
        c = serving.estuary_of_size(n, 10, seed=1)
        assert abs(c.maxele - n) / n < 0.05


def test_prepared_round_trip_is_exact():
    case = serving.estuary_of_size(3000, 5, seed=2, river_q=20.0)
    packed = pickle.loads(pickle.dumps(driver.pack_prepared(case)))
    p1, m1, s1 = driver.unpack_prepared(packed)
    p0, m0, s0 = prep.build(case)
    from gom_release_amd import abi
    assert abi.params_to_dict(p1) == abi.params_to_dict(p0)
    for a, b in ((m1, m0), (s1, s0)):
        for k in b.arr:
            assert np.array_equal(a[k], b[k]), k


def test_autotune_cache_round_trip(tmp_path):
    from gom_release_amd import abi, autotune
    p = abi.Params()
    p.maxele, p.maxlayer, p.turbulence_flag = 200200, 20, 3
    path = str(tmp_path / "launch.json")
    assert autotune.load(p, path) is None
    cfg = {"elm": 256, "mom_rhs": 512}
    autotune.save(p, cfg, path)
    assert autotune.load(p, path) == cfg
    q = abi.Params()
    q.maxele, q.maxlayer, q.turbulence_flag = 190000, 20, 3    # same size class
    assert autotune.load(q, path) == cfg
    q.maxlayer = 30
    assert autotune.load(q, path) is None
