"""Host forcing of the transport boundaries and the river discharge, against values worked
out by hand from the reference's loops.

solve_transport_equation_v19.f90:229-397 interpolates every salt/temperature series at
u2 = (julian_day*86400 - dt) + (subcycle+1)*dt2 with the series times shifted by
reference_diff_days*86400, in a local v2 that carries over between boundaries, tracers and
sub-cycles when a series does not bracket u2. calculate_Q.f90:21-45 does the same for the
river discharge, with Q_add keeping its previous value.
"""
from __future__ import annotations

import numpy as np

from gom_release_amd import forcing as fz
from gom_release_amd import gomio, meshgen


def _case():
    case = meshgen.estuary_case(ncol=6, nrep=2, nlayer=3, ndt=3, river_q=25.0)
    case.trans_sub_iter = 2
    case.start = (2012, 3, 5, 0, 0, 0)
    case.data_start_year, case.data_time_shift = 2011, 10    # 365 - 10 = 355 days
    return case


def test_reference_diff_days_of_the_test_case():
    assert _case().reference_diff_days() == 355


def test_tvd_boundary_concentrations_shift_and_carry():
    case = _case()
    rdd = 355 * 86400.0
    jd = 64.5                                   # julian day of the step's end
    u2 = [(jd * 86400.0 - case.dt) + (s + 1) * case.dt / 2 for s in range(2)]
    # series 1 covers the step after the shift; series 2 ends before it (no bracket)
    t0 = (u2[0] + rdd - 3600.0) / 86400.0
    case.salt_series = [gomio.Series([t0, t0 + 1.0], [10.0, 34.0]),
                        gomio.Series([0.0, 1.0], [99.0, 99.0])]
    nob = len(case.open_boundaries)
    assert nob >= 3
    for b in case.open_boundaries:
        b.salt_ser_id = 1
    case.open_boundaries[1].salt_ser_id = 2     # keeps the v2 of boundary 0
    case.open_boundaries[2].salt_ser_id = 0     # never set in the call: 0
    for r in case.rivers:
        r.salt_ser_id = 2                       # keeps the v2 of the last open boundary
    sp = 0.75
    ob, qb = fz.tvd_boundary_concentrations(case, jd, sp, 1, [1, 0])
    ob = ob.reshape(2, 1, nob, order="F")
    qb = qb.reshape(2, 1, max(len(case.rivers), 1), order="F")
    for s in range(2):
        want = 10.0 + 24.0 * (u2[s] - (t0 * 86400.0 - rdd)) / 86400.0
        assert np.isclose(ob[s, 0, 0], sp * want, rtol=1e-14)
        assert ob[s, 0, 1] == ob[s, 0, 0]
        assert ob[s, 0, 2] == 0.0
        assert np.all(ob[s, 0, 3:] == ob[s, 0, 0])
        assert np.all(qb[s, 0, :len(case.rivers)] == ob[s, 0, 0])
    # without the shift the same step falls outside series 1: every value carries the
    # start-of-call v2 (0)
    case.data_start_year, case.data_time_shift = None, 0
    ob0, _ = fz.tvd_boundary_concentrations(case, jd, sp, 1, [1, 0])
    assert np.all(ob0 == 0.0)


def test_river_discharge_shift_and_previous_value():
    case = _case()
    rdd = 355 * 86400.0
    jd = 64.5
    t0 = (jd * 86400.0 + rdd - 7200.0) / 86400.0
    case.q_series = [gomio.Series([t0, t0 + 0.5], [100.0, 200.0])]
    q = fz.river_discharge(case, jd)
    want = 100.0 + 100.0 * 7200.0 / (0.5 * 86400.0)
    for i, r in enumerate(case.rivers):
        assert np.isclose(q[i], want * r.portion, rtol=1e-14)
    # a later step outside the series keeps the previous Q_add
    q2 = fz.river_discharge(case, jd + 3.0, q)
    assert np.array_equal(q2, q)
    # with no previous value (start of the run) the allocation value 0 stays
    assert np.all(fz.river_discharge(case, jd + 3.0) == 0.0)
