"""GOM on-disk formats: the project reader/writer (gom_release_amd.gomio) and the restart
and dump writers (gom_release_amd.gomout).

Pinned against files the reference itself wrote (tests/golden/formats_estuary.npz, made by
the GOM v1.1.0 build driven by oracle/ref_harness.f90): the restart and dump files must be
reproduced byte for byte from the states the reference wrote them from, a restart read
back must give those states exactly, and a hot start must give the reference's start-up
state. The list-directed reader is checked on known answers of the Fortran rules, and on
the reference's own Mobile Bay project when /root/reference is present.
"""
from __future__ import annotations

import dataclasses
import os
import tempfile

import numpy as np
import pytest

from golden_io import FormatsGolden, formats_cases
from gom_release_amd import forcing as fz
from gom_release_amd import gomio, gomout, meshgen, prep

MB = "/root/reference/GOM_Rv1.1.0/Projects/MB_test/2_baroclinic_test"


@pytest.fixture(scope="module")
def fx():
    return FormatsGolden()


# ----------------------------------------------------------------- list-directed input

def _rd(text):
    return gomio.ListReader(text.splitlines(), "t")


def test_list_directed_known_answers():
    rd = _rd("  1, 2.5d-3  'it''s'  ! rest ignored\n 7\n")
    assert rd.read("irc") == [1, 2.5e-3, "it's"]
    assert rd.read("i") == [7]
    # values continue on the next records; the rest of the last record is dropped
    rd = _rd("1\n\n2 3 4\n5\n")
    assert rd.read("iii") == [1, 2, 3]
    assert rd.read("i") == [5]
    # repeat counts, null values and the slash terminator
    assert _rd("3*1.5 2*\n").read("rrrrr") == [1.5, 1.5, 1.5, None, None]
    assert _rd("1,,3\n").read("iii") == [1, None, 3]
    assert _rd("1 / 2 3\n").read("iii") == [1, None, None]
    # Fortran exponent forms
    assert _rd("1.0D2 -3.5-2 .5 1.e-4\n").read("rrrr") == [100.0, -0.035, 0.5, 1e-4]
    with pytest.raises(gomio.FormatError):
        _rd("1.5\n").read("i")
    with pytest.raises(gomio.FormatError):
        _rd("1\n").read("ii")


def test_decimal_to_binary_is_correctly_rounded():
    # the tie-breaking cases of decimal conversion that a naive parser gets wrong
    vals = ["0.1", "2.2250738585072011e-308", "9007199254740993", "1.7976931348623157e308",
            "5e-324", "0.30000000000000004"]
    got = _rd(" ".join(vals)).read("r" * len(vals))
    assert [float.hex(g) for g in got] == [float.hex(float(v)) for v in vals]


def test_comment_lines_only_before_a_card():
    """read_main_inp.f90 skips comment lines before a card only: a comment line between
    the data lines of one card is read as data, and the reference stops on it."""
    case = meshgen.channel_case(nx=3, ny=2, nlayer=2, ndt=1)
    with tempfile.TemporaryDirectory() as d:
        gomio.write_project(case, d)
        path = os.path.join(d, "input", "main.inp")
        lines = open(path).read().splitlines()
        i = lines.index("C21\ttide_spinup tide_spinup_period / baroclinic_spinup baroclinic_spinup_period")
        good = lines[:i] + ["! extra comment", "C another"] + lines[i:]
        open(path, "w").write("\n".join(good) + "\n")
        gomio.read_main_inp(path)
        bad = lines[:i + 2] + ["! between the two data lines of C21"] + lines[i + 2:]
        open(path, "w").write("\n".join(bad) + "\n")
        with pytest.raises(gomio.FormatError):
            gomio.read_main_inp(path)


# ------------------------------------------------------------------------ round trips

def _cases():
    ch = meshgen.channel_case(nx=5, ny=3, nlayer=3, ndt=7)
    es = meshgen.estuary_case(ncol=6, nrep=2, nlayer=4, ndt=5, river_q=40.0, dry_land=True,
                              tide_phase=-90.0, elm_backtrace_flag=1, elm_max_iter=9)
    # elevation-series boundaries, shifted series times, output cards, odd constants
    ts = meshgen.estuary_case(ncol=6, nrep=2, nlayer=3, ndt=3, river_q=25.0)
    for k, b in enumerate(ts.open_boundaries):
        b.eta_type, b.harmonic_id, b.eta_ser_id = 2, 0, 1 + k % 2
    ts.harmonics = []
    ts.eta_series = [gomio.Series([0.0, 0.5, 1.0], [0.1, -0.2, 0.3], 86400.0, 3.0, 0.01, 1.5),
                     gomio.Series([-1.0, 40.0], [0.0, 0.25], 3600.0, 0.0, 1.0, 0.0)]
    # the interpolation column round-trips; the last series read sets the reference's global
    # q_interp_method (read_q_ser.f90:46), which stays linear
    ts.q_series = [gomio.Series([0.0, 10.0], [20.0, 30.0], 86400.0, 0.0, 1.0, 0.0, 2),
                   gomio.Series([0.0, 10.0], [5.0, 6.0], 86400.0, 0.0, 1.0, 0.0, 1)]
    ts.start = (2012, 3, 5, 6, 30, 0)
    ts.data_start_year, ts.data_time_shift = 2010, 3
    ts.gravity, ts.rho_o, ts.rho_a, ts.max_no_neighbor_node = 9.80665, 1024.5, 1.225, 12
    ts.h_node_adjust = 0.25
    ts.restart_out, ts.restart_freq = 1, 2
    ts.dump2d = gomio.DumpSpec(0, 1, 4, 2, 1, 3)
    ts.dump3d = gomio.DumpSpec(1, 3, 10, 5, 2, 9)
    # the ELM transport solver: C7 names solver 2 and no TVD card C7_1 follows
    el = meshgen.estuary_case(ncol=6, nrep=2, nlayer=4, ndt=3, river_q=40.0, dry_land=True,
                              transport_solver=2)
    return [ch, es, ts, el]


def _assert_same_case(a: gomio.Case, b: gomio.Case):
    for f in dataclasses.fields(gomio.Case):
        if a.transport_solver == 2 and f.name in ("trans_sub_iter", "h_flux_limiter", "v_flux_limiter"):
            continue    # TVD settings: the format carries them for solver 1 only
        va, vb = getattr(a, f.name), getattr(b, f.name)
        if isinstance(va, np.ndarray) or isinstance(vb, np.ndarray):
            assert np.array_equal(np.asarray(va), np.asarray(vb)), f.name
        elif f.name in ("harmonics",):
            assert [(s, [tuple(c) for c in cons]) for s, cons in va] == \
                   [(s, [tuple(c) for c in cons]) for s, cons in vb], f.name
        else:
            assert va == vb, f.name


@pytest.mark.parametrize("k", range(4))
def test_project_round_trip(k):
    case = _cases()[k]
    with tempfile.TemporaryDirectory() as d:
        gomio.write_project(case, d)
        back = gomio.read_case(d)
    _assert_same_case(case, back)


def test_hot_start_project_round_trip(fx):
    a, b_of = formats_cases()
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "r.out")
        open(p, "wb").write(fx.files("A")["restart_0000004.out"])
        b = b_of(gomout.read_restart(p, a.maxele, a.maxnod, a.maxlayer))
        gomio.write_project(b, d)
        assert open(os.path.join(d, "input", "restart.inp"), "rb").read() == \
            fx.files("A")["restart_0000004.out"]
        back = gomio.read_case(d)
    assert back.restart_in == 1
    for f in gomout.RESTART_FIELDS:
        assert np.array_equal(back.hot_start[f], b.hot_start[f]), f
    # a hot start does not read the initial tracer files (read_main_inp.f90:316-321)
    assert back.salt_init is None
    back.hot_start, back.salt_init = b.hot_start, b.salt_init
    _assert_same_case(b, back)


# -------------------------------------------------------------- restart and dumps

def test_restart_bytes_match_reference(fx):
    for run, files in (("A", {"restart_0000004.out": 4}),
                       ("B", {"restart_0000001.out": 1, "restart_0000002.out": 2})):
        ref = fx.files(run)
        for fn, it in files.items():
            st = fx.state(run, f"s{it}_post")
            assert gomout.restart_bytes(st) == ref[fn], (run, fn)


def test_restart_read_back_exact(fx):
    a, _ = formats_cases()
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "restart.inp")
        open(p, "wb").write(fx.files("A")["restart_0000004.out"])
        got = gomout.read_restart(p, a.maxele, a.maxnod, a.maxlayer)
        st = fx.state("A", "s4_post")
        for f in gomout.RESTART_FIELDS:
            assert np.array_equal(got[f], st[f]), f
        # framing: every record holds its byte count before and after the data
        recs = gomout.read_records(p)
        J = got["un_face"].size // a.maxlayer
        assert len(recs) == 3 * a.maxele + 4 * a.maxnod + 2 * J
        # a damaged marker is refused
        raw = bytearray(open(p, "rb").read())
        raw[4 + 16] ^= 1                     # trailing marker of the first record
        open(p, "wb").write(bytes(raw))
        with pytest.raises(ValueError):
            gomout.read_restart(p, a.maxele, a.maxnod, a.maxlayer)


def _drive_dumps(fx, run, case, steps):
    """Run the DumpWriter schedule over the fixture's per-step states."""
    out = {}
    with tempfile.TemporaryDirectory() as d:
        writers = [gomout.DumpWriter(d, kind, spec, case.maxlayer)
                   for kind, spec in (("2D", case.dump2d), ("3D", case.dump3d))]
        for it in range(1, steps + 1):
            st = fx.state(run, f"s{it}_post")
            for w in writers:
                w.step(it, float(st["julian_day"][0]), lambda names: st)
        for w in writers:
            w.close()
        for fn in os.listdir(d):
            out[fn] = open(os.path.join(d, fn), "rb").read()
    return out


def test_dump_files_match_reference(fx):
    a, b_of = formats_cases()
    got = _drive_dumps(fx, "A", a, a.ndt)
    ref = {k: v for k, v in fx.files("A").items() if k.startswith("dump")}
    assert sorted(got) == sorted(ref) == ["dump2D_0001.dat", "dump2D_0002.dat", "dump3D_0001.dat"]
    for fn in ref:
        assert got[fn] == ref[fn], fn
    b = b_of(None)
    got = _drive_dumps(fx, "B", b, b.ndt)
    ref = {k: v for k, v in fx.files("B").items() if k.startswith("dump")}
    assert sorted(got) == sorted(ref)
    for fn in ref:
        assert got[fn] == ref[fn], fn


def test_dump_read_back(fx):
    a, _ = formats_cases()
    L, P = a.maxlayer, a.maxnod
    with tempfile.TemporaryDirectory() as d:
        for fn, binary, kind in (("dump2D_0001.dat", 0, "2D"), ("dump3D_0001.dat", 1, "3D")):
            p = os.path.join(d, fn)
            open(p, "wb").write(fx.files("A")[fn])
            dumps = gomout.read_dump(p, kind, P, L, binary)
            its = [1, 2] if kind == "2D" else [2, 4]
            assert len(dumps) == len(its)
            for (t, rows), it in zip(dumps, its):
                st = fx.state("A", f"s{it}_post")
                want = gomout.dump_rows(kind, st, L)
                if binary:
                    assert np.array_equal(rows, want)
                else:   # five significant digits
                    np.testing.assert_allclose(rows, want, rtol=5e-5, atol=0)


def test_fortran_edit_descriptors():
    e = gomout.fortran_e
    assert e(1.0) == "  0.10000E+0001"
    assert e(-1.0) == " -0.10000E+0001"
    assert e(0.0) == "  0.00000E+0000"
    assert e(123456.789) == "  0.12346E+0006"
    assert e(9.999995) == "  0.10000E+0002"          # carries into the exponent
    assert e(2.5e-7) == "  0.25000E-0006"
    assert e(1e-300) == "  0.10000E-0299"
    assert e(1.03125) == "  0.10312E+0001"           # exact tie: to even
    assert e(1.03375) == "  0.10337E+0001"           # below the tie in binary
    assert gomout.fortran_f(212.5) == "      212.50000"
    assert gomout.fortran_f(18316800.0) == " 18316800.00000"


def test_hot_start_matches_reference_startup(fx):
    """prep.build with restart_in = 1 against the reference's state after prepare_gom of a
    hot start from the same restart file (layers from the surface before the restart is
    read, restart fields, untouched allocation values elsewhere)."""
    a, b_of = formats_cases()
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "restart.inp")
        open(p, "wb").write(fx.files("A")["restart_0000004.out"])
        b = b_of(gomout.read_restart(p, a.maxele, a.maxnod, a.maxlayer))
    _p, _m, st = prep.build(b)
    ref = fx.state("B", "s0_post")
    skip = {"rho_cell", "rho_node", "rho_face", "ubar_node", "vbar_node", "sbar_node", "tbar_node",
            "rbar_node", "Gamma_B", "Cdb", "q2", "xl"}
    n = 0
    for f in st.arr:
        if f in skip or f not in ref:
            continue
        assert np.array_equal(st[f], ref[f]), f
        n += 1
    assert n > 40


# ------------------------------------------------------------------------ forcing

def test_reference_diff_days():
    c = meshgen.channel_case(nx=2, ny=2, nlayer=2, ndt=1)
    c.start = (2010, 8, 1, 0, 0, 0)
    assert c.reference_diff_days() == 0
    c.data_time_shift = 212
    assert c.reference_diff_days() == -212
    c.data_start_year = 2007      # 2007 (365) + 2008 (divisible by 4: 364) + 2009 (365)
    assert c.reference_diff_days() == 365 + 364 + 365 - 212


def test_elevation_series_boundaries():
    """compute_boundary_eta.f90:92-124 on a hand-made series: linear interpolation at
    julian_day*86400 - dt and julian_day*86400, times shifted by reference_diff_days."""
    c = _cases()[2]
    c.tide_spinup = 0
    s1, s2 = c.eta_series
    t1 = [(t + 3.0) * 86400.0 for t in s1.times_days]
    v1 = [(v + 1.5) * 0.01 for v in s1.values]
    rdd = c.reference_diff_days() * 86400.0
    jd = (t1[0] - rdd) / 86400.0 + 0.6          # 0.6 day into the first series
    old, new = fz.boundary_eta(c, 3, 0.5, jd)

    def lin(t, v, u):
        for i in range(1, len(t)):
            if t[i - 1] - rdd <= u <= t[i] - rdd:
                return v[i - 1] + (u - (t[i - 1] - rdd)) * (v[i] - v[i - 1]) / (t[i] - t[i - 1])
        raise AssertionError("not covered")
    for i, b in enumerate(c.open_boundaries):
        if b.eta_ser_id == 1:
            np.testing.assert_allclose(old[i], 0.5 * lin(t1, v1, jd * 86400.0 - c.dt), rtol=1e-14)
            np.testing.assert_allclose(new[i], 0.5 * lin(t1, v1, jd * 86400.0), rtol=1e-14)


# ----------------------------------------------------- the reference's own project

@pytest.mark.skipif(not os.path.isdir(MB), reason="the reference's example project is not here")
def test_reads_the_reference_project():
    """Mobile Bay (GOM_Rv1.1.0/Projects/MB_test/2_baroclinic_test/input): values read off
    its main.inp, node.inp, cell.inp and series files by eye."""
    pr = gomio.read_project(MB)
    m = pr.main
    assert m.project_name == "Mobile_Bay:_GOM_v84_synchronized"
    assert (m.maxnod, m.maxele, m.maxlayer) == (18718, 35127, 5)
    assert m.z_level == [-100.0, -12.0, -9.0, -6.0, -3.0, 20.0]
    assert (m.dt, m.ndt, m.data_start_year, m.data_time_shift) == (150.0, 576, 2010, 212)
    assert (m.max_iteration_pcg, m.error_tolerance) == (1000, 5e-6)
    assert (m.smagorinsky_parameter, m.Ah_0, m.Kh_0, m.Av_0, m.Kv_0) == (0.021, 0.0, 0.0, 0.0, 1e-4)
    assert (m.trans_sub_iter, m.is_tran, m.num_tran_ser) == (30, [1, 1], [3, 3])
    assert (m.num_ob_cell, m.num_eta_bc, m.num_eta_ser, m.eta_ser_shape) == (75, 75, 2, 2)
    assert m.ob_rows[0] == [1, 18718, 18716, 2, 0, 1, 1, 1]
    assert m.ob_rows[19] == [20, 16941, 16940, 2, 0, 2, 2, 2]
    assert m.q_rows == [[1, 151, 141, 1, 1.0, 3, 3]]
    assert (m.hurricane_flag, m.hurricane_dt, m.hurricane_start) == (1, 21600, (2010, 8, 1))
    assert [s[3] for s in m.tser_stations] == ["MSD", "DPI", "MBL", "BOS", "MEP"]
    assert pr.nodes.shape == (18718, 3) and pr.cells.shape == (35127, 4)
    assert np.array_equal(pr.nodes[0], [0.4053641000e06, 0.3431161830e07, 0.555e01])
    assert list(pr.cells[0]) == [2, 1, 3, 0] and not pr.cells[:, 3].any()
    assert pr.salt_init.shape == (18718, 5)
    assert [len(s.times_days) for s in pr.eta_series] == [1464, 1464]
    assert pr.eta_series[0].values[:2] == [-0.011, -0.022]
    assert [s.values[0] for s in pr.salt_series] == [32.2, 35.3, 35.3]
    assert pr.q_series[0].values[0] == 287.45 and pr.q_series[0].interp == 1
    bad = gomio.unsupported(pr)
    assert bad == ["C7_3 surface heat flux (heat_option with air_ser.inp)",
                   "C31 hurricane fields: input/hurricane_ser.inp is missing"]
    with pytest.raises(NotImplementedError):
        gomio.case_from_project(pr)
    case, dropped = gomio.case_from_project(pr, strict=False)
    assert dropped == bad and case.is_temp == 0 and len(case.open_boundaries) == 75
    assert case.hurricane_flag == 0


# ------------------------------------------- the host driver's output files, on the oracle

def test_driver_on_oracle_writes_the_reference_files(fx):
    """The host driver (forcing, output schedule, writers) stepping the CPU oracle writes
    byte for byte the files the reference wrote: run A from a cold start, run B from a hot
    start on the reference's restart file."""
    from oracle_model import OracleModel
    from gom_release_amd.driver import Scenario
    a, b_of = formats_cases()
    with tempfile.TemporaryDirectory() as d:
        rp = os.path.join(d, "restart.inp")
        open(rp, "wb").write(fx.files("A")["restart_0000004.out"])
        b = b_of(gomout.read_restart(rp, a.maxele, a.maxnod, a.maxlayer))
        for run, case in (("A", a), ("B", b)):
            out = os.path.join(d, run)
            sc = Scenario(case, model_factory=OracleModel)
            sc.open_outputs(out)
            sc.run(case.ndt)
            sc.close_outputs()
            ref = fx.files(run)
            assert sorted(os.listdir(out)) == sorted(ref)
            for fn in ref:
                assert open(os.path.join(out, fn), "rb").read() == ref[fn], (run, fn)
