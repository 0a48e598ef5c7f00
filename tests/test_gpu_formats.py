"""Output files and hot start on the device against the files the reference wrote.

Run A of tests/golden/formats_estuary.npz is stepped on the GPU with its output cards
(restart at the last step, 2-D formatted dumps rolling over after two dumps, 3-D
unformatted dumps every second step); run B starts hot from the reference's own restart
file. The files written from the device state must have the reference's names, sizes and
record framing, and their values must agree with the reference's within the tolerances
the whole-step tests use (the free-surface PCG sums in another order: eta within 1e-9
relative). Formatted values carry five significant digits and are compared at that
precision. The start-up state of the hot start must equal the reference's exactly.
"""
from __future__ import annotations

import os
import tempfile

import numpy as np
import pytest

from golden_io import FormatsGolden, formats_cases
from gom_release_amd import gomout

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1e-7, 1e-10


@pytest.fixture(scope="module")
def gpu_available():
    from gom_release_amd import host
    if host.device_count() < 1:
        pytest.fail("no GPU visible: the HIP library cannot run")
    return True


def _compare_files(d, ref_files, case):
    L, P, E = case.maxlayer, case.maxnod, case.maxele
    got = sorted(f for f in os.listdir(d) if f.startswith(("restart_", "dump")))
    assert got == sorted(ref_files)
    for fn in got:
        mine = open(os.path.join(d, fn), "rb").read()
        ref = ref_files[fn]
        print(f"[gpu-formats] {fn}: {len(mine)} bytes", flush=True)
        if fn.startswith("restart_"):
            assert len(mine) == len(ref), fn
            rp = os.path.join(d, "_ref_" + fn)
            open(rp, "wb").write(ref)
            a = gomout.read_restart(os.path.join(d, fn), E, P, L)
            b = gomout.read_restart(rp, E, P, L)
            os.remove(rp)
            for f in gomout.RESTART_FIELDS:
                tol = (1e-9, 1e-12) if f.startswith("eta") else (RTOL, ATOL)
                np.testing.assert_allclose(a[f], b[f], rtol=tol[0], atol=tol[1], err_msg=f"{fn} {f}")
            continue
        kind = "2D" if fn.startswith("dump2D") else "3D"
        binary = 0 if mine[:1] == b" " else 1
        rp = os.path.join(d, "_ref_" + fn)
        open(rp, "wb").write(ref)
        a = gomout.read_dump(os.path.join(d, fn), kind, P, L, binary)
        b = gomout.read_dump(rp, kind, P, L, binary)
        os.remove(rp)
        assert len(a) == len(b), fn
        if binary:
            assert len(mine) == len(ref), fn
        for (ta, ra), (tb, rb) in zip(a, b):
            assert ta == tb, fn
            if binary:
                np.testing.assert_allclose(ra, rb, rtol=RTOL, atol=ATOL, err_msg=fn)
            else:
                np.testing.assert_allclose(ra, rb, rtol=2e-5, atol=1e-12, err_msg=fn)


def test_gpu_output_files_match_reference(gpu_available):
    from gom_release_amd.driver import Scenario
    fx = FormatsGolden()
    a, _ = formats_cases()
    print("[gpu-formats] run A: 4 steps with restart and dumps", flush=True)
    with tempfile.TemporaryDirectory() as d:
        sc = Scenario(a)
        sc.open_outputs(d)
        sc.run(a.ndt)
        sc.close_outputs()
        # the restart holds exactly the device state it was written from
        st = sc.fields(gomout.RESTART_FIELDS)
        assert open(os.path.join(d, "restart_0000004.out"), "rb").read() == gomout.restart_bytes(st)
        sc.model.close()
        _compare_files(d, fx.files("A"), a)


def test_gpu_hot_start_matches_reference(gpu_available):
    from gom_release_amd.driver import Scenario
    fx = FormatsGolden()
    a, b_of = formats_cases()
    with tempfile.TemporaryDirectory() as d:
        rp = os.path.join(d, "restart.inp")
        open(rp, "wb").write(fx.files("A")["restart_0000004.out"])
        b = b_of(gomout.read_restart(rp, a.maxele, a.maxnod, a.maxlayer))
    print("[gpu-formats] run B: hot start, 2 steps with restarts and dumps", flush=True)
    with tempfile.TemporaryDirectory() as d:
        sc = Scenario(b)
        ref0 = fx.state("B", "s0_post")
        s0 = sc.fields([f for f in ref0 if f in gomout.RESTART_FIELDS])
        for f in gomout.RESTART_FIELDS:
            if f.startswith("rho"):     # recomputed by the start-up EOS, as prepare_gom does
                np.testing.assert_allclose(s0[f], ref0[f], rtol=1e-13, atol=0, err_msg=f)
            else:
                assert np.array_equal(s0[f], ref0[f]), f
        sc.open_outputs(d)
        sc.run(b.ndt)
        sc.close_outputs()
        sc.model.close()
        _compare_files(d, fx.files("B"), b)
