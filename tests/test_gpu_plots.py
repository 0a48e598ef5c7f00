"""Contour and time-series output written from the device (asynchronous snapshots,
gomgpu_snapshot) against the files the reference build wrote (tests/golden/plots_estuary.npz):
the same files, line for line the same text, every number within the rounding of the
formatted digits of the device's state (which equals the reference's within the whole-step
tolerances, PCG summation order aside); and the asynchronous path writes the same bytes as
reading the state synchronously."""
from __future__ import annotations

import os
import re

import pytest

from golden_io import PlotsGolden, plots_cases
from gom_release_amd.driver import Scenario

pytestmark = pytest.mark.gpu

PREFIXES = ("tec2D_", "tec3D_surf_", "vtk2D_", "vtk3D_surf_", "eta_tser", "H_tser", "u_tser", "v_tser",
            "salt_tser", "temp_tser")
NUM = re.compile(rb"-?\d*\.\d+(?:E[+-]\d+)?")


def _written(d):
    return {fn: open(os.path.join(d, fn), "rb").read() for fn in os.listdir(d) if fn.startswith(PREFIXES)}


def _close(a: bytes, b: bytes) -> bool:
    x, y = float(a), float(b)
    return abs(x - y) <= 2.0e-5 * max(1.0, abs(y))


def _run(case, d, asynchronous):
    print(f"[gpu-plots] {case.name} -> {d} async={asynchronous}", flush=True)
    sc = Scenario(case)
    sc.open_outputs(d, asynchronous=asynchronous)
    sc.run(case.ndt)
    sc.close_outputs()
    sc.model.close()
    return _written(d)


def test_device_plot_files_match_the_reference(tmp_path):
    g = PlotsGolden()
    for run, case in plots_cases().items():
        got = _run(case, str(tmp_path / run), True)
        want = g.files(run)
        assert sorted(got) == sorted(want), run
        for fn, ref in want.items():
            la, lb = ref.split(b"\n"), got[fn].split(b"\n")
            assert len(la) == len(lb), (run, fn)
            for i, (x, y) in enumerate(zip(la, lb)):
                if x == y:
                    continue
                assert NUM.sub(b"#", x) == NUM.sub(b"#", y), (run, fn, i)
                nx, ny = NUM.findall(x), NUM.findall(y)
                assert len(nx) == len(ny) and all(_close(b_, a_) for a_, b_ in zip(nx, ny)), (run, fn, i)


def test_asynchronous_output_equals_synchronous(tmp_path):
    case = plots_cases()["A"]
    a = _run(case, str(tmp_path / "async"), True)
    b = _run(case, str(tmp_path / "sync"), False)
    assert a == b
