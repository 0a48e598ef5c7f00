"""Contour and time-series output (gom_release_amd.gomplot) against the files the reference
build wrote (tests/golden/plots_estuary.npz, two runs; golden_io.plots_cases): the host driver
stepping the CPU oracle writes every Tecplot, VTK and time-series file byte for byte, through
the asynchronous writer thread and without it."""
from __future__ import annotations

import os

import pytest

from golden_io import PlotsGolden, plots_cases
from gom_release_amd import gomio, gomplot
from gom_release_amd.driver import Scenario

PREFIXES = ("tec2D_", "tec3D_surf_", "vtk2D_", "vtk3D_surf_", "eta_tser", "H_tser", "u_tser", "v_tser",
            "salt_tser", "temp_tser")


def _written(d):
    return {fn: open(os.path.join(d, fn), "rb").read() for fn in os.listdir(d) if fn.startswith(PREFIXES)}


@pytest.mark.parametrize("asynchronous", [True, False])
def test_driver_on_oracle_writes_the_reference_plot_files(tmp_path, asynchronous):
    from oracle_model import OracleModel
    g = PlotsGolden()
    for run, case in plots_cases().items():
        d = str(tmp_path / run)
        sc = Scenario(case, model_factory=OracleModel)
        sc.open_outputs(d, asynchronous=asynchronous)
        sc.run(case.ndt)
        sc.close_outputs()
        want, got = g.files(run), _written(d)
        assert sorted(got) == sorted(want), run
        for fn in want:
            assert got[fn] == want[fn], (run, fn)


def test_fortran_f0_edit():
    # the reference runtime's F0.5: no zero before the point, ties to even on the exact value
    assert gomplot.f0(0.0) == ".00000"
    assert gomplot.f0(-0.25) == "-.25000"
    assert gomplot.f0(1200.0) == "1200.00000"
    assert gomplot.f0(0.015625) == ".01562"        # 2^-6: an exact tie, to even
    assert gomplot.f0(0.046875) == ".04688"        # 3 * 2^-6: an exact tie, to even
    assert gomplot.f0_row([0.5, -1.5, 2.0]) == ".50000 -1.50000 2.00000"


def test_options_outside_the_writers_are_refused(tmp_path):
    case = plots_cases()["A"]
    from gom_release_amd import prep
    p, mesh, _ = prep.build(case)
    g = gomplot.OutputMesh(mesh, p)
    with pytest.raises(NotImplementedError):
        gomplot.ContourWriter(str(tmp_path), "3D", gomio.ContourSpec(full=1), g)
    with pytest.raises(NotImplementedError):
        gomplot.TserWriter(str(tmp_path), gomio.TserSpec(stations=[(1, 1, "a")], fields=(1, 0, 0, 0, 0, 0, 1)), g)
    with pytest.raises(ValueError):
        gomplot.TserWriter(str(tmp_path), gomio.TserSpec(stations=[(p.maxele + 1, 1, "a")]), g)
