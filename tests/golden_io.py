"""Loading of the golden fixtures (tests/golden/*.npz) into the ABI arrays."""
from __future__ import annotations

import json
import os
from typing import Dict, List

import numpy as np

from gom_release_amd import abi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STAGES = [("friction", "pre"), ("elm", "elm"), ("momentum", "mom"), ("freesurf", "fs"),
          ("velocity", "vel"), ("layers", "lay"), ("nodevel", "nodev"), ("transport", "tvd"),
          ("eos", "eos"), ("update", "post")]
FIXTURES = ["channel_16x8x5", "estuary_wetdry", "hurricane_estuary", "elm_estuary", "estuary_limiters",
            "estuary_superbee"]


def _flat(a):
    a = np.asarray(a)
    return np.asfortranarray(a).ravel(order="F") if a.ndim == 2 else a.ravel()


class Golden:
    def __init__(self, name: str):
        self.name = name
        z = np.load(os.path.join(GOLDEN, name + ".npz"))
        self.data = {k: z[k] for k in z.files}
        self.meta = json.load(open(os.path.join(GOLDEN, name + ".json")))
        self.steps: List[int] = self.meta["steps"]
        g = {k[5:]: v for k, v in self.data.items() if k.startswith("geom/")}
        self.geom = g
        import refdump  # oracle helper: dump dict -> ABI structs
        self.params = refdump.params_from_geom(g)
        self.mesh = refdump.mesh_from_geom(g, self.params)

    def group(self, tag: str) -> Dict[str, np.ndarray]:
        pre = tag + "/"
        return {k[len(pre):]: v for k, v in self.data.items() if k.startswith(pre)}

    def state_in(self, it: int) -> "abi.Arrays":
        s = abi.state_arrays(self.params)
        for k, v in self.group(f"s{it}_in").items():
            if k in s.arr:
                s[k] = _flat(v)
        return s

    def forcing(self, it: int, case=None) -> "abi.ForcingArrays":
        import refdump
        pre = self.group(f"s{it}_pre")
        ob = qb = None
        if case is not None:
            from gom_release_amd import forcing as fz
            p = self.params
            if p.transport_solver == 2:
                # the river values without a series keep their zero start in these fixtures
                ob, qb = fz.elm_boundary_concentrations(case, float(pre["julian_day"]),
                                                        float(pre["spinup_function_baroclinic"]))
            else:
                ob, qb = fz.tvd_boundary_concentrations(case, float(pre["julian_day"]),
                                                        float(pre["spinup_function_baroclinic"]),
                                                        p.maxtran2, [p.tran_id1, p.tran_id2])
        return refdump.forcing_from_dump(pre, self.params, ob, qb)

    def expected(self, it: int, tag: str) -> Dict[str, np.ndarray]:
        out = {}
        for k, v in self.group(f"s{it}_{tag}").items():
            out[k] = _flat(v)
        return out

    def pcg_iterations(self, it: int):
        """The reference's PCG iteration count of step `it` (cgit2, pre_conj_grad.f90:
        224-229), when the fixture recorded it."""
        v = self.meta.get("pcg_iterations", {}).get(str(it))
        return None if v is None else int(v)

    def case(self):
        return case_of(self.meta)

    def step_forcing(self, it: int, case) -> "abi.ForcingArrays":
        """The forcing a whole step `it` starts from: as `forcing`, but with the hurricane
        fields as they were before the step (the step's wind stage recomputes the wind
        stress and air pressure), the snapshot the step reads, and the step's phase."""
        f = self.forcing(it, case)
        if self.params.hurricane_flag != 1:
            return f
        from gom_release_amd import forcing as fz
        prev = self.group(f"s{it}_in")
        pre = self.group(f"s{it}_pre")
        for n in HURRICANE_FIELDS:
            f[n] = _flat(prev[n])
        f["wind_u_next"], f["wind_v_next"], f["air_p_next"] = (_flat(pre["wind_u2"]), _flat(pre["wind_v2"]),
                                                                _flat(pre["air_p2"]))
        f.scalars["hurricane_phase"] = float(fz.hurricane_phase(it, float(pre["julian_day"]),
                                                                fz.hurricane_window(case),
                                                                self.params.hurricane_read_interval))
        return f


HURRICANE_FIELDS = ["wind_u0", "wind_v0", "air_p0", "wind_u1", "wind_v1", "air_p1", "wind_u2",
                    "wind_v2", "air_p2", "wind_u_at_face", "wind_v_at_face"]
WIND_OUTPUTS = HURRICANE_FIELDS + ["Gamma_T", "wind_stress_normal", "wind_stress_tangnt"]


def case_of(spec):
    """The generated case of a fixture description (kind, args, elm_tracers, hurricane)."""
    from gom_release_amd import meshgen
    if spec["kind"] == "channel":
        case = meshgen.channel_case(**spec["args"])
    else:
        case = meshgen.estuary_case(**spec["args"])
    if spec.get("elm_tracers"):
        meshgen.elm_tracers(case, t=0.0, temp_input=False)
    if spec.get("stratify"):
        meshgen.stratify(case, **spec["stratify"])
    if spec.get("hurricane"):
        # a Holland cyclone (950 hPa, r_max 300 m) crossing the small estuary at 5 m/s,
        # fields every third step; the model starts at 23:53 so the window opens at step 4
        case.start = (2020, 1, 1, 23, 53, 0)
        case.hurricane_flag, case.hurricane_dt = 1, int(3 * case.dt)
        case.hurricane_start, case.hurricane_end = (2020, 1, 2), (2020, 1, 3)
        x0, y0 = float(case.x.min()), float(0.5 * (case.y.min() + case.y.max()))
        times = [k * case.hurricane_dt for k in range(case.ndt // 3 + 3)]
        case.hurricane = meshgen.holland_snapshots(case.x, case.y, times, (x0 - 200.0, y0 - 150.0),
                                                   (5.0, 1.0), p_min=950.0, r_max=300.0, lat=case.lat)
    return case


def apply(state: "abi.Arrays", values: Dict[str, np.ndarray]) -> None:
    for k, v in values.items():
        if k in state.arr:
            state[k] = v


# Output formats and hot start (formats_estuary.npz): run A writes a restart at its last
# step and state dumps (2-D formatted with a file roll-over after two dumps, 3-D
# unformatted every second step); run B starts hot from A's restart file and writes the
# other forms, with a restart after every step.
FORMATS = dict(kind="estuary", args=dict(ncol=8, nrep=2, nlayer=4, ndt=4, river_q=50.0,
                                          dry_land=True, tide_phase=-90.0))


def formats_cases():
    """(case A, function making case B from A's restart fields)."""
    from dataclasses import replace
    from gom_release_amd import gomio, meshgen
    a = meshgen.estuary_case(**FORMATS["args"])
    a.restart_out = -1
    a.dump2d = gomio.DumpSpec(binary=0, time=1, file_freq=2, frequency=1)
    a.dump3d = gomio.DumpSpec(binary=1, time=4, file_freq=99999999, frequency=2)

    def b_of(restart):
        return replace(a, ndt=2, restart_in=1, hot_start=restart, restart_out=1, restart_freq=1,
                       dump2d=gomio.DumpSpec(binary=1, time=3, file_freq=99999999, frequency=1),
                       dump3d=gomio.DumpSpec(binary=0, time=2, file_freq=99999999, frequency=1))
    return a, b_of


PLOTS = dict(kind="estuary", args=dict(ncol=12, nrep=2, nlayer=4, ndt=10, river_q=50.0,
                                        dry_land=True, tide_phase=-90.0))


def plots_cases():
    """The two runs of plots_estuary.npz: A with time series at cells and Tecplot + VTK
    output in 2-D and of the 3-D surface (files rolling over every two outputs); B with time
    series at nodes, 2-D Tecplot only in km and minutes, 3-D surface VTK only in seconds."""
    from dataclasses import replace
    from gom_release_amd import gomio, meshgen
    a = meshgen.estuary_case(**PLOTS["args"])
    stations = [(3, 5, "st_a"), (40, 30, "st_b"), (57, 45, "st_c")]
    a.tser = gomio.TserSpec(stations=stations, hloc=1, time=4, frequency=1, format=3)
    a.out2d = gomio.ContourSpec(format=3, file_freq=4, time=4, frequency=2, unit_conv=1.0)
    a.out3d = gomio.ContourSpec(format=3, file_freq=4, time=3, frequency=2, unit_conv=0.001, surf=1, full=0)
    b = replace(a, ndt=6,
                tser=gomio.TserSpec(stations=stations[:2], hloc=2, time=2, frequency=2, format=3),
                out2d=gomio.ContourSpec(format=1, file_freq=99999999, time=2, frequency=1, start=2, end=5,
                                        unit_conv=0.001, variables=(1, 0, 1, 1, 0, 1)),
                out3d=gomio.ContourSpec(format=2, file_freq=99999999, time=1, frequency=3, unit_conv=1.0,
                                        surf=1, full=0))
    return {"A": a, "B": b}


class PlotsGolden:
    """plots_estuary.npz: the contour and time-series files the reference wrote (bytes)."""

    def __init__(self):
        z = np.load(os.path.join(GOLDEN, "plots_estuary.npz"))
        self.data = {k: z[k] for k in z.files}

    def files(self, run: str) -> Dict[str, bytes]:
        pre = f"{run}/files/"
        return {k[len(pre):]: v.tobytes() for k, v in self.data.items() if k.startswith(pre)}


class FormatsGolden:
    """formats_estuary.npz: reference-written files (bytes) and the states behind them."""

    def __init__(self):
        z = np.load(os.path.join(GOLDEN, "formats_estuary.npz"))
        self.data = {k: z[k] for k in z.files}

    def state(self, run: str, tag: str) -> Dict[str, np.ndarray]:
        pre = f"{run}/{tag}/"
        return {k[len(pre):]: _flat(v) for k, v in self.data.items() if k.startswith(pre)}

    def files(self, run: str) -> Dict[str, bytes]:
        pre = f"{run}/files/"
        return {k[len(pre):]: v.tobytes() for k, v in self.data.items() if k.startswith(pre)}
