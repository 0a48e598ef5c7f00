"""Time loop of one scenario on the device (the role of main.f90:53-304).

A Scenario owns a device context, computes the host forcing of each step (spin-up,
harmonic open-boundary elevation, boundary concentrations of the transport sub-cycles,
river discharge, uniform wind) and advances the step through the C ABI. The state stays
on the device; only O(boundary) forcing crosses the bus each step.
"""
from __future__ import annotations

import math
import os
from typing import Optional

import numpy as np

from . import abi, forcing as fz, host, prep
from .gomio import Case

DYNAMIC = ["eta_at_ob_old", "eta_at_ob_new", "con_at_ob", "con_at_Qb", "Q_add"]


class HostForcing:
    """Per-step host forcing of a generated case (the 2-D routines of main.f90:84-161)."""

    def __init__(self, case: Case, params: "abi.Params", mesh: "abi.Arrays", wind=(0.0, 0.0)):
        self.case, self.params, self.mesh = case, params, mesh
        self.forcing = abi.ForcingArrays(params)
        if wind[0] or wind[1]:
            gt, wsn, wst = fz.wind_stress(mesh, *wind, rho_a=case.rho_a, rho_o=case.rho_o)
            self.forcing["Gamma_T"] = gt
            self.forcing["wind_stress_normal"] = wsn
            self.forcing["wind_stress_tangnt"] = wst
        self.julian_day = fz.julian_day0(case)
        self.tran_ids = [params.tran_id1, params.tran_id2]
        self.uploads = list(DYNAMIC)      # forcing arrays the coming step needs uploaded
        # set when the model computes the boundary forcing itself (GOMGPU_STAGE_BOUNDARY):
        # then a step needs only its scalars
        self.device_boundary = False
        # set when the model generates the Holland cyclone's snapshots itself: then a read step
        # sends the centre of its snapshot, not the node fields
        self.device_holland = False
        self._snap_index = 0
        if case.hurricane_flag == 1:
            # read_hurricane_ser_0.f90: the first snapshot pair; the interpolated field,
            # air_p1 and the face winds start at zero (allocate_variables.f90:507-519)
            from .gomio import hurricane_snapshots
            self._snaps = hurricane_snapshots(case)
            f = self.forcing
            for tag in ("0", "2"):
                u, v, pr = self._next_snapshot(0)
                f["wind_u" + tag], f["wind_v" + tag], f["air_p" + tag] = u, v, pr
            self.window = fz.hurricane_window(case)

    def _next_snapshot(self, it: int):
        try:
            snap = next(self._snaps)
        except StopIteration:
            raise RuntimeError(f"the hurricane fields end before step {it}") from None
        self._snap_index += 1
        return snap

    def advance(self, it: int, eta_node: Optional[np.ndarray] = None) -> "abi.ForcingArrays":
        """Fill the forcing of step `it`. With eta_node the river face velocity is computed
        here (compute_boundary_velocity.f90:32-46); otherwise the device computes it."""
        c = self.case
        elapsed = it * c.dt
        self.julian_day = self.julian_day + c.dt / 86400.0
        sp_tide = fz.spinup(c.tide_spinup, c.tide_spinup_period, elapsed)
        sp_bc = fz.spinup(c.baroclinic_spinup, c.baroclinic_spinup_period, elapsed) \
            if c.baroclinic_flag == 1 else 1.0
        f = self.forcing
        f.scalars["elapsed_time"] = elapsed
        f.scalars["spinup_function_tide"] = sp_tide
        f.scalars["spinup_function_baroclinic"] = sp_bc
        f.scalars["julian_day"] = self.julian_day
        self.uploads = [] if self.device_boundary else list(DYNAMIC)
        if c.hurricane_flag == 1:
            ph = fz.hurricane_phase(it, self.julian_day, self.window, self.params.hurricane_read_interval)
            f.scalars["hurricane_phase"] = float(ph)
            if ph == 2 and self.device_holland:
                tr = c.hurricane
                if self._snap_index >= len(tr):
                    raise RuntimeError(f"the hurricane fields end before step {it}")
                cx, cy = tr.centre(tr.times[self._snap_index])
                self._snap_index += 1
                f.scalars["hurricane_centre_x"], f.scalars["hurricane_centre_y"] = cx, cy
            elif ph == 2:
                f["wind_u_next"], f["wind_v_next"], f["air_p_next"] = self._next_snapshot(it)
                self.uploads += ["wind_u_next", "wind_v_next", "air_p_next"]
        if self.device_boundary:
            return f
        old, new = fz.boundary_eta(c, it, sp_tide, self.julian_day)
        f["eta_at_ob_old"] = old
        f["eta_at_ob_new"] = new
        if c.transport_solver == 2:
            nqb = max(len(c.rivers), 1)
            prev = self.forcing["con_at_Qb"].reshape(-1, 2, nqb, order="F")[0]
            ob, qb = fz.elm_boundary_concentrations(c, self.julian_day, sp_bc, prev)
        else:
            ob, qb = fz.tvd_boundary_concentrations(c, self.julian_day, sp_bc, self.params.maxtran2,
                                                    self.tran_ids)
        f["con_at_ob"] = ob
        f["con_at_Qb"] = qb
        if c.rivers:
            q = fz.river_discharge(c, self.julian_day, self.forcing["Q_add"])
            f["Q_add"] = q
            if eta_node is not None:
                m = self.mesh
                isf = m["isflowside3"]
                nf = m.view2("nodenum_at_face")
                qu = np.zeros(max(len(c.rivers), 1))
                for j in np.flatnonzero(isf > 0):
                    i = isf[j] - 1
                    depth = m["h_face"][j] + (eta_node[nf[0, j] - 1] + eta_node[nf[1, j] - 1]) * 0.5
                    qu[i] = -q[i] / (m["face_length"][j] * depth)
                f["Qu_boundary"] = qu
        return f


def _compact(arrays):
    """Arrays whose entries are all one bit pattern travel as (value, size): most of a cold
    start (Av, Kv, q2, xl, uniform tracers, zeros) is constant, and the pickled transfer from
    a preparation worker is otherwise gigabytes."""
    out = {}
    for k, v in arrays.items():
        bits = v.view(np.uint64) if v.dtype == np.float64 else v
        if v.size and (bits == bits[0]).all():
            out[k] = ("const", v.dtype.str, v.size, v[0].item())
        else:
            out[k] = v
    return out


def _expand(dst, packed):
    for k, v in packed.items():
        if isinstance(v, tuple):
            _, dt, n, val = v
            a = dst.arr[k]
            if a.dtype != np.dtype(dt) or a.size != n:
                raise ValueError(f"{k}: packed array does not match the parameters")
            if not (val == 0 and not (isinstance(val, float) and math.copysign(1.0, val) < 0)):
                a[...] = val   # the arrays start as +0
        else:
            dst.arr[k] = v


def pack_prepared(case: Case, precision: int = 0):
    """prep.build in a picklable, compact form (for preparing scenarios in worker processes)."""
    p, m, st = prep.build(case, precision)
    return abi.params_to_dict(p), _compact(m.arr), _compact(st.arr)


def unpack_prepared(packed):
    pd, marr, sarr = packed
    p = abi.params_from_dict(pd)
    m = abi.mesh_arrays(p)
    s = abi.state_arrays(p)
    _expand(m, marr)
    _expand(s, sarr)
    return p, m, s


class Scenario:
    """One scenario advanced on the device."""

    def __init__(self, case: Case, precision: int = 0, wind=(0.0, 0.0), model: Optional[host.GpuModel] = None,
                 prepared=None, model_factory=None, device_boundary: bool = True, device_holland: bool = True):
        self.case = case
        if prepared is None:
            self.params, self.mesh, state0 = prep.build(case, precision)
        else:
            self.params, self.mesh, state0 = unpack_prepared(prepared)
        if case.rivers:
            self.params.river_on_device = 1
        self.model = model or (model_factory or host.GpuModel)(self.mesh, self.params)
        # a model that does not derive the river face velocity itself gets it from the host
        self.host_river = bool(case.rivers) and getattr(self.model, "host_river_velocity", False)
        self.model.set_state(state0)
        self.model.stage("eos")          # initial density and depth means, as prepare_gom does
        self.model.stage("bars")
        self.host_forcing = HostForcing(case, self.params, self.mesh, wind)
        self.forcing = self.host_forcing.forcing
        self.model.set_forcing(self.forcing)   # static fields once
        # a model that computes the boundary forcing on the device gets its tables once; the
        # host then sends only each step's scalars
        if device_boundary and hasattr(self.model, "set_boundary") and not self.host_river:
            self.model.set_boundary(fz.boundary_tables(case))
            self.host_forcing.device_boundary = True
        # a Holland cyclone's snapshots generated on the device
        from .meshgen import HollandTrack
        if device_holland and case.hurricane_flag == 1 and isinstance(case.hurricane, HollandTrack) \
                and hasattr(self.model, "set_holland"):
            self.model.set_holland(case.hurricane)
            self.host_forcing.device_holland = True
        self.it = 0

    def stage_forcing(self) -> None:
        """Host forcing of the next step, uploaded asynchronously (before a batch step)."""
        self.it += 1
        hf = self.host_forcing
        if self.host_river:
            hf.advance(self.it, eta_node=self.fields(["eta_node"])["eta_node"])
            self.model.set_forcing(self.forcing, only=hf.uploads + ["Qu_boundary"])
        else:
            hf.advance(self.it)
            self.model.set_forcing(self.forcing, only=hf.uploads)

    def step(self) -> None:
        self.stage_forcing()
        self.model.step()
        if self.outputs is not None:
            self.write_outputs()

    def run(self, n: int) -> None:
        for _ in range(n):
            self.step()

    # ---- output files (prepare_gom.f90:186-227 opens them; main.f90:274-295 writes)
    outputs = None
    plots = None
    async_out = None

    def fields(self, names) -> "abi.Arrays":
        """The named state fields, read back from the device."""
        s = abi.state_arrays(self.params)
        self.model.get_state(s, only=list(names))
        return s

    def open_outputs(self, out_dir: str, asynchronous: bool = True) -> None:
        """Write the case's output into `out_dir` after each step: state dumps and restart
        files (gomout), and the time series and Tecplot/VTK contour files of cards C41-C43
        (gomplot), whose start-up part is written here. With `asynchronous` the contour and
        time-series files are formatted and written by a thread of their own from device
        snapshots, while the following steps run."""
        from . import gomout, gomplot
        os.makedirs(out_dir, exist_ok=True)
        self.outputs = out_dir
        self.dumps = [gomout.DumpWriter(out_dir, kind, spec, self.params.maxlayer)
                      for kind, spec in (("2D", self.case.dump2d), ("3D", self.case.dump3d))
                      if spec is not None]
        c = self.case
        self.plots = self.async_out = None
        if (c.tser is not None and c.tser.stations) or c.out2d is not None or c.out3d is not None:
            self.plots = gomplot.PlotOutputs(c, self.mesh, self.params, out_dir)
            self.plots.open(self.host_forcing.julian_day, self.it * c.dt, self.fields(self.plots.fields()))
            if asynchronous:
                self.async_out = gomplot.AsyncOutput(self.plots)

    def _snapshot(self, names):
        """A callable returning the named fields as they are now: a device snapshot when the
        model takes one (the read happens on the caller of the callable), else a read now."""
        if hasattr(self.model, "snapshot"):
            return self.model.snapshot(list(names))
        s = self.fields(names)
        return lambda: s

    def write_outputs(self) -> list:
        """After step `it`: the dumps the schedule asks for, then the restart file when
        restart_out is -1 at the last step or 1 every restart_freq steps. Returns the
        files written."""
        from . import gomout
        c, it = self.case, self.it
        out = []
        for d in self.dumps:
            if d.step(it, self.host_forcing.julian_day, self.fields):
                out.append(d.files[-1])
        if self.plots is not None:
            due = self.plots.due(it)
            if due:
                names = []
                for w in due:
                    names += [f for f in w.fields() if f not in names]
                fetch = self._snapshot(names)
                jd, elapsed = self.host_forcing.julian_day, it * c.dt
                if self.async_out is not None:
                    self.async_out.submit(due, jd, elapsed, fetch)
                else:
                    self.plots.write(due, jd, elapsed, fetch())
        if (c.restart_out == -1 and it == c.ndt) or (c.restart_out == 1 and it % c.restart_freq == 0):
            out.append(gomout.write_restart(gomout.restart_name(self.outputs, it),
                                            self.fields(gomout.RESTART_FIELDS)))
        return out

    def close_outputs(self) -> None:
        """Close the output files (waiting for the output thread to finish its work)."""
        for d in getattr(self, "dumps", []):
            d.close()
        if getattr(self, "async_out", None) is not None:
            self.async_out.close()
        elif getattr(self, "plots", None) is not None:
            self.plots.close()
        self.plots = self.async_out = None
        self.outputs = None

    def active_layers(self):
        """(face-layers, cell-layers, node-layers) of the wet entities in the current state."""
        s = abi.state_arrays(self.params)
        self.model.get_state(s, only=["top_layer_at_element", "top_layer_at_face", "top_layer_at_node"])
        out = []
        for ent in ("face", "element", "node"):
            top = s[f"top_layer_at_{ent}"].astype(np.int64)
            bot = self.mesh[f"bottom_layer_at_{ent}"].astype(np.int64)
            out.append(int(np.sum(np.where(top > 0, top - bot + 1, 0))))
        return tuple(out)

    def active_cell_layers(self) -> int:
        """Sum over wet cells of (top - bottom + 1) for the current state."""
        s = abi.state_arrays(self.params)
        self.model.get_state(s, only=["top_layer_at_element"])
        top = s["top_layer_at_element"].astype(np.int64)
        bot = self.mesh["bottom_layer_at_element"].astype(np.int64)
        return int(np.sum(np.where(top > 0, top - bot + 1, 0)))
