"""GOM's contour and time-series output (Tecplot and VTK), written from device snapshots.

Files and their reference writers (GOM_Rv1.1.0/source):
* ``tec2D_NNNN.dat``: write_tec2D_head.f90 / write_tec2D_body.f90, a head zone when a file
  opens (open_2D_out_files.f90, at start-up) and body zones after, rolled over every
  IS2D_File_freq / IS2D_frequency outputs (write_output_files.f90:26-126);
* ``vtk2D_NNNN.vtk``: write_vtk2D.f90, one file per output;
* ``tec3D_surf_NNNN.dat`` and ``vtk3D_surf_NNNN.vtk``: write_tec3D_surf_head/_body.f90 and the
  surface part of write_vtk3D.f90 (write_output_files.f90:131-297);
* ``<var>_tser.dat`` / ``.txt``: open_tser_out_files.f90 (headers and the start-up row) and
  write_tser_out_files.f90 (a row every tser_frequency steps) for eta, H, u, v, salt, temp.

The text is the reference's byte for byte: ``F0.5`` as the reference's Fortran runtime writes
it (no leading zero before the point), ``F15.5``, ``E15.5E4`` and the integer edits, values
rounded to nearest (ties to even) on the exact binary value; a list-directed empty write is
an empty line. What the reference computes on the way is kept as it stands:
* the surface VTK file's third coordinate is eta_cell(i) of node index i
  (write_vtk3D.f90:71-77), so it needs maxnod <= maxele;
* the tec3D surface body's zone title holds elapsed_time, its SOLUTIONTIME julian_day
  (write_tec3D_surf_body.f90:22-25);
* the H time series at nodes at start-up reads h_node at the station's cell index
  (open_tser_outfiles_sub.f90:68-72).
Not built: 3-D full-depth output (write_tec3D_full_*, the full part of write_vtk3D, which
interpolate to sigma levels through zto_sigma_*_values), the air-pressure series (the
reference writes it to a unit it never opens), binary Tecplot, flood maps.

Output is asynchronous (AsyncOutput): after a step the scenario takes a device snapshot of the
fields the due writers need (gomgpu_snapshot: a device-side copy on the step's stream, then a
copy to pinned host memory on a stream of its own) and a writer thread formats and writes the
files while the next steps run.
"""
from __future__ import annotations

import os
import queue
import threading
from typing import Callable, Dict, List, Mapping, Optional, Sequence

import numpy as np

from .gomio import Case, ContourSpec, TserSpec
from .gomout import fortran_e, fortran_f

TIME_CONV = {1: 86400.0, 2: 1440.0, 3: 24.0, 4: 1.0}
TIME_CHAR = {1: " sec", 2: " min", 3: "  hr", 4: " day"}


# ----------------------------------------------------------------------------- edits

def f0(v: float, d: int = 5) -> str:
    """Fortran ``F0.d`` as the reference's runtime writes it: the shortest field, with no
    zero before the point of a value below one in magnitude."""
    s = f"{v:.{d}f}"
    if s.startswith("0."):
        return s[1:]
    if s.startswith("-0."):
        return "-" + s[2:]
    return s


def f0_row(vals) -> str:
    """One record of ``(n(F0.5,1x))``: the values one blank apart (the last blank is a
    position edit with nothing after it, so the record ends at the last value)."""
    return " ".join(f0(float(v)) for v in vals)


def _blocks(vals, per: int = 100) -> List:
    vals = np.asarray(vals)
    return [vals[i:i + per] for i in range(0, vals.size, per)]


def f0_block(vals) -> str:
    """A node field in records of 100 ``F0.5`` values, then the empty record."""
    return "".join(f0_row(b) + "\n" for b in _blocks(vals)) + "\n"


def e_block(vals) -> str:
    """A node field in records of 100 ``E15.5E4`` values (the 3-D surface file)."""
    return "".join("".join(fortran_e(float(v)) for v in b) + "\n" for b in _blocks(vals))


def i_w(v: int, w: int) -> str:
    s = str(int(v))
    return s.rjust(w) if len(s) <= w else "*" * w


# ----------------------------------------------------------------------------- geometry

class OutputMesh:
    """What the writers read of the static mesh."""

    def __init__(self, mesh, params):
        self.P, self.E, self.L = int(params.maxnod), int(params.maxele), int(params.maxlayer)
        self.x = np.asarray(mesh["x_node"], dtype=np.float64)
        self.y = np.asarray(mesh["y_node"], dtype=np.float64)
        self.h_node = np.asarray(mesh["h_node"], dtype=np.float64)
        self.h_cell = np.asarray(mesh["h_cell"], dtype=np.float64)
        self.tq = np.asarray(mesh["tri_or_quad"], dtype=np.int64)
        self.cells = mesh.view2("nodenum_at_cell").astype(np.int64)      # (4, E), 1-based
        self.bot_cell = np.asarray(mesh["bottom_layer_at_element"], dtype=np.int64)
        self.bot_node = np.asarray(mesh["bottom_layer_at_node"], dtype=np.int64)
        self.xmin, self.ymin = float(self.x.min()), float(self.y.min())
        tec = self.cells.copy()           # read_cell_inp.f90:107-111: a triangle repeats node 3
        tri = self.tq == 3
        tec[3, tri] = tec[2, tri]
        self.tec_cells = tec
        self.conn_size = int(self.E + self.tq.sum())


def _time(jd: float, conv: float) -> float:
    return jd * conv


# ----------------------------------------------------------------------------- 2-D

def tec2d_zone(g: OutputMesh, st: Mapping, t: float, spec: ContourSpec, zone: int, head: bool) -> str:
    """One zone of tec2D (write_tec2D_head.f90 when head, else write_tec2D_body.f90)."""
    var = spec.variables
    out = []
    if head:
        out.append('Title = "2D contour plot"\n')
        out.append('Variables = "X [m]", "Y [m]"' if spec.unit_conv > 0.5 else 'Variables = "X [km]", "Y [km]"')
        names = (', "eta [m]"', ', "u [m/s]"', ', "v [m/s]"', ', "salt [ppt]"', ', "temp [C]"', ', "rho [kg/m3]"')
        out.append("".join(n for n, on in zip(names, var) if on == 1) + "\n")
        out.append('ZONE T = "' + fortran_f(t) + '", N =' + i_w(g.P, 10) + ", E = " + i_w(g.E, 10)
                   + ", DATAPACKING=BLOCK, ZONETYPE=FEQUADRILATERAL, SOLUTIONTIME=" + fortran_f(t)
                   + ", STRANDID=1\n")
        out.append(f0_block((g.x - g.xmin) * spec.unit_conv))
        out.append(f0_block((g.y - g.ymin) * spec.unit_conv))
    else:
        out.append('ZONE T = "' + fortran_f(t) + '", N =' + i_w(g.P, 10) + ", E = " + i_w(g.E, 10)
                   + ", DATAPACKING=BLOCK, ZONETYPE=FEQUADRILATERAL, VARSHARELIST=([1-2]=1), "
                   + "CONNECTIVITYSHAREZONE=1, SOLUTIONTIME=" + fortran_f(t) + ", \tSTRANDID=1\n")
    for on, f in zip(var, ("eta_node", "ubar_node", "vbar_node", "sbar_node", "tbar_node", "rbar_node")):
        if on == 1:
            out.append(f0_block(st[f]))
    if head:
        tc = g.tec_cells
        out.append("".join(i_w(tc[0, e], 10) + i_w(tc[1, e], 10) + i_w(tc[2, e], 10) + i_w(tc[3, e], 10) + "\n"
                           for e in range(g.E)))
    out.append('TEXT CS=FRAME, HU=FRAME, X=50, Y=95, H=2.5, AN=MIDCENTER, T="TIME = ' + fortran_f(t)
               + TIME_CHAR[spec.time] + '", ZN =' + i_w(zone, 5) + "\n\n")
    return "".join(out)


def _vtk_cells(g: OutputMesh) -> str:
    rows = []
    for e in range(g.E):
        n = int(g.tq[e])
        rows.append(i_w(n, 2) + "".join(i_w(g.cells[j, e] - 1, 10) for j in range(n)) + "\n")
    return "".join(rows)


def _vtk_types(g: OutputMesh) -> str:
    return "".join(i_w(5 if n == 3 else 9, 2) + "\n" for n in g.tq)


def vtk2d(g: OutputMesh, st: Mapping, t: float, spec: ContourSpec) -> str:
    """write_vtk2D.f90."""
    u = spec.unit_conv
    out = ["# vtk DataFile Version 3.0\n", 'Title = "2D contour plot"\n', "ASCII\n", "\n",
           "DATASET UNSTRUCTURED_GRID\n", "FIELD FieldData 1\n", "TIME 1 1 float\n", fortran_f(t) + "\n",
           "POINTS " + i_w(g.P, 10) + " float\n"]
    zero = fortran_f(0.0)
    out.append("".join(fortran_f((x - g.xmin) * u) + fortran_f((y - g.ymin) * u) + zero + "\n"
                       for x, y in zip(g.x, g.y)))
    out.append("\n")
    out.append("CELLS " + i_w(g.E, 10) + i_w(g.conn_size, 10) + "\n")
    out.append(_vtk_cells(g))
    out.append("\n")
    out.append("CELL_TYPES " + i_w(g.E, 10) + "\n")
    out.append(_vtk_types(g))
    var = spec.variables
    if max(var) > 0:
        out.append("\n")
        out.append("POINT_DATA " + i_w(g.P, 10) + "\n")

    def scal(name, f):
        out.append("\nSCALARS " + name + " float\nLOOKUP_TABLE default\n")
        out.append("".join(fortran_f(float(v)) + "\n" for v in st[f]))
    if var[0] == 1:
        scal("Eta(m)", "eta_node")
    if var[1] == 1 or var[2] == 1:
        out.append("\nVECTORS Velocity(m/s) float\n")
        out.append("".join(fortran_f(float(a)) + fortran_f(float(b)) + zero + "\n"
                           for a, b in zip(st["ubar_node"], st["vbar_node"])))
    if var[3] == 1:
        scal("Salt(psu)", "sbar_node")
    if var[4] == 1:
        scal("Temp(C)", "tbar_node")
    if var[5] == 1:
        scal("Rho(kg/m3)", "rbar_node")
    return "".join(out)


# ----------------------------------------------------------------------------- 3-D surface

def tec3d_surf_zone(g: OutputMesh, st: Mapping, jd: float, elapsed: float, spec: ContourSpec, zone: int,
                    head: bool) -> str:
    """One zone of tec3D_surf (write_tec3D_surf_head.f90 / write_tec3D_surf_body.f90)."""
    conv = TIME_CONV[spec.time]
    t = jd * conv
    out = []
    if head:
        out.append('Title = "3D contour plot"\n')
        out.append('Variables = "X [m]", "Y [m]", "Z [m]"\n' if spec.unit_conv > 0.5
                   else 'Variables = "X [km]", "Y [km]", "Z [m]"\n')
        out.append('ZONE T = "' + fortran_f(t) + '", N =' + i_w(g.P, 10) + ", E = " + i_w(g.E, 10)
                   + ", DATAPACKING=BLOCK, ZONETYPE=FEQUADRILATERAL, SOLUTIONTIME=" + fortran_f(t)
                   + ", STRANDID=1\n")
        out.append(e_block((g.x - g.xmin) * spec.unit_conv) + "\n")
        out.append(e_block((g.y - g.ymin) * spec.unit_conv) + "\n")
        out.append(e_block(st["eta_node"]) + "\n")
        tc = g.tec_cells
        out.append("".join(i_w(tc[0, e], 10) + i_w(tc[1, e], 10) + i_w(tc[2, e], 10) + i_w(tc[3, e], 10) + "\n"
                           for e in range(g.E)))
    else:
        out.append('ZONE T = "' + fortran_f(elapsed * conv) + '", N =' + i_w(g.P, 10) + ", E = "
                   + i_w(g.E, 10) + ", DATAPACKING=BLOCK, ZONETYPE=FEQUADRILATERAL, VARSHARELIST=([1-2]=1),  "
                   + "CONNECTIVITYSHAREZONE=1, SOLUTIONTIME=" + fortran_f(t) + ", STRANDID=1\n")
        out.append(e_block(st["eta_node"]))
    out.append('TEXT CS=FRAME, HU=FRAME, X=50, Y=95, H=2.5, AN=MIDCENTER, T="TIME = ' + fortran_f(t)
               + TIME_CHAR[spec.time] + '", ZN =' + i_w(zone, 5) + "\n\n")
    return "".join(out)


def vtk3d_surf(g: OutputMesh, st: Mapping, spec: ContourSpec) -> str:
    """The surface file of write_vtk3D.f90 (its third coordinate is eta_cell of the node's
    index, as written there)."""
    u = spec.unit_conv
    eta_c = st["eta_cell"]
    out = ["# vtk DataFile Version 3.0\n", 'Title = "3D contour plot"\n', "ASCII\n", "\n",
           "DATASET UNSTRUCTURED_GRID\n", "POINTS " + i_w(g.P, 10) + " float\n"]
    out.append("".join(fortran_f((g.x[i] - g.xmin) * u) + fortran_f((g.y[i] - g.ymin) * u)
                       + fortran_f(float(eta_c[i])) + "\n" for i in range(g.P)))
    out.append("\n")
    out.append("CELLS " + i_w(g.E, 10) + i_w(g.conn_size, 10) + "\n")
    out.append(_vtk_cells(g))
    out.append("\n")
    out.append("CELL_TYPES " + i_w(g.E, 10) + "\n")
    out.append(_vtk_types(g))
    return "".join(out)


# ----------------------------------------------------------------------------- time series

TSER = [  # (tser_list, file stem, flag index in TserSpec.fields, title)
    (1, "eta_tser_from_msl", 0, "Water elevation time series from msl"),
    (2, "H_tser", 1, "Total water depth (H) time series"),
    (3, "u_tser", 2, "Velocity 2d u time series"),
    (4, "v_tser", 3, "Velocity 2d v time series"),
    (5, "salt_tser", 4, "Salinity time series"),
    (6, "temp_tser", 5, "Temperature time series"),
]
TIME_NAME = {1: "sec", 2: "min", 3: "hr", 4: "day"}


def _tser_row(t: float, vals) -> str:
    return fortran_e(t) + "".join("," + fortran_e(float(v)) for v in vals) + "\n"


class TserWriter:
    """The time-series files of open_tser_out_files.f90 and write_tser_out_files.f90."""

    def __init__(self, out_dir: str, spec: TserSpec, g: OutputMesh):
        if len(spec.fields) > 6 and spec.fields[6] == 1:
            raise NotImplementedError("tser_airp: the reference writes the air-pressure series to a unit "
                                      "open_tser_out_files never opens")
        self.spec, self.g = spec, g
        self.conv = TIME_CONV[spec.time]
        self.cells = np.array([s[0] for s in spec.stations], dtype=np.int64) - 1
        self.nodes = np.array([s[1] for s in spec.stations], dtype=np.int64) - 1
        if self.cells.min() < 0 or self.cells.max() >= g.E or self.nodes.min() < 0 or self.nodes.max() >= g.P:
            raise ValueError("time-series station outside the mesh")
        self.files = {}
        for lst, stem, flag, title in TSER:
            if spec.fields[flag] != 1:
                continue
            for ext in ((".dat",) if spec.format == 1 else (".txt",) if spec.format == 2 else (".dat", ".txt")):
                self.files[(lst, ext)] = open(os.path.join(out_dir, stem + ext), "w", newline="\n")
        self.buff = 0

    def fields(self) -> List[str]:
        f = ["eta_cell", "eta_node", "top_layer_at_element", "top_layer_at_node"]
        on = self.spec.fields
        if on[2] or on[3]:
            f += ["u_node", "v_node", "ubar_node", "vbar_node"]
        if on[4]:
            f += ["salt_cell", "salt_node", "sbar_node"]
        if on[5]:
            f += ["temp_cell", "temp_node", "tbar_node"]
        return f

    def _cell_profile(self, st, lst):
        """Station rows of u, v (layer means of the cell's node values) or salt, temp (cell
        values) at cells: the depth mean, then layers 1..maxlayer (zeros where dry or below
        the bed), as write_tser_out_files.f90:60-234 forms them."""
        g, L = self.g, self.g.L
        top = np.asarray(st["top_layer_at_element"])
        row = []
        for ie in self.cells:
            prof = np.zeros(L)
            bar = 0.0
            tl, bl = int(top[ie]), int(g.bot_cell[ie])
            if tl != 0:
                if lst in (3, 4):
                    f = st["u_node" if lst == 3 else "v_node"].reshape(-1, L + 1)   # (P, 0:L)
                    n = int(g.tq[ie])
                    for k in range(bl, tl + 1):
                        s = 0.0
                        for l in range(n):
                            nd = g.cells[l, ie] - 1
                            s = s + f[nd, k - 1] + f[nd, k]
                        prof[k - 1] = s / (n * 2)
                    src = prof
                else:
                    src = st["salt_cell" if lst == 5 else "temp_cell"].reshape(-1, L)[ie]
                for k in range(bl, tl + 1):
                    bar = bar + src[k - 1]
                bar = bar / (tl - bl + 1)
                if lst in (5, 6):
                    prof = np.asarray(src, dtype=np.float64).copy()
            elif lst in (5, 6):
                prof = np.asarray(st["salt_cell" if lst == 5 else "temp_cell"].reshape(-1, L)[ie], dtype=np.float64)
            row.append(bar)
            row.extend(prof)
        return row

    def _node_start(self, st, lst):
        """The start-up row at nodes (open_tser_outfiles_sub.f90, tser_hloc 2): the routine's
        own depth means. u and v: levels bottom-1..top over (layers + 1). Salinity and
        temperature: layers bottom-1..top over the layer count, where layer 0 of node nd is the
        array element before layer 1, i.e. layer maxlayer of node nd-1 (0 for the first
        node). A dry node's mean is 0."""
        g, L = self.g, self.g.L
        top = st["top_layer_at_node"]
        prof_name = {3: "u_node", 4: "v_node", 5: "salt_node", 6: "temp_node"}[lst]
        flat = np.asarray(st[prof_name], dtype=np.float64)
        lo = 0 if lst in (3, 4) else 1
        ld = L + 1 if lst in (3, 4) else L
        row = []
        for nd in self.nodes:
            tl, bl = int(top[nd]), int(g.bot_node[nd])
            bar = 0.0
            if tl != 0:
                nv = tl - bl + 1
                for k in range(bl - 1, tl + 1):
                    idx = (k - lo) + nd * ld
                    bar = bar + (float(flat[idx]) if idx >= 0 else 0.0)
                bar = bar / (nv + 1) if lst in (3, 4) else bar / nv
            row.append(bar)
            row.extend(flat[nd * ld + (1 - lo):nd * ld + (1 - lo) + L])
        return row

    def _values(self, st, lst, start: bool):
        g, sp = self.g, self.spec
        L = g.L
        if lst == 1:
            return st["eta_cell"][self.cells] if sp.hloc == 1 else st["eta_node"][self.nodes]
        if lst == 2:
            if sp.hloc == 1:
                return g.h_cell[self.cells] + st["eta_cell"][self.cells]
            # at start-up the reference reads h_node at the station's cell index
            hidx = self.cells if start else self.nodes
            return g.h_node[hidx] + st["eta_node"][self.nodes]
        if sp.hloc == 1:
            return self._cell_profile(st, lst)
        if start:
            return self._node_start(st, lst)
        bar = {3: "ubar_node", 4: "vbar_node", 5: "sbar_node", 6: "tbar_node"}[lst]
        prof = {3: "u_node", 4: "v_node", 5: "salt_node", 6: "temp_node"}[lst]
        row = []
        for nd in self.nodes:
            row.append(st[bar][nd])
            if lst in (3, 4):
                row.extend(st[prof].reshape(-1, L + 1)[nd, 1:L + 1])
            else:
                row.extend(st[prof].reshape(-1, L)[nd, :L])
        return row

    def open(self, jd: float, st) -> None:
        """Headers and the start-up row (open_tser_outfiles_tec/_vtk/_sub.f90)."""
        sp = self.spec
        names = [s[2] for s in sp.stations]
        L = self.g.L
        for (lst, ext), fh in self.files.items():
            title = [t for (l, _s, _f, t) in TSER if l == lst][0]
            if ext == ".dat":
                fh.write('Title = "' + title + (' at cell"' if sp.hloc == 1 else ' at node"') + "\n")
                fh.write('Variables = "Elapsed Time [' + TIME_NAME[sp.time] + ']"')
            else:
                fh.write('"Elapsed Time [' + TIME_NAME[sp.time] + ']"')
            if lst <= 2:
                fh.write("".join(', "' + n.strip() + '"' for n in names) + "\n")
            else:
                fh.write("".join(', "' + n.strip() + "_(" + f"{k:02d}" + ')"' for n in names for k in range(L + 1))
                         + "\n")
            fh.write(_tser_row(jd * self.conv, self._values(st, lst, True)))

    def due(self, it: int) -> bool:
        return it % self.spec.frequency == 0

    def write(self, jd: float, st) -> None:
        rows = {}
        for (lst, ext), fh in self.files.items():
            if lst not in rows:
                rows[lst] = _tser_row(jd * self.conv, self._values(st, lst, False))
            fh.write(rows[lst])
            fh.flush()

    def close(self) -> None:
        for fh in self.files.values():
            fh.close()


# ----------------------------------------------------------------------------- schedules

class ContourWriter:
    """The 2-D or 3-D-surface contour files of one run with write_output_files.f90's
    schedule: an output every `frequency` steps in [start, end]; a new Tecplot file (head
    zone) every File_freq / frequency outputs, body zones between; one VTK file per output."""

    def __init__(self, out_dir: str, kind: str, spec: ContourSpec, g: OutputMesh):
        if kind == "3D" and spec.full == 1:
            raise NotImplementedError("3-D full-depth output (write_tec3D_full_*, zto_sigma_*_values) is not built")
        if kind == "3D" and spec.format in (2, 3) and g.P > g.E:
            raise ValueError("the surface VTK file reads eta_cell at node indices: it needs maxnod <= maxele")
        self.out_dir, self.kind, self.spec, self.g = out_dir, kind, spec, g
        self.conv = TIME_CONV[spec.time]
        self.data_in = int(spec.file_freq / spec.frequency)
        if self.data_in < 1:
            raise ValueError(f"{kind} output: File_freq {spec.file_freq} is below frequency {spec.frequency}")
        self.tec = spec.format in (1, 3) and (kind == "2D" or spec.surf == 1)
        self.vtk = spec.format in (2, 3) and (kind == "2D" or spec.surf == 1)
        self.buff = self.buff2 = 0
        self.file_num = 1
        self.zone = 1
        self.vtk_num = 0
        self.fh = None
        self.files: List[str] = []

    def fields(self) -> List[str]:
        if self.kind == "2D":
            return ["eta_node", "ubar_node", "vbar_node", "sbar_node", "tbar_node", "rbar_node"]
        return ["eta_node", "eta_cell"]

    def _tec_name(self) -> str:
        stem = "tec2D_" if self.kind == "2D" else "tec3D_surf_"
        return os.path.join(self.out_dir, f"{stem}{self.file_num:04d}.dat")

    def _zone(self, jd, elapsed, st, head: bool) -> str:
        if self.kind == "2D":
            return tec2d_zone(self.g, st, jd * self.conv, self.spec, self.zone, head)
        return tec3d_surf_zone(self.g, st, jd, elapsed, self.spec, self.zone, head)

    def open(self, jd: float, elapsed: float, st) -> None:
        """open_2D_out_files.f90 / open_3D_out_files.f90: the first Tecplot file and its head."""
        if self.tec:
            path = self._tec_name()
            self.fh = open(path, "w", newline="\n")
            self.files.append(path)
            self.fh.write(self._zone(jd, elapsed, st, True))
            self.fh.flush()

    def due(self, it: int) -> bool:
        sp = self.spec
        if not sp.start <= it <= sp.end:
            return False
        self.buff += 1
        return self.buff % sp.frequency == 0

    def write(self, jd: float, elapsed: float, st) -> None:
        """The output of a due step (call due() first)."""
        self.buff2 += 1
        if self.tec:
            if self.buff2 != 1 and self.buff2 % self.data_in == 1:
                self.fh.close()
                self.file_num += 1
                self.zone = 1
                path = self._tec_name()
                self.fh = open(path, "w", newline="\n")
                self.files.append(path)
                self.fh.write(self._zone(jd, elapsed, st, True))
            else:
                self.zone += 1
                self.fh.write(self._zone(jd, elapsed, st, False))
            self.fh.flush()
        if self.vtk:
            self.vtk_num += 1
            stem = "vtk2D_" if self.kind == "2D" else "vtk3D_surf_"
            path = os.path.join(self.out_dir, f"{stem}{self.vtk_num:04d}.vtk")
            text = vtk2d(self.g, st, jd * self.conv, self.spec) if self.kind == "2D" else vtk3d_surf(self.g, st,
                                                                                                    self.spec)
            with open(path, "w", newline="\n") as f:
                f.write(text)
            self.files.append(path)

    def close(self) -> None:
        if self.fh is not None:
            self.fh.close()
            self.fh = None


class PlotOutputs:
    """The time-series and contour writers a case asks for (cards C41-C43), fed state
    snapshots: open() with the start-up state, then step() after each step."""

    def __init__(self, case: Case, mesh, params, out_dir: str):
        self.g = OutputMesh(mesh, params)
        self.tser = TserWriter(out_dir, case.tser, self.g) if case.tser is not None and case.tser.stations else None
        self.c2 = ContourWriter(out_dir, "2D", case.out2d, self.g) if case.out2d is not None else None
        self.c3 = ContourWriter(out_dir, "3D", case.out3d, self.g) if case.out3d is not None else None

    def fields(self) -> List[str]:
        out: List[str] = []
        for w in (self.tser, self.c2, self.c3):
            if w is not None:
                out += [f for f in w.fields() if f not in out]
        return out

    def open(self, jd: float, elapsed: float, st) -> None:
        # prepare_gom.f90:186-210: time series, then 2-D, then 3-D
        if self.tser is not None:
            self.tser.open(jd, st)
        if self.c2 is not None:
            self.c2.open(jd, elapsed, st)
        if self.c3 is not None:
            self.c3.open(jd, elapsed, st)

    def due(self, it: int):
        """Which writers step `it` feeds (advances the contour schedules)."""
        return [w for w in (self.tser, self.c2, self.c3)
                if w is not None and (w.due(it) if w is not self.tser else self.tser.due(it))]

    def write(self, writers, jd: float, elapsed: float, st) -> None:
        for w in writers:
            if w is self.tser:
                w.write(jd, st)
            else:
                w.write(jd, elapsed, st)

    def close(self) -> None:
        for w in (self.tser, self.c2, self.c3):
            if w is not None:
                w.close()

    def files(self) -> List[str]:
        out = []
        for w in (self.c2, self.c3):
            if w is not None:
                out += w.files
        if self.tser is not None:
            out += [f.name for f in self.tser.files.values()]
        return out


class AsyncOutput:
    """Formats and writes PlotOutputs' files in a thread of its own, from snapshots the
    scenario hands over, so that output steps do not wait for the text to be written."""

    def __init__(self, plots: PlotOutputs, depth: int = 4):
        self.plots = plots
        self.q: "queue.Queue" = queue.Queue(maxsize=depth)
        self.err: Optional[BaseException] = None
        self.t = threading.Thread(target=self._run, daemon=True)
        self.t.start()

    def _run(self) -> None:
        while True:
            job = self.q.get()
            if job is None:
                return
            try:
                writers, jd, elapsed, fetch = job
                self.plots.write(writers, jd, elapsed, fetch())
            except BaseException as e:  # reported by the next submit or close
                self.err = e

    def submit(self, writers, jd: float, elapsed: float, fetch: Callable[[], Mapping]) -> None:
        if self.err is not None:
            raise self.err
        self.q.put((writers, jd, elapsed, fetch))

    def close(self) -> None:
        self.q.put(None)
        self.t.join()
        self.plots.close()
        if self.err is not None:
            raise self.err
