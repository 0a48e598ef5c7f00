"""GOM output formats written from the model state: restart files and state dumps.

* Restart (``write_restart.f90:25-72``; read back by ``setup_hot_start.f90:14-62``): a
  Fortran unformatted sequential file. Every record is framed by its byte count as a
  4-byte little-endian integer before and after the data. There is one record per entity
  in nine groups, each record holding the group's fields for that entity, layer-fastest:

  ==========  =======  ==================================================
  group       entity   record
  ==========  =======  ==================================================
  1           cell     eta_cell, eta_cell_new
  2           node     eta_node
  3           face     un_face(1:L), vn_face(1:L)
  4           node     u_node(0:L), v_node(0:L)
  5           cell     wn_cell(0:L)
  6           node     w_node(0:L)
  7           cell     salt_cell(1:L), temp_cell(1:L), rho_cell(1:L)
  8           node     salt_node(1:L), temp_node(1:L), rho_node(1:L)
  9           face     salt_face(1:L), temp_face(1:L), rho_face(1:L)
  ==========  =======  ==================================================

  The file is ``restart_<it, 7 digits>.out`` in the output folder.

* State dumps (``write_dump2D.f90``, ``write_dump3D.f90``): a time record
  ``julian_day * conv`` (conv 86400, 1440, 24 or 1 for seconds, minutes, hours, days;
  ``open_dump2D.f90:13-21``) then one record per node: 2-D ``eta ubar vbar sbar tbar rbar``;
  3-D ``eta u(0:L) v(0:L) w(0:L) salt(1:L) temp(1:L) rho(1:L)``. Formatted files write
  the time as ``F15.5`` and the values as ``E15.5E4``; unformatted files use the record
  framing above. Files ``dump2D_<nnnn>.dat`` / ``dump3D_<nnnn>.dat`` roll over after
  ``File_freq / frequency`` dumps, on the schedule of ``write_output_files.f90:299-362``.

Arrays are taken and returned in the C ABI's layout: 1-D per entity, or layer-fastest
``(levels, entity)`` flattened in Fortran order, the layout of the reference's modules.
"""
from __future__ import annotations

import math
import os
from typing import Callable, Dict, List, Mapping, Optional, Sequence

import numpy as np

from .gomio import DumpSpec

# (entity code, fields, levels per field): "E" cells, "P" nodes, "J" faces;
# levels 0 = one value, "L" = layers 1..L, "L1" = levels 0..L
RESTART_GROUPS = [
    ("E", ("eta_cell", "eta_cell_new"), 0),
    ("P", ("eta_node",), 0),
    ("J", ("un_face", "vn_face"), "L"),
    ("P", ("u_node", "v_node"), "L1"),
    ("E", ("wn_cell",), "L1"),
    ("P", ("w_node",), "L1"),
    ("E", ("salt_cell", "temp_cell", "rho_cell"), "L"),
    ("P", ("salt_node", "temp_node", "rho_node"), "L"),
    ("J", ("salt_face", "temp_face", "rho_face"), "L"),
]
RESTART_FIELDS = [f for _, fs, _ in RESTART_GROUPS for f in fs]

DUMP_TIME_CONV = {1: 86400.0, 2: 1440.0, 3: 24.0, 4: 1.0}
DUMP2D_FIELDS = ("eta_node", "ubar_node", "vbar_node", "sbar_node", "tbar_node", "rbar_node")
DUMP3D_FIELDS = ("eta_node", "u_node", "v_node", "w_node", "salt_node", "temp_node", "rho_node")


def _nlev(code, L: int) -> int:
    return 1 if code == 0 else (L if code == "L" else L + 1)


def _matrix(arrs: Sequence[np.ndarray], nlev: int, n: int) -> np.ndarray:
    """(n, sum of levels) matrix: for each entity the fields' levels back to back."""
    cols = []
    for a in arrs:
        a = np.asarray(a, dtype=np.float64)
        if a.size != nlev * n:
            raise ValueError(f"expected {nlev} x {n} values, got {a.size}")
        cols.append(a.reshape((nlev, n), order="F").T if nlev > 1 else a.reshape(n, 1))
    return np.concatenate(cols, axis=1) if len(cols) > 1 else np.ascontiguousarray(cols[0])


def records_bytes(mat: np.ndarray) -> bytes:
    """One unformatted sequential record per row of `mat` (float64 values)."""
    n, m = mat.shape
    rec = np.empty(n, dtype=[("a", "<i4"), ("v", "<f8", (m,)), ("b", "<i4")])
    rec["a"] = m * 8
    rec["b"] = m * 8
    rec["v"] = mat
    return rec.tobytes()


def read_records(path: str) -> List[bytes]:
    """The records of an unformatted sequential file (markers checked)."""
    raw = open(path, "rb").read()
    out = []
    pos = 0
    while pos < len(raw):
        if pos + 4 > len(raw):
            raise ValueError(f"{path}: truncated record marker at byte {pos}")
        n = int(np.frombuffer(raw, "<i4", 1, pos)[0])
        if n < 0 or pos + 8 + n > len(raw):
            raise ValueError(f"{path}: bad record length {n} at byte {pos}")
        tail = int(np.frombuffer(raw, "<i4", 1, pos + 4 + n)[0])
        if tail != n:
            raise ValueError(f"{path}: record at byte {pos} ends with marker {tail}, not {n}")
        out.append(raw[pos + 4:pos + 4 + n])
        pos += 8 + n
    return out


def _sizes(fields: Mapping[str, np.ndarray]):
    E = np.asarray(fields["eta_cell"]).size
    P = np.asarray(fields["eta_node"]).size
    L = np.asarray(fields["u_node"]).size // P - 1
    J = np.asarray(fields["un_face"]).size // L
    return dict(E=E, P=P, J=J), L


def restart_bytes(fields: Mapping[str, np.ndarray]) -> bytes:
    """The bytes write_restart.f90:25-72 writes for this state."""
    n, L = _sizes(fields)
    parts = []
    for ent, names, lev in RESTART_GROUPS:
        nl = _nlev(lev, L)
        parts.append(records_bytes(_matrix([fields[f] for f in names], nl, n[ent])))
    return b"".join(parts)


def write_restart(path: str, fields: Mapping[str, np.ndarray]) -> str:
    with open(path, "wb") as f:
        f.write(restart_bytes(fields))
    return path


def restart_name(out_dir: str, it: int) -> str:
    """write_restart.f90:18-20: restart_<it as I7.7>.out."""
    return os.path.join(out_dir, f"restart_{it:07d}.out")


def read_restart(path: str, maxele: int, maxnod: int, maxlayer: int) -> Dict[str, np.ndarray]:
    """setup_hot_start.f90:14-62: the fields of a restart file, in the ABI layout.

    The face count is the one the file's size implies; every record marker is checked."""
    raw = open(path, "rb").read()
    E, P, L = maxele, maxnod, maxlayer
    n = dict(E=E, P=P, J=0)
    fixed = 0
    per_face = 0
    for ent, names, lev in RESTART_GROUPS:
        rl = 8 + 8 * len(names) * _nlev(lev, L)
        if ent == "J":
            per_face += rl
        else:
            fixed += rl * n[ent]
    if per_face == 0 or (len(raw) - fixed) % per_face or len(raw) < fixed:
        raise ValueError(f"{path}: {len(raw)} bytes do not fit a restart of "
                         f"{E} cells, {P} nodes and {L} layers")
    n["J"] = (len(raw) - fixed) // per_face
    out: Dict[str, np.ndarray] = {}
    pos = 0
    for ent, names, lev in RESTART_GROUPS:
        nl = _nlev(lev, L)
        m = nl * len(names)
        dt = np.dtype([("a", "<i4"), ("v", "<f8", (m,)), ("b", "<i4")])
        cnt = n[ent]
        rec = np.frombuffer(raw, dt, cnt, pos)
        pos += cnt * dt.itemsize
        if cnt and (np.any(rec["a"] != 8 * m) or np.any(rec["b"] != 8 * m)):
            raise ValueError(f"{path}: record markers of the {'/'.join(names)} group are not {8 * m}")
        v = rec["v"].reshape(cnt, m)
        for i, f in enumerate(names):
            blk = v[:, i * nl:(i + 1) * nl]
            out[f] = np.ascontiguousarray(blk).ravel() if nl > 1 else blk[:, 0].copy()
    return out


# ----------------------------------------------------------------------------- dumps

def fortran_e(v: float, d: int = 5, e: int = 4, w: int = 15) -> str:
    """Fortran ``Ew.dEe`` output of one value: ``0.ddddd`` mantissa with a sign only when
    negative, exponent sign and `e` digits, right-justified in `w` columns; rounded to
    nearest (ties to even) on the exact binary value."""
    if math.isnan(v):
        return "NaN".rjust(w)
    if math.isinf(v):
        return ("-Infinity" if v < 0 else "Infinity").rjust(w)
    neg = math.copysign(1.0, v) < 0
    a = abs(v)
    if a == 0.0:
        digits, ex = "0" * d, 0
    else:
        s = f"{a:.{d - 1}e}"
        mant, xs = s.split("e")
        digits = mant.replace(".", "")
        ex = int(xs) + 1
    body = ("-" if neg else "") + "0." + digits + "E" + ("-" if ex < 0 else "+") + f"{abs(ex):0{e}d}"
    return body.rjust(w) if len(body) <= w else "*" * w


def fortran_f(v: float, d: int = 5, w: int = 15) -> str:
    """Fortran ``Fw.d`` output of one value."""
    s = f"{v:.{d}f}"
    return s.rjust(w) if len(s) <= w else "*" * w


def dump_rows(kind: str, fields: Mapping[str, np.ndarray], L: int) -> np.ndarray:
    """(P, values) matrix of one dump's node records."""
    if kind == "2D":
        return _matrix([fields[f] for f in DUMP2D_FIELDS], 1, np.asarray(fields["eta_node"]).size)
    P = np.asarray(fields["eta_node"]).size
    parts = [_matrix([fields["eta_node"]], 1, P),
             _matrix([fields[f] for f in ("u_node", "v_node", "w_node")], L + 1, P),
             _matrix([fields[f] for f in ("salt_node", "temp_node", "rho_node")], L, P)]
    return np.concatenate(parts, axis=1)


def dump_bytes(kind: str, time_value: float, fields: Mapping[str, np.ndarray], L: int,
               binary: int) -> bytes:
    """One dump (write_dump2D.f90 / write_dump3D.f90) as the bytes it appends to the file."""
    rows = dump_rows(kind, fields, L)
    if binary == 1:
        return records_bytes(np.array([[time_value]])) + records_bytes(rows)
    out = [fortran_f(time_value) + "\n"]
    for r in rows:
        out.append("".join(fortran_e(float(x)) for x in r) + "\n")
    return "".join(out).encode()


class DumpWriter:
    """The state dumps of one kind ("2D" or "3D") over a run: file naming and roll-over as
    open_dump2D.f90 / open_dump3D.f90 and write_output_files.f90:299-362 do them."""

    def __init__(self, out_dir: str, kind: str, spec: DumpSpec, maxlayer: int):
        if kind not in ("2D", "3D"):
            raise ValueError("kind is '2D' or '3D'")
        self.out_dir, self.kind, self.spec, self.L = out_dir, kind, spec, maxlayer
        self.conv = DUMP_TIME_CONV[spec.time]
        self.data_in = int(spec.file_freq / spec.frequency)
        if self.data_in < 1:
            raise ValueError(f"dump {kind}: File_freq {spec.file_freq} is below frequency "
                             f"{spec.frequency} (the reference divides by zero)")
        self.buff = 0
        self.buff2 = 0
        self.file_num = 1
        self.files: List[str] = []
        self._open()

    @property
    def fields(self):
        return DUMP2D_FIELDS if self.kind == "2D" else DUMP3D_FIELDS

    def _open(self) -> None:
        path = os.path.join(self.out_dir, f"dump{self.kind}_{self.file_num:04d}.dat")
        self.f = open(path, "wb")
        self.files.append(path)

    def due(self, it: int) -> bool:
        """Advance the schedule by step `it`; True when this step writes a dump."""
        sp = self.spec
        if not sp.start <= it <= sp.end:
            return False
        self.buff += 1
        if self.buff % sp.frequency != 0:
            return False
        self.buff2 += 1
        if self.buff2 != 1 and self.buff2 % self.data_in == 1:
            self.f.close()
            self.file_num += 1
            self._open()
        return True

    def step(self, it: int, julian_day: float, get: Callable[[Sequence[str]], Mapping[str, np.ndarray]]) -> bool:
        """After step `it`: write a dump if the schedule says so; `get(names)` supplies
        the fields (called only then). Returns whether a dump was written."""
        if not self.due(it):
            return False
        self.f.write(dump_bytes(self.kind, julian_day * self.conv, get(self.fields), self.L,
                                self.spec.binary))
        self.f.flush()
        return True

    def close(self) -> None:
        self.f.close()


def read_dump(path: str, kind: str, maxnod: int, maxlayer: int, binary: int):
    """The dumps of one file: list of (time_value, (P, values) matrix)."""
    nv = 6 if kind == "2D" else 6 * maxlayer + 4
    out = []
    if binary == 1:
        recs = read_records(path)
        per = 1 + maxnod
        if len(recs) % per:
            raise ValueError(f"{path}: {len(recs)} records are not whole dumps of {maxnod} nodes")
        for k in range(0, len(recs), per):
            t = float(np.frombuffer(recs[k], "<f8")[0])
            rows = np.array([np.frombuffer(r, "<f8") for r in recs[k + 1:k + per]])
            if rows.shape != (maxnod, nv):
                raise ValueError(f"{path}: node records do not hold {nv} values")
            out.append((t, rows))
        return out
    lines = open(path).read().splitlines()
    per = 1 + maxnod
    for k in range(0, len(lines), per):
        t = float(lines[k])
        rows = np.array([[float(ln[i:i + 15]) for i in range(0, 15 * nv, 15)]
                         for ln in lines[k + 1:k + per]])
        out.append((t, rows))
    return out
