"""ctypes mirror of the C ABI in ``include/gomgpu.h``.

The field lists are read from the header's X-macro lists at import time, so the Python
structs cannot drift from the C ones. Arrays are numpy buffers in the reference's Fortran
layout (column-major ``(layer, entity)``), flattened; :class:`Arrays` owns them and fills
the struct pointers.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from typing import Dict, List, Tuple

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gomgpu.h")


def _parse_lists(path: str) -> Dict[str, List[Tuple[str, ...]]]:
    text = open(path).read()
    out: Dict[str, List[Tuple[str, ...]]] = {}
    for mac in re.finditer(r"#define (GOMGPU_[A-Z_]+)\(X\)((?:.*\\\n)*.*)", text):
        name, body = mac.group(1), mac.group(2).replace("\\\n", " ")
        items = re.findall(r"X\(([^)]*)\)", body)
        out[name] = [tuple(s.strip() for s in it.split(",")) for it in items]
    return out


LISTS = _parse_lists(HEADER)
MESH_REAL = LISTS["GOMGPU_MESH_REAL"]
MESH_INT = LISTS["GOMGPU_MESH_INT"]
STATE_REAL = LISTS["GOMGPU_STATE_REAL"]
STATE_INT = LISTS["GOMGPU_STATE_INT"]
FORCING_REAL = LISTS["GOMGPU_FORCING_REAL"]
PARAMS_INT = [t[0] for t in LISTS["GOMGPU_PARAMS_INT"]]
PARAMS_REAL = [t[0] for t in LISTS["GOMGPU_PARAMS_REAL"]]
FORCING_SCALARS = [t[0] for t in LISTS["GOMGPU_FORCING_SCALARS"]]

PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int32)


class Mesh(C.Structure):
    _fields_ = [(n, PD) for n, _ in MESH_REAL] + [(n, PI) for n, _ in MESH_INT]


class Params(C.Structure):
    _fields_ = [(n, C.c_int32) for n in PARAMS_INT] + [(n, C.c_double) for n in PARAMS_REAL]


class State(C.Structure):
    _fields_ = [(n, PD) for n, _ in STATE_REAL] + [(n, PI) for n, _ in STATE_INT]


class Forcing(C.Structure):
    _fields_ = [(n, C.c_double) for n in FORCING_SCALARS] + [(n, PD) for n, _ in FORCING_REAL]


class ForcingOut(C.Structure):
    _fields_ = [(n, PD) for n, _ in FORCING_REAL]


STAGES = ["friction", "elm", "momentum", "freesurf", "velocity", "layers", "nodevel",
          "transport", "eos", "update", "turbulence", "wind", "boundary", "bars"]


class Holland(C.Structure):
    """gomgpu_holland: a Holland cyclone generated on the device."""
    _fields_ = [(n, C.c_double) for n in ("p_min", "p_env", "r_max", "b", "coriolis", "rho_air",
                                           "surface_factor", "vx", "vy")]


class Boundary(C.Structure):
    """gomgpu_boundary: the tables of the device boundary forcing (GOMGPU_STAGE_BOUNDARY)."""
    _fields_ = [("num_harmonic_sets", C.c_int32), ("harmonic_first", PI), ("tidal_phase_shift", PD),
                ("constituents", PD), ("num_series", C.c_int32), ("series_first", PI), ("series_time", PD),
                ("series_value", PD), ("num_ob", C.c_int32), ("ob_eta_type", PI), ("ob_eta_id", PI),
                ("ob_tracer_series", PI), ("num_qb", C.c_int32), ("qb_q_series", PI), ("qb_portion", PD),
                ("qb_tracer_series", PI), ("is_salt", C.c_int32), ("is_temp", C.c_int32),
                ("ref_salt", C.c_double), ("ref_temp", C.c_double), ("reference_diff_days", C.c_int32)]


def extent(code: str, p: Params) -> int:
    """Number of elements of an array whose shape code is `code` (see the header)."""
    P, E, J, L = p.maxnod, p.maxele, p.maxface, p.maxlayer
    NB = p.max_no_neighbor_node
    nt = max(p.trans_sub_iter, 1)
    tab = {
        "P": P, "E": E, "J": J, "L": L, "L1": L + 1, "4E": 4 * E, "2J": 2 * J,
        "NBP": NB * P, "NB1P": (NB + 1) * P, "OB": max(p.num_ob_cell, 1),
        "QB": max(p.num_Qb_cell, 1), "LJ": L * J, "L1J": (L + 1) * J, "LE": L * E,
        "L1E": (L + 1) * E, "LP": L * P, "L1P": (L + 1) * P,
        "NTOB": nt * max(p.maxtran2, 1) * max(p.num_ob_cell, 1),
        "NTQB": nt * max(p.maxtran2, 1) * max(p.num_Qb_cell, 1),
        "HP": P if p.hurricane_flag == 1 else 1, "HJ": J if p.hurricane_flag == 1 else 1,
    }
    return tab[code]


def leading(code: str, p: Params) -> int:
    """Leading (layer) dimension of a 2-D array code, 1 for 1-D arrays."""
    if code in ("LJ", "LE", "LP"):
        return p.maxlayer
    if code in ("L1J", "L1E", "L1P"):
        return p.maxlayer + 1
    if code == "4E":
        return 4
    if code == "2J":
        return 2
    if code == "NBP":
        return p.max_no_neighbor_node
    if code == "NB1P":
        return p.max_no_neighbor_node + 1
    return 1


class Arrays:
    """numpy storage for one of the structs; attribute access by field name."""

    def __init__(self, fields_real, fields_int, p: Params, struct_cls):
        self.p = p
        self.codes: Dict[str, str] = {}
        self.arr: Dict[str, np.ndarray] = {}
        for n, code in fields_real:
            self.arr[n] = np.zeros(extent(code, p), dtype=np.float64)
            self.codes[n] = code
        for n, code in fields_int:
            self.arr[n] = np.zeros(extent(code, p), dtype=np.int32)
            self.codes[n] = code
        self.struct_cls = struct_cls

    def __getitem__(self, name: str) -> np.ndarray:
        return self.arr[name]

    def __setitem__(self, name: str, value) -> None:
        a = self.arr[name]
        v = np.asarray(value, dtype=a.dtype).ravel(order="F")
        if v.shape != a.shape:
            raise ValueError(f"{name}: expected {a.shape[0]} values, got {v.shape[0]}")
        a[...] = v

    def view2(self, name: str) -> np.ndarray:
        """(layer, entity) view of a 2-D field (Fortran order)."""
        a = self.arr[name]
        ld = leading(self.codes[name], self.p)
        return a.reshape((ld, a.shape[0] // ld), order="F")

    def struct(self, only=None):
        s = self.struct_cls()
        for name, a in self.arr.items():
            if only is not None and name not in only:
                continue
            ptr = a.ctypes.data_as(PD if a.dtype == np.float64 else PI)
            setattr(s, name, ptr)
        return s

    def copy(self) -> "Arrays":
        other = Arrays.__new__(Arrays)
        other.p = self.p
        other.codes = dict(self.codes)
        other.arr = {k: v.copy() for k, v in self.arr.items()}
        other.struct_cls = self.struct_cls
        return other


def mesh_arrays(p: Params) -> Arrays:
    return Arrays(MESH_REAL, MESH_INT, p, Mesh)


def state_arrays(p: Params) -> Arrays:
    return Arrays(STATE_REAL, STATE_INT, p, State)


class ForcingArrays(Arrays):
    def __init__(self, p: Params):
        super().__init__(FORCING_REAL, [], p, Forcing)
        self.scalars = {n: 0.0 for n in FORCING_SCALARS}
        self.scalars["spinup_function_tide"] = 1.0
        self.scalars["spinup_function_baroclinic"] = 1.0
        self.scalars["spinup_function_wind"] = 1.0

    def struct(self, only=None):
        s = super().struct(only)
        for n, v in self.scalars.items():
            setattr(s, n, float(v))
        return s

    def copy(self) -> "ForcingArrays":
        other = ForcingArrays(self.p)
        other.arr = {k: v.copy() for k, v in self.arr.items()}
        other.scalars = dict(self.scalars)
        return other


def params_from_dict(d: Dict[str, float]) -> Params:
    p = Params()
    for n in PARAMS_INT:
        setattr(p, n, int(d.get(n, 0)))
    for n in PARAMS_REAL:
        setattr(p, n, float(d.get(n, 0.0)))
    return p


def params_to_dict(p: Params) -> Dict[str, float]:
    return {n: getattr(p, n) for n in PARAMS_INT + PARAMS_REAL}
