"""ctypes loader for liboracle.so, the C restatement of the GOM step (test oracle only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module, and
only as the checker or the CPU baseline -- never as the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
LIB_OMP = os.path.join(HERE, "liboracle_omp.so")   # the same source threaded with OpenMP
_libs = {}


def build() -> None:
    subprocess.run(["make", "-C", HERE, "liboracle.so", "liboracle_omp.so"], check=True,
                   stdout=subprocess.DEVNULL)


def lib(threaded: bool = False):
    """The serial build, or with threaded the OpenMP build (OMP_NUM_THREADS threads), which
    gives the same bits."""
    path = LIB_OMP if threaded else LIB
    if path not in _libs:
        if not os.path.exists(path):
            build()
        _lib = C.CDLL(path)
        _libs[path] = _lib
        from gom_release_amd import abi
        args = [C.POINTER(abi.Mesh), C.POINTER(abi.Params), C.POINTER(abi.State),
                C.POINTER(abi.Forcing)]
        for st in abi.STAGES + ["step"]:
            if st == "boundary":  # the boundary forcing's check is the host forcing (forcing.py)
                continue
            fn = getattr(_lib, "oracle_" + st)
            fn.argtypes = args
            fn.restype = C.c_int
        _lib.oracle_rho_unesco.argtypes = [C.c_double, C.c_double]
        _lib.oracle_rho_unesco.restype = C.c_double
        _lib.oracle_last_pcg_iterations.restype = C.c_int
        _lib.oracle_last_error_entity.restype = C.c_int
        _lib.oracle_tridiag.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 5
        _lib.oracle_pcg.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_int, C.c_double]
        _lib.oracle_check_intersection.argtypes = [C.c_double] * 8 + [C.c_void_p, C.c_void_p]
        _lib.oracle_bilinear.argtypes = [C.c_double] * 10 + [C.c_void_p] * 3
        _lib.oracle_elm_record.argtypes = [C.c_void_p]
        _lib.oracle_elm_record.restype = None
        _lib.oracle_elm_near_edge.argtypes = [C.c_int]
        _lib.oracle_elm_near_edge.restype = C.c_longlong
        _lib.oracle_elm_near_reset.restype = None
    return _libs[path]


class Oracle:
    """Runs oracle stages on numpy-backed ABI arrays (mesh, params, state, forcing). With
    threaded the stages run in the OpenMP build; the static accessors (iteration counts,
    error entities, near-edge counts) read the serial build's."""

    def __init__(self, mesh, params, state, forcing, threaded: bool = False):
        self.mesh, self.params, self.state, self.forcing = mesh, params, state, forcing
        self._m = mesh.struct()
        self._s = state.struct()
        self._lib = lib(threaded)

    def run(self, stage: str) -> int:
        f = self.forcing.struct()
        rc = getattr(self._lib, "oracle_" + stage)(C.byref(self._m), C.byref(self.params),
                                                   C.byref(self._s), C.byref(f))
        return rc

    def step(self) -> int:
        return self.run("step")

    def elm_departures(self):
        """Run the ELM stage recording its departures: returns (rc, nnel, jlev), each array
        [maxlayer, maxface], nnel negated for a point within 1e-12 of a cell edge."""
        L, J = self.params.maxlayer, self.params.maxface
        dep = np.zeros((2, L, J), dtype=np.int32)
        self._lib.oracle_elm_record(dep.ctypes.data)
        try:
            rc = self.run("elm")
        finally:
            self._lib.oracle_elm_record(None)
        return rc, dep[0], dep[1]

    @staticmethod
    def elm_near_edge(reset: bool = False):
        """Points within 1e-12 of a cell edge: (momentum ELM, tracer ELM)."""
        out = (int(lib().oracle_elm_near_edge(0)), int(lib().oracle_elm_near_edge(1)))
        if reset:
            lib().oracle_elm_near_reset()
        return out

    @staticmethod
    def last_error_entity() -> int:
        """The first failing entity of the last stage that stopped."""
        return lib().oracle_last_error_entity()

    @staticmethod
    def pcg_iterations() -> int:
        return lib().oracle_last_pcg_iterations()
