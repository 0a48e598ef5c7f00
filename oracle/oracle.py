"""ctypes loader for liboracle.so, the C restatement of the GOM step (test oracle only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module, and
only as the checker or the CPU baseline -- never as the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None


def build() -> None:
    subprocess.run(["make", "-C", HERE, "liboracle.so"], check=True,
                   stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        from gom_release_amd import abi
        args = [C.POINTER(abi.Mesh), C.POINTER(abi.Params), C.POINTER(abi.State),
                C.POINTER(abi.Forcing)]
        for st in abi.STAGES + ["step"]:
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
    return _lib


class Oracle:
    """Runs oracle stages on numpy-backed ABI arrays (mesh, params, state, forcing)."""

    def __init__(self, mesh, params, state, forcing):
        self.mesh, self.params, self.state, self.forcing = mesh, params, state, forcing
        self._m = mesh.struct()
        self._s = state.struct()

    def run(self, stage: str) -> int:
        f = self.forcing.struct()
        rc = getattr(lib(), "oracle_" + stage)(C.byref(self._m), C.byref(self.params),
                                               C.byref(self._s), C.byref(f))
        return rc

    def step(self) -> int:
        return self.run("step")

    @staticmethod
    def pcg_iterations() -> int:
        return lib().oracle_last_pcg_iterations()
