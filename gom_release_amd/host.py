"""Python host of the HIP step: ctypes binding of libgomgpu.so (the C ABI of
include/gomgpu.h).

This plays the role of the Fortran time loop (GOM_Rv1.1.0/source/main.f90:53-304) for
tests, the benchmark and the serving loop: it owns host copies of the module arrays in the
reference's layout and calls the same per-stage entry points a Fortran host would bind with
ISO_C_BINDING. There is no CPU fallback: if the library is missing or no GPU is visible,
construction fails with an error.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Optional, Sequence

import numpy as np

from . import abi

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libgomgpu.so")
_lib = None

STAGE_IDS = {name: i for i, name in enumerate(abi.STAGES)}
# tunable launch shapes (GOMGPU_LAUNCH_* of include/gomgpu.h) and their largest block
LAUNCH = {"elm": 0, "mom_rhs": 1, "mom_solve": 2, "tvd_rhs": 3, "tvd_solve": 4, "tvd_lim": 5,
          "node_vel": 6, "turb_prep": 7, "turb_solve": 8, "stream": 9}
LAUNCH_MAX = {k: (256 if k in ("turb_solve", "mom_rhs", "elm") else 1024) for k in LAUNCH}


class GomGpuError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gomgpu error {code}: {msg}")
        self.code = code


def lib():
    """Load libgomgpu.so (built in-tree by gom_release_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m gom_release_amd.build` first")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.gomgpu_create.argtypes = [C.POINTER(abi.Mesh), C.POINTER(abi.Params), C.POINTER(vp)]
        L.gomgpu_set_state.argtypes = [vp, C.POINTER(abi.State)]
        L.gomgpu_get_state.argtypes = [vp, C.POINTER(abi.State)]
        L.gomgpu_set_forcing.argtypes = [vp, C.POINTER(abi.Forcing)]
        L.gomgpu_set_boundary.argtypes = [vp, C.POINTER(abi.Boundary)]
        L.gomgpu_snapshot.argtypes = [vp, C.POINTER(abi.State)]
        L.gomgpu_set_holland.argtypes = [vp, C.POINTER(abi.Holland)]
        L.gomgpu_snapshot_wait.argtypes = [vp]
        L.gomgpu_get_forcing.argtypes = [vp, C.POINTER(abi.ForcingOut)]
        L.gomgpu_stage.argtypes = [vp, C.c_int]
        L.gomgpu_step.argtypes = [vp]
        L.gomgpu_run.argtypes = [vp, C.c_int]
        L.gomgpu_sync.argtypes = [vp]
        L.gomgpu_budgets.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.gomgpu_pcg_iterations.argtypes = [vp]
        L.gomgpu_elm_record.argtypes = [vp, C.c_int]
        L.gomgpu_elm_departures.argtypes = [vp, vp, vp]
        L.gomgpu_elm_near_edge.argtypes = [vp, C.POINTER(C.c_longlong), C.c_int]
        L.gomgpu_last_error.argtypes = [vp]
        L.gomgpu_last_error.restype = C.c_char_p
        L.gomgpu_timings.argtypes = [vp, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_double), C.c_int]
        L.gomgpu_destroy.argtypes = [vp]
        L.gomgpu_destroy.restype = None
        L.gomgpu_set_launch.argtypes = [vp, C.c_int, C.c_int]
        L.gomgpu_get_launch.argtypes = [vp, C.c_int]
        L.gomgpu_set_pcg_rows.argtypes = [vp, C.c_int]
        L.gomgpu_footprint.argtypes = [C.POINTER(abi.Params)]
        L.gomgpu_footprint.restype = C.c_longlong
        L.gomgpu_device_bytes.argtypes = [vp]
        L.gomgpu_device_bytes.restype = C.c_longlong
        L.gomgpu_batch_create.argtypes = [C.POINTER(vp)]
        for fn in ("gomgpu_batch_add", "gomgpu_batch_remove"):
            getattr(L, fn).argtypes = [vp, vp]
        L.gomgpu_batch_layout.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        for fn in ("gomgpu_batch_size", "gomgpu_batch_step", "gomgpu_batch_sync"):
            getattr(L, fn).argtypes = [vp]
        L.gomgpu_batch_last_error.argtypes = [vp]
        L.gomgpu_batch_last_error.restype = C.c_char_p
        L.gomgpu_batch_destroy.argtypes = [vp]
        L.gomgpu_batch_destroy.restype = None
        L.gomgpu_pcg_solve.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp, vp, C.c_int, C.c_double,
                                       C.POINTER(C.c_int), vp]
        L.gomgpu_pcg_solve_noalloc.argtypes = L.gomgpu_pcg_solve.argtypes
        L.gomgpu_pcg_workspace_bytes.argtypes = [C.c_int]
        L.gomgpu_pcg_workspace_bytes.restype = C.c_longlong
        L.gomgpu_pcg_last_error.argtypes = []
        L.gomgpu_pcg_last_error.restype = C.c_char_p
        # device memory for the entry points that take device pointers, through the HIP
        # runtime the library itself links
        L.hipMalloc.argtypes = [C.POINTER(vp), C.c_size_t]
        L.hipFree.argtypes = [vp]
        L.hipMemcpy.argtypes = [vp, vp, C.c_size_t, C.c_int]
        L.hipDeviceSynchronize.argtypes = []
        _lib = L
    return _lib


EXPORTED = ["gomgpu_create", "gomgpu_set_state", "gomgpu_set_forcing", "gomgpu_set_boundary", "gomgpu_stage",
            "gomgpu_step", "gomgpu_run", "gomgpu_get_state", "gomgpu_get_forcing", "gomgpu_budgets",
            "gomgpu_pcg_iterations", "gomgpu_sync", "gomgpu_last_error", "gomgpu_timings",
            "gomgpu_destroy", "gomgpu_set_launch", "gomgpu_get_launch", "gomgpu_set_pcg_rows",
            "gomgpu_footprint",
            "gomgpu_device_bytes", "gomgpu_batch_create", "gomgpu_batch_add", "gomgpu_batch_remove",
            "gomgpu_batch_size", "gomgpu_batch_layout", "gomgpu_batch_step", "gomgpu_batch_sync",
            "gomgpu_batch_last_error",
            "gomgpu_batch_destroy", "gomgpu_pcg_solve", "gomgpu_pcg_workspace_bytes",
            "gomgpu_pcg_last_error", "gomgpu_pcg_solve_noalloc", "gomgpu_elm_record",
            "gomgpu_elm_departures", "gomgpu_elm_near_edge", "gomgpu_snapshot", "gomgpu_snapshot_wait", "gomgpu_set_holland"]

HIP_H2D, HIP_D2H = 1, 2
_BND_INT = {"harmonic_first", "series_first", "ob_eta_type", "ob_eta_id", "ob_tracer_series", "qb_q_series",
            "qb_tracer_series"}


def footprint(params: "abi.Params") -> int:
    """Device bytes a context for these parameters allocates."""
    n = lib().gomgpu_footprint(C.byref(params))
    if n < 0:
        raise GomGpuError(1, "gomgpu_footprint: invalid parameters or no device")
    return int(n)


def device_count() -> int:
    """Number of visible HIP devices (through the HIP runtime the library links)."""
    n = C.c_int(0)
    rc = lib().hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


class DeviceArray:
    """A numpy array copied into device memory (hipMalloc of the library's HIP runtime);
    `ptr` is the device address, `download()` copies it back."""

    def __init__(self, a: np.ndarray):
        self.host = np.ascontiguousarray(a)
        self._L = lib()
        self.ptr = C.c_void_p()
        nbytes = max(self.host.nbytes, 1)
        if self._L.hipMalloc(C.byref(self.ptr), nbytes) != 0:
            raise GomGpuError(3, f"hipMalloc of {nbytes} bytes failed")
        if self.host.nbytes and self._L.hipMemcpy(self.ptr, self.host.ctypes.data_as(C.c_void_p),
                                                  self.host.nbytes, HIP_H2D) != 0:
            raise GomGpuError(2, "hipMemcpy to the device failed")

    def download(self) -> np.ndarray:
        out = np.empty_like(self.host)
        if out.nbytes and self._L.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes, HIP_D2H) != 0:
            raise GomGpuError(2, "hipMemcpy from the device failed")
        return out

    def free(self) -> None:
        if self.ptr:
            self._L.hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def pcg_solve(vals: np.ndarray, cols: np.ndarray, rowstr: np.ndarray, b: np.ndarray, x0: np.ndarray,
              max_iteration: int, tolerance: float, noalloc: bool = False):
    """pre_conj_grad (pre_conj_grad.f90:9) through gomgpu_pcg_solve (or, with `noalloc`,
    gomgpu_pcg_solve_noalloc): the 1-based CSR system is copied to the device, solved there,
    and (return code, iterations, x) returned; the code is 0, or 5
    (GOMGPU_W_PCG_NOT_CONVERGED); other failures raise."""
    n = int(b.size)
    bufs = [DeviceArray(np.asarray(vals, dtype=np.float64)), DeviceArray(np.asarray(cols, dtype=np.int32)),
            DeviceArray(np.asarray(rowstr, dtype=np.int32)), DeviceArray(np.asarray(b, dtype=np.float64)),
            DeviceArray(np.asarray(x0, dtype=np.float64))]
    try:
        it = C.c_int(0)
        fn = lib().gomgpu_pcg_solve_noalloc if noalloc else lib().gomgpu_pcg_solve
        rc = fn(n, int(np.asarray(vals).size), *[d.ptr for d in bufs], int(max_iteration), float(tolerance),
                C.byref(it), None)
        if rc not in (0, 5):
            raise GomGpuError(rc, lib().gomgpu_pcg_last_error().decode())
        return rc, it.value, bufs[4].download()
    finally:
        for d in bufs:
            d.free()


class GpuModel:
    """Device-resident model state driven through the C ABI."""

    def __init__(self, mesh: "abi.Arrays", params: "abi.Params"):
        self.params = params
        self.mesh = mesh
        self._L = lib()
        self._ctx = C.c_void_p()
        ms = mesh.struct()
        rc = self._L.gomgpu_create(C.byref(ms), C.byref(params), C.byref(self._ctx))
        if rc != 0:
            raise GomGpuError(rc, "gomgpu_create failed (see stderr)")

    # -------------------------------------------------------------- data movement
    def set_state(self, state: "abi.Arrays", only: Optional[Sequence[str]] = None) -> None:
        st = state.struct(only)
        self._check(self._L.gomgpu_set_state(self._ctx, C.byref(st)))

    def get_state(self, state: "abi.Arrays", only: Optional[Sequence[str]] = None) -> None:
        st = state.struct(only)
        self._check(self._L.gomgpu_get_state(self._ctx, C.byref(st)))

    def set_forcing(self, forcing: "abi.ForcingArrays", only: Optional[Sequence[str]] = None) -> None:
        f = forcing.struct(only)
        self._check(self._L.gomgpu_set_forcing(self._ctx, C.byref(f)))

    def set_boundary(self, tables: dict) -> None:
        """Upload the boundary tables (forcing.boundary_tables): from then on every step
        computes its open-boundary elevations, river discharge and boundary concentrations on
        the device, and the host uploads only the per-step scalars."""
        keep = {k: np.ascontiguousarray(v, dtype=np.int32 if k in _BND_INT else np.float64)
                for k, v in tables.items() if isinstance(v, np.ndarray)}
        b = abi.Boundary()
        for name, typ in abi.Boundary._fields_:
            v = keep.get(name, tables.get(name))
            if isinstance(v, np.ndarray):
                setattr(b, name, v.ctypes.data_as(typ))
            elif v is not None:
                setattr(b, name, v)
        self._bnd_keep = keep
        self._check(self._L.gomgpu_set_boundary(self._ctx, C.byref(b)))

    def set_holland(self, track) -> None:
        """Generate the hurricane snapshots of a Holland cyclone (meshgen.HollandTrack) on the
        device: the host then sends each read step's centre instead of the node fields."""
        h = abi.Holland(p_min=track.p_min, p_env=track.p_env, r_max=track.r_max, b=track.b,
                        coriolis=track.coriolis(), rho_air=track.rho_air, surface_factor=track.surface_factor,
                        vx=track.velocity[0], vy=track.velocity[1])
        self._check(self._L.gomgpu_set_holland(self._ctx, C.byref(h)))

    def snapshot(self, names: Sequence[str]):
        """Start an asynchronous snapshot of the named state fields (gomgpu_snapshot); returns
        a callable that completes it (gomgpu_snapshot_wait, from any thread) and returns the
        fields."""
        s = abi.state_arrays(self.params)
        st = s.struct(list(names))
        self._check(self._L.gomgpu_snapshot(self._ctx, C.byref(st)))

        def wait():
            self._check(self._L.gomgpu_snapshot_wait(self._ctx))
            return s
        return wait

    def get_forcing(self, forcing: "abi.ForcingArrays", only: Optional[Sequence[str]] = None) -> None:
        """Download the device's forcing arrays (the wind stage's outputs among them)."""
        out = abi.ForcingOut()
        for name, a in forcing.arr.items():
            if only is None or name in only:
                setattr(out, name, a.ctypes.data_as(abi.PD))
        self._check(self._L.gomgpu_get_forcing(self._ctx, C.byref(out)))

    # -------------------------------------------------------------- execution
    def stage(self, name: str) -> None:
        self._check(self._L.gomgpu_stage(self._ctx, STAGE_IDS[name]))

    def step(self) -> None:
        self._check(self._L.gomgpu_step(self._ctx))

    def run(self, nsteps: int) -> None:
        self._check(self._L.gomgpu_run(self._ctx, int(nsteps)))

    def sync(self, allow_pcg_warning: bool = False) -> int:
        rc = self._L.gomgpu_sync(self._ctx)
        if rc == 5 and allow_pcg_warning:
            return rc
        self._check(rc)
        return rc

    def pcg_iterations(self) -> int:
        return int(self._L.gomgpu_pcg_iterations(self._ctx))

    def budgets(self):
        v, s = C.c_double(), C.c_double()
        self._check(self._L.gomgpu_budgets(self._ctx, C.byref(v), C.byref(s)))
        return v.value, s.value

    def elm_record(self, on: bool = True):
        """Keep (on) or stop keeping the momentum ELM's departure cell and level per face-layer."""
        self._check(self._L.gomgpu_elm_record(self._ctx, 1 if on else 0))

    def elm_departures(self):
        """(nnel, jlev) of the last recorded ELM stage, each [maxlayer, maxface] (layer k-1,
        face j-1); nnel is 1-based and negated for a point within 1e-12 of a cell edge, 0
        where nothing was backtraced."""
        L, J = self.params.maxlayer, self.params.maxface
        nnel = np.zeros((L, J), dtype=np.int32)
        jlev = np.zeros((L, J), dtype=np.int32)
        self._check(self._L.gomgpu_elm_departures(self._ctx, nnel.ctypes.data, jlev.ctypes.data))
        return nnel, jlev

    def elm_near_edge(self, reset: bool = False):
        """Departure points within 1e-12 of a cell edge: (momentum ELM, tracer ELM)."""
        c = (C.c_longlong * 2)()
        self._check(self._L.gomgpu_elm_near_edge(self._ctx, c, 1 if reset else 0))
        return int(c[0]), int(c[1])

    def timings(self, enable: bool = True) -> Dict[str, float]:
        names = (C.c_char_p * 32)()
        ms = (C.c_double * 32)()
        n = self._L.gomgpu_timings(self._ctx, 1 if enable else 0, names, ms, 32)
        return {names[i].decode(): ms[i] for i in range(n)}

    def last_error(self) -> str:
        return self._L.gomgpu_last_error(self._ctx).decode()

    def device_bytes(self) -> int:
        return int(self._L.gomgpu_device_bytes(self._ctx))

    def set_launch(self, which: str, block: int) -> None:
        self._check(self._L.gomgpu_set_launch(self._ctx, LAUNCH[which], int(block)))

    def set_pcg_rows(self, rows: int) -> None:
        """Rows per thread of the PCG partition (0 automatic, or 1, 2, 4; 4 for batching)."""
        self._check(self._L.gomgpu_set_pcg_rows(self._ctx, int(rows)))

    def get_launch(self, which: str) -> int:
        return int(self._L.gomgpu_get_launch(self._ctx, LAUNCH[which]))

    def launch_config(self) -> Dict[str, int]:
        return {k: self.get_launch(k) for k in LAUNCH}

    def close(self) -> None:
        if self._ctx:
            self._L.gomgpu_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise GomGpuError(rc, self.last_error())


class GpuBatch:
    """Several device-resident models advanced together (gomgpu_batch_*)."""

    def __init__(self):
        self._L = lib()
        self._b = C.c_void_p()
        rc = self._L.gomgpu_batch_create(C.byref(self._b))
        if rc != 0:
            raise GomGpuError(rc, "gomgpu_batch_create failed")
        self.members = []

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise GomGpuError(rc, self._L.gomgpu_batch_last_error(self._b).decode())

    def add(self, model: "GpuModel") -> None:
        self._check(self._L.gomgpu_batch_add(self._b, model._ctx))
        self.members.append(model)

    def remove(self, model: "GpuModel") -> None:
        self._check(self._L.gomgpu_batch_remove(self._b, model._ctx))
        self.members.remove(model)

    def __len__(self) -> int:
        return self._L.gomgpu_batch_size(self._b)

    def step(self) -> None:
        self._check(self._L.gomgpu_batch_step(self._b))

    def layout(self):
        """(fused PCG launches, members solved alone) of the last step."""
        g, s = C.c_int(), C.c_int()
        self._check(self._L.gomgpu_batch_layout(self._b, C.byref(g), C.byref(s)))
        return g.value, s.value

    def sync(self) -> None:
        self._check(self._L.gomgpu_batch_sync(self._b))

    def close(self) -> None:
        if self._b:
            self._L.gomgpu_batch_destroy(self._b)
            self._b = C.c_void_p()
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
