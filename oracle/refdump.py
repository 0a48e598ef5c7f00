"""Reader for the reference harness dumps and runner for the reference harness.

Test infrastructure only (part of the oracle): it reads the full-precision records that
``oracle/ref_harness.f90`` writes from the reference's module state, and runs that
harness (built into ``oracle/_ref`` from the sources under /root/reference) on a project
directory written by ``gom_release_amd.gomio``.
"""
from __future__ import annotations

import os
import subprocess
from typing import Dict, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(HERE, "_ref", "gom_ref_harness")


def read_dump(path: str) -> Dict[str, np.ndarray]:
    """Parse one dump file into {name: array}; 2-D arrays keep Fortran (layer, entity) order
    as a (extent1, extent2) array with C index [k - lower1, entity - 1]."""
    out: Dict[str, np.ndarray] = {}
    raw = open(path, "rb").read()
    pos = 0
    hdr_dt = np.dtype([("name", "S32"), ("kind", "<i4"), ("rank", "<i4"),
                       ("lo", "<i4", 3), ("ex", "<i4", 3)])
    while pos < len(raw):
        h = np.frombuffer(raw, dtype=hdr_dt, count=1, offset=pos)[0]
        pos += hdr_dt.itemsize
        name = h["name"].decode().strip()
        dt = np.float64 if h["kind"] == 1 else np.int32
        ex = [int(v) for v in h["ex"]]
        rank = int(h["rank"])
        n = ex[0] * ex[1] * ex[2]
        a = np.frombuffer(raw, dtype=dt, count=n, offset=pos).copy()
        pos += n * np.dtype(dt).itemsize
        if rank == 0:
            out[name] = a[0]
        elif rank == 1:
            out[name] = a
        else:
            # Fortran order (ex0 fastest) -> array indexed [entity, k]; transpose to [k, entity]
            out[name] = a.reshape(ex[1], ex[0]).T.copy()
            out[name + "@lo"] = np.int32(h["lo"][0])
    return out


def harness_available() -> bool:
    return os.path.isfile(HARNESS) and os.access(HARNESS, os.X_OK)


def run_reference(project: str, dumpdir: Optional[str], first: int = 1, last: int = 0,
                  extra: Sequence[int] = (), threads: int = 1, timeout: float = 3600.0,
                  perturb: int = 255, no_heat: bool = False) -> str:
    """Run the reference harness in `project`. With dumpdir None it only times the run.
    With `no_heat` the harness leaves out the surface heat model (phi_sz stays 0).

    MALLOC_PERTURB_=255 makes every heap allocation start zero-filled, so the automatic
    arrays the reference reads before writing (the TVD limiter block) read zeros.
    """
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(threads)
    env["MALLOC_PERTURB_"] = str(perturb)
    env["GLIBC_TUNABLES"] = "glibc.malloc.tcache_count=0"
    env.setdefault("OMP_STACKSIZE", "512M")
    if no_heat:
        env["GOM_HARNESS_NO_HEAT"] = "1"
    if dumpdir is not None:
        os.makedirs(dumpdir, exist_ok=True)
        args = [HARNESS, os.path.abspath(dumpdir), str(first), str(last)] + [str(e) for e in extra]
    else:
        args = [HARNESS, "-"]
    p = subprocess.run(args, cwd=project, env=env, capture_output=True, text=True,
                       timeout=timeout)
    if p.returncode != 0:
        raise RuntimeError(f"reference harness failed ({p.returncode}):\n{p.stdout[-3000:]}\n{p.stderr[-3000:]}")
    # A Fortran STOP exits with status 0: detect it from the log of the run.
    log = os.path.join(project, "output", "run.log")
    tail = open(log).read()[-400:] if os.path.exists(log) else ""
    want = [last] if last >= first else []
    want += [e for e in extra]
    if dumpdir is not None and any(not os.path.exists(os.path.join(dumpdir, f"s{w}_post.bin"))
                                   for w in want if w > 0):
        raise RuntimeError("reference stopped early: " + tail)
    return p.stdout


# ----------------------------------------------------------------------------------
# Conversion of dumps into the ABI structs (Fortran layout, flattened).

def _flat(a: np.ndarray) -> np.ndarray:
    """[k, entity] dump array -> Fortran-order flat buffer (layer fastest)."""
    return np.asfortranarray(a).ravel(order="F") if a.ndim == 2 else np.ravel(a)


def params_from_geom(g: Dict[str, np.ndarray]):
    from gom_release_amd import abi
    tid = list(g["tran_id"]) + [0, 0]
    d = dict(
        maxnod=g["maxnod"], maxele=g["maxele"], maxface=g["maxface"], maxlayer=g["maxlayer"],
        max_no_neighbor_node=g["max_no_neighbor_node"], num_ob_cell=g["num_ob_cell"],
        num_Qb_cell=g["num_Qb_cell"], maxtran2=g["maxtran2"], tran_id1=tid[0], tran_id2=tid[1],
        advection_flag=g["advection_flag"], ELM_backtrace_flag=g["ELM_backtrace_flag"],
        ELM_min_iter=g["ELM_min_iter"], ELM_max_iter=g["ELM_max_iter"], bf_flag=g["bf_flag"],
        baroclinic_flag=g["baroclinic_flag"], ana_density=g["ana_density"],
        i_density_flag=g["i_density_flag"], transport_flag=g["transport_flag"],
        trans_sub_iter=g["trans_sub_iter"], h_flux_limiter=g["h_flux_limiter"],
        v_flux_limiter=g["v_flux_limiter"], heat_option=g["heat_option"],
        max_iteration_pcg=g["max_iteration_pcg"], precision=0,
        dt=g["dt"], theta=g["theta"], gravity=g["gravity"], rho_o=g["rho_o"], MSL=g["MSL"],
        dry_depth=g["dry_depth"], adv_onoff_depth=g["adv_onoff_depth"],
        smagorinsky_parameter=g["smagorinsky_parameter"], Ah_0=g["Ah_0"], Kh_0=g["Kh_0"],
        von_Karman=g["von_Karman"], error_tolerance=g["error_tolerance"],
        denominator_min_for_matrix=g["denominator_min_for_matrix"], delta_z_min=g["delta_z_min"])
    # dumped by the harness since the wind stage exists (older fixtures lack them)
    for k in ("rho_a", "wind_formular", "hurricane_flag", "hurricane_read_interval", "transport_solver"):
        if k in g:
            d[k] = g[k]
    if d["transport_flag"] == 1 and "transport_solver" not in d:
        d["transport_solver"] = 1   # fixtures dumped before the card was read: the TVD solver
    if d.get("transport_solver") == 2 and d["transport_flag"] == 1:
        # the ABI's convention for the ELM solver, which carries both tracers
        d.update(maxtran2=2, tran_id1=1, tran_id2=2)
    return abi.params_from_dict(d)


def mesh_from_geom(g: Dict[str, np.ndarray], p):
    from gom_release_amd import abi
    m = abi.mesh_arrays(p)
    for name, _ in abi.MESH_REAL + abi.MESH_INT:
        if name == "Qb_cell_id":
            if p.num_Qb_cell > 0:
                m[name] = g["Q_boundary"][:, 0][: p.num_Qb_cell] if g["Q_boundary"].ndim == 2 else g["Q_boundary"]
            continue
        if name == "ob_cell_id" and p.num_ob_cell == 0:
            continue
        if name not in g:
            continue
        m[name] = _flat(g[name])
    return m


def _qb_first_column(qb: np.ndarray, n: int) -> np.ndarray:
    return qb[:, 0][:n]


def state_from_dump(d: Dict[str, np.ndarray], p):
    from gom_release_amd import abi
    s = abi.state_arrays(p)
    for name, _ in abi.STATE_REAL + abi.STATE_INT:
        if name in d:
            s[name] = _flat(d[name])
    return s


def forcing_from_dump(d: Dict[str, np.ndarray], p, con_at_ob=None, con_at_Qb=None):
    from gom_release_amd import abi
    f = abi.ForcingArrays(p)
    for name, _ in abi.FORCING_REAL:
        if name in d:
            f[name] = _flat(d[name])
    f.scalars["elapsed_time"] = float(d["elapsed_time"])
    f.scalars["spinup_function_tide"] = float(d["spinup_function_tide"])
    f.scalars["spinup_function_baroclinic"] = float(d["spinup_function_baroclinic"])
    if "spinup_function_wind" in d:
        f.scalars["spinup_function_wind"] = float(d["spinup_function_wind"])
    if con_at_ob is not None:
        f["con_at_ob"] = con_at_ob
    if con_at_Qb is not None:
        f["con_at_Qb"] = con_at_Qb
    return f
