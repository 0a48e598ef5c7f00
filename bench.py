"""Benchmark of the GOM semi-implicit step on one MI355X.

Workload (default, --config storm): the baroclinic storm-surge configuration -- 1,200,990
mixed orthogonal cells (equilateral 100 m triangles stitched to 100x100 m squares), 30
z-layers, a seeded fractal bathymetry of 1-40 m, a Holland cyclone (950 hPa, r_max 40 km,
B 1.5) crossing the domain at 5 m/s with its fields every 10 steps, M2+K1 tide at the open
boundary, wet/dry, stratified temperature and salinity (10 C, 3 psu) transported by TVD
(minmod, 4 sub-cycles), GLS k-kl turbulence closure, dt = 60 s. --config estuary selects the
200,200-cell x 20-layer synthetic estuary (500 m^3/s river, uniform 10 m/s wind, 2
sub-cycles, dt = 120 s). Inputs are generated (no network) and resident on the device when
the timed region starts; per step only the O(boundary) host forcing (and a hurricane
snapshot every 10 steps) crosses the bus.

A "step" is one full semi-implicit time step (wind, friction, ELM, momentum, free surface +
PCG, velocities, layers, node velocities, closure, transport, EOS, update). The metric is
active cell-layer-steps per second: sum over wet cells of their layer count, per step.

Prints one JSON line. Extra fields: "roofline" for the dominant kernel (HBM-bound; bytes
are algorithmic, durations from HIP events on the context stream) and "cpu_baseline" (the
CPU restatement of the same step on a bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


TURBULENCE = {"off": dict(turbulence_flag=0),
              "gls": dict(turbulence_flag=3, turbulence_model="kl", stability_function="GA")}


# The PCG stop test of pre_conj_grad.f90:199 (sqrt(delta^2/delta_0^2) < tol) at the value of
# the reference's own project input (GOM_Rv1.1.0/Projects/MB_test/2_baroclinic_test/input/
# main.inp:232, card C6: max_iteration_pcg 1000, error_tolerance 5.e-6).
PCG_TOLERANCE = 5.0e-6


def storm_case(pcg_tolerance: float = PCG_TOLERANCE, turbulence: str = "gls", ncol: int = 980, nrep: int = 245):
    """The baroclinic storm-surge configuration: 1,200,990 mixed tri/quad cells (100 m), 30
    z-layers, seeded fractal bathymetry 1-40 m, a Holland cyclone (950 hPa, r_max 40 km,
    B 1.5) crossing at 5 m/s, M2+K1 tide, wet/dry, stratified T/S (10 C, 3 psu) transported
    by TVD (minmod, 4 sub-cycles), GLS k-kl closure, dt = 60 s, 2,880 steps."""
    from gom_release_amd import meshgen
    return meshgen.storm_surge_case(ncol=ncol, nrep=nrep, nlayer=30, dt=60.0, ndt=2880, h_flux_limiter=1,
                                    v_flux_limiter=1, error_tolerance=pcg_tolerance, max_iteration_pcg=1000,
                                    **TURBULENCE[turbulence])


def make_case(cells: str, turbulence: str = "gls", pcg_tolerance: float = PCG_TOLERANCE,
              storm_steps: int = 0, transport: str = "tvd"):
    """The estuary of `cells`; with storm_steps > 0 the storm-surge physics of the 1.2M-cell
    configuration: a Holland cyclone (950 hPa, r_max 40 km, B 1.5) crossing at 5 m/s, its
    fields every 10 steps for storm_steps steps, and stratified temperature (10 C) and
    salinity (3 psu) transported. transport "elm" selects the ELM transport solver
    (transport_solver 2) in place of the TVD one, with a uniform 20 C temperature and a
    5 psu river: that scheme adds the horizontal diffusion of temperature to salinity (and of
    salinity to temperature), which with a cold river drives salinity well below zero at the
    mouth, and has no clamp, so that a fresh river's salinity dips just below zero even with
    uniform temperature, where the S^1.5 term of the UNESCO equation is undefined (the
    linear equation of state is no way out: the reference stops at its first call, on the
    face densities it never computes)."""
    from gom_release_amd import meshgen
    kw = dict(nlayer=20, dt=120.0, river_q=500.0, trans_sub_iter=2, h_flux_limiter=1, v_flux_limiter=1,
              error_tolerance=pcg_tolerance, max_iteration_pcg=1000, **TURBULENCE[turbulence])
    if transport == "elm":
        kw["transport_solver"] = 2
    if cells == "200k":
        case = meshgen.estuary_case(ncol=400, nrep=100, **kw)
    elif cells == "20k":
        case = meshgen.estuary_case(ncol=100, nrep=40, **kw)
    elif cells == "1200k":  # the size of the 1.2M-cell x 30-layer configuration
        kw.update(nlayer=30, dt=60.0)
        case = meshgen.estuary_case(ncol=980, nrep=245, **kw)
    else:
        raise ValueError(cells)
    if transport == "elm":
        meshgen.elm_tracers(case)
    if storm_steps > 0:
        meshgen.add_cyclone(case, storm_steps, seed=3, p_min=950.0, r_max=40.0e3, b=1.5, speed=5.0)
        meshgen.stratify(case, d_temp=10.0, d_salt=3.0)
    return case


# HBM bytes of one launch of k_pcg_cg1m (the solve of the storm-surge configuration's
# 1,200,990 rows), from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of
# `bench.py --counter-pass 3 --copy-calibration 1024` (three solves of 58, 62 and 68
# iterations), each counter scaled by its reading of the 1 GiB copy (FETCH_SIZE x2.0,
# WRITE_SIZE x1.0): a part per launch plus a part per iteration, least squares over the three
# launches (residuals below 1e-4).
PCG_TRAFFIC_CELLS = 1200990
PCG_TRAFFIC_LAUNCH = 2.33982e8
PCG_TRAFFIC_PER_ITERATION = 2.273454e8


def pcg_bytes_per_iteration(mesh, params) -> float:
    """Algorithmic bytes of one PCG iteration: E*(12*zbar + 108), zbar = nnz per row."""
    import numpy as np
    adj = mesh.view2("adj_cellnum_at_cell")
    nnz = params.maxele + int(np.count_nonzero(adj))
    zbar = nnz / params.maxele
    return params.maxele * (12.0 * zbar + 108.0)


def step_bytes(F: int, C: int, N: int, J: int, E: int, zbar: float, nt: int, ntr: int, iters: float,
               s: int = 8, closure: bool = False) -> float:
    """Algorithmic (compulsory) HBM bytes of one step: every stage's irreducible 3-D inputs
    read once and outputs written once, dz/dzhalf derived, the new->old swap free; F, C, N
    active face-, cell- and node-layers, T tracers, Nt sub-cycles, I PCG iterations; the
    two-equation closure adds q2, xl read and written and Av written per face-layer and Kv
    per cell-layer."""
    b3 = s * ((20 + 3 * (nt - 1)) * F + (15 + (3 + 2 * ntr) * (nt - 1)) * C + 12 * N)
    if closure:
        b3 += s * (5 * F + C)
    return b3 + 24.0 * J + E * (12.0 * zbar + 20.0) + iters * E * (12.0 * zbar + 108.0)


def cpu_baseline(case, wind, budget_s: float = 15.0, max_steps: int = 12, label: str = ""):
    """Time the CPU restatement on the same workload: the oracle's OpenMP build
    (liboracle_omp.so, the same bits as the serial build), with OMP_NUM_THREADS threads or,
    unset, one per CPU the process may use."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = int(os.environ.get("OMP_NUM_THREADS") or nproc)
    os.environ["OMP_NUM_THREADS"] = str(threads)     # read when the OpenMP runtime loads
    from oracle import Oracle
    from gom_release_amd import driver, prep
    p, m, s = prep.build(case)
    hf = driver.HostForcing(case, p, m, wind)
    o = Oracle(m, p, s, hf.forcing, threaded=True)
    if o.run("eos") != 0:
        raise RuntimeError("oracle eos failed")
    t_total = 0.0
    work = 0
    n = 0
    import numpy as np
    bot = m["bottom_layer_at_element"].astype(np.int64)
    while n < max_steps and (t_total < budget_s or n == 0):
        hf.advance(n + 1, eta_node=s["eta_node"])
        top = s["top_layer_at_element"].astype(np.int64)
        t0 = time.perf_counter()
        rc = o.step()
        t_total += time.perf_counter() - t0
        if rc != 0:
            raise RuntimeError(f"oracle step failed: {rc}")
        work += int(np.sum(np.where(top > 0, top - bot + 1, 0)))
        n += 1
    return {"value": work / t_total, "unit": "cell-layer-steps/s", "cores": threads, "nproc": nproc,
            "kind": "port",
            "sample": f"first {n} steps of the same workload ({label}), FP64 C restatement built with "
                      f"OpenMP, {threads} threads, {t_total:.1f} s"}


def copy_calibration(nbytes: int) -> None:
    """One device-to-device hipMemcpy of nbytes (reads nbytes, writes nbytes) through the
    HIP runtime the library links, so profiler counters can be scaled to known traffic."""
    import ctypes as C
    from gom_release_amd import host
    hip = host.lib()
    src, dst = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(src, 1, C.c_size_t(nbytes)) or hip.hipDeviceSynchronize():
            raise RuntimeError("hipMemset failed")
        if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) or hip.hipDeviceSynchronize():  # 3: device to device
            raise RuntimeError("hipMemcpy failed")
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)


def _log_uniform(rng, lo, hi, n):
    return [float(x) for x in np.exp(rng.uniform(np.log(lo), np.log(hi), n))]


def serving_sample(n_req: int = 12, rate: float = 2.0, steps: int = 30, seed: int = 5):
    """Bounded sample of the serving workload: Poisson arrivals at `rate`, scenarios of
    10k-200k cells and 10-30 layers, `steps` steps each, admission by device footprint
    against a 64 GiB budget, continuous batching."""
    from gom_release_amd import serving
    rng = np.random.default_rng(seed)
    sizes = _log_uniform(rng, 10e3, 200e3, n_req)
    layers = rng.integers(10, 31, n_req)
    arrivals = serving.poisson_trace(n_req, rate, seed=seed)
    from gom_release_amd import meshgen
    reqs = [serving.Request(i, meshgen.add_cyclone(
                serving.estuary_of_size(sizes[i], int(layers[i]), seed=100 + i, dt=60.0, river_q=200.0,
                                        tide_phase=float(rng.uniform(-180, 180)), turbulence_flag=3),
                steps, seed=300 + i, p_min=float(rng.uniform(940.0, 990.0))),
                            steps, arrivals[i]) for i in range(n_req)]
    results, summary = serving.ServingLoop(64 << 30, workers=4).run(reqs)
    summary.update({"lambda_per_s": rate, "steps_per_request": steps,
                    "cells_min_max": [min(r.cells for r in results), max(r.cells for r in results)],
                    "sample": f"{n_req} storm-surge requests of 10k-200k cells x 10-30 layers, {steps} steps "
                              f"each (the 24 h scenarios shortened), own Holland cyclone (940-990 hPa) and "
                              f"tidal phase, GLS closure, 64 GiB budget"})
    return summary


def ensemble_sample(n_mem: int = 8, steps: int = 10, seed: int = 7):
    """Members of 20k-200k cells and 10-30 layers with their own tidal phase and their own
    Holland cyclone (track and 940-990 hPa central pressure): stepped together in one batch
    (fused PCG, captured graph) against one after another."""
    from gom_release_amd import driver, host, meshgen, serving
    rng = np.random.default_rng(seed)
    sizes = _log_uniform(rng, 20e3, 200e3, n_mem)
    layers = rng.integers(10, 31, n_mem)
    cases = [meshgen.add_cyclone(
                 serving.estuary_of_size(sizes[i], int(layers[i]), seed=200 + i, dt=60.0, river_q=200.0,
                                         tide_phase=float(rng.uniform(-180, 180)), turbulence_flag=3),
                 2 * steps + 4, seed=400 + i, p_min=float(rng.uniform(940.0, 990.0)))
             for i in range(n_mem)]
    scs = [driver.Scenario(c) for c in cases]
    for sc in scs:   # the batch partition, alone and batched alike
        sc.model.set_pcg_rows(serving.BATCH_PCG_ROWS)
    # serial: each member alone
    for sc in scs:
        sc.run(2)
        sc.model.sync()
    t0 = time.perf_counter()
    work = 0
    for sc in scs:
        a = sc.active_cell_layers()
        sc.run(steps)
        sc.model.sync()
        work += a * steps
    t_serial = time.perf_counter() - t0
    b = host.GpuBatch()
    for sc in scs:
        b.add(sc.model)
    for _ in range(2):
        for sc in scs:
            sc.stage_forcing()
        b.step()
    b.sync()
    t0 = time.perf_counter()
    workb = 0
    for sc in scs:
        workb += sc.active_cell_layers() * steps
    for _ in range(steps):
        for sc in scs:
            sc.stage_forcing()
        b.step()
    b.sync()
    t_batch = time.perf_counter() - t0
    b.close()
    for sc in scs:
        sc.model.close()
    return {"members": n_mem, "steps": steps, "cells": [int(c.maxele) for c in cases],
            "layers": [int(x) for x in layers],
            "serial_cell_layer_steps_per_s": work / t_serial, "batched_cell_layer_steps_per_s": workb / t_batch,
            "speedup": (workb / t_batch) / (work / t_serial)}


MIXED_FIELDS = ["eta_cell", "salt_cell", "temp_cell", "un_face", "vn_face", "wn_cell", "top_layer_at_element"]


def field_errors(a, b) -> dict:
    """Error of state a (mixed precision) against b (FP64): eta and velocities as relative L2,
    tracers as the largest absolute difference; wet/dry masks exact or not."""
    rel = lambda x, y: float(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300))  # noqa: E731
    return {"eta_rel_l2": rel(a["eta_cell"], b["eta_cell"]),
            "salt_max_abs_psu": float(np.max(np.abs(a["salt_cell"] - b["salt_cell"]))),
            "temp_max_abs_C": float(np.max(np.abs(a["temp_cell"] - b["temp_cell"]))),
            "un_rel_l2": rel(a["un_face"], b["un_face"]), "vn_rel_l2": rel(a["vn_face"], b["vn_face"]),
            "wn_rel_l2": rel(a["wn_cell"], b["wn_cell"]),
            "wet_dry_masks_equal": bool(np.array_equal(a["top_layer_at_element"], b["top_layer_at_element"]))}


def mixed_sample(case, wind, warmup: int, steps: int, ref_fields, fp64_ms: float, launch) -> dict:
    """The same workload with GOMGPU_MIXED (FP32 storage of the transport work arrays and of
    the PCG inner solve, FP64 refinement), the same warm-up and timed steps; its throughput
    and its state against the FP64 run's after the same steps."""
    from gom_release_amd.driver import Scenario
    sc = Scenario(case, precision=1, wind=wind)
    if launch:
        for k, v in launch.items():
            sc.model.set_launch(k, v)
    sc.run(warmup)
    sc.model.sync()
    a0 = sc.active_cell_layers()
    t0 = time.perf_counter()
    sc.run(steps)
    sc.model.sync()
    el = time.perf_counter() - t0
    a1 = sc.active_cell_layers()
    errs = field_errors(sc.fields(MIXED_FIELDS), ref_fields)
    errs["salt_rms_psu"] = float(np.sqrt(np.mean((sc.fields(["salt_cell"])["salt_cell"] - ref_fields["salt_cell"]) ** 2)))
    iters = []
    for _ in range(3):
        sc.step()
        iters.append(sc.model.pcg_iterations())
    sc.model.close()
    ms = 1e3 * el / steps
    # the same FP64 step with the PCG stop tolerance halved: how far the FP64 state itself
    # moves within the solver's tolerance (the scale the mixed-precision errors compare to)
    tight = dataclasses.replace(case, error_tolerance=0.5 * case.error_tolerance)
    sc = Scenario(tight, precision=0, wind=wind)
    sc.run(warmup + steps)
    sens = field_errors(sc.fields(MIXED_FIELDS), ref_fields)
    sc.model.close()
    return {"value": 0.5 * (a0 + a1) * steps / el, "unit": "cell-layer-steps/s", "ms_per_step": ms,
            "speedup_vs_fp64": fp64_ms / ms, "pcg_inner_iterations": sum(iters) / len(iters),
            "errors_vs_fp64_after_timed_steps": errs,
            "fp64_half_tolerance_vs_fp64": sens}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--config", choices=["storm", "estuary"], default="storm",
                    help="storm: the 1.2M-cell baroclinic storm-surge configuration (default); "
                         "estuary: the synthetic estuary of --cells")
    ap.add_argument("--cells", default="200k")
    ap.add_argument("--turbulence", choices=sorted(TURBULENCE), default="gls",
                    help="vertical closure of the workload (gls: k-kl GLS, Galperin stability functions)")
    ap.add_argument("--storm", action="store_true",
                    help="cyclone forcing and stratified T/S (the physics of the 1.2M-cell configuration)")
    ap.add_argument("--transport", choices=["tvd", "elm"], default="tvd",
                    help="scalar transport solver: TVD (transport_solver 1) or ELM (2)")
    ap.add_argument("--pcg-tolerance", type=float, default=PCG_TOLERANCE,
                    help="PCG stop tolerance on r.M^-1 r / r0.M^-1 r0 (default: the reference's main.inp)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=5,
                    help="untimed steps after the timed ones, for the stage and kernel timers (>= 1)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-loop sample")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble sample")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision sample")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run one device-to-device copy of MIB MiB (calibrates rocprofv3 "
                         "FETCH_SIZE/WRITE_SIZE against known traffic)")
    ap.add_argument("--graph", action="store_true",
                    help="submit each timed step as a captured hipGraph (a batch of this one scenario)")
    ap.add_argument("--counter-pass", type=int, default=0, metavar="STEPS",
                    help="instead of the benchmark, run STEPS steps (default launch shapes) and print "
                         "each step's PCG iteration count, for normalising rocprofv3 --pmc counters")
    args = ap.parse_args()
    if args.profile_steps < 1:
        ap.error("--profile-steps must be at least 1: the roofline is read from those steps' timers")
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (got --gpus {args.gpus}); "
              "the step would shard by domain decomposition, which is not built.", file=sys.stderr)
        return 2
    from gom_release_amd import host
    from gom_release_amd.driver import Scenario
    if host.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        return 1
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration << 20)
        print(f"[bench] calibration copy of {args.copy_calibration} MiB done", file=sys.stderr, flush=True)
    storm_cfg = args.config == "storm"
    wind = (0.0, 0.0) if (args.storm or storm_cfg) else (10.0, 0.0)
    if storm_cfg:
        case = storm_case(args.pcg_tolerance, args.turbulence)
    else:
        case = make_case(args.cells, args.turbulence, args.pcg_tolerance,
                         storm_steps=(4000 + args.warmup + args.steps + args.profile_steps) if args.storm else 0,
                         transport=args.transport)
    prec = 0 if args.precision == "fp64" else 1
    print(f"[bench] building {case.maxele}-cell x {case.maxlayer}-layer case", file=sys.stderr, flush=True)
    sc = Scenario(case, precision=prec, wind=wind)
    model = sc.model
    if args.counter_pass > 0:
        its = []
        for _ in range(args.counter_pass):
            sc.step()
            its.append(model.pcg_iterations())
            print(f"[bench] counter pass step {len(its)}: {its[-1]} PCG iterations", file=sys.stderr, flush=True)
        print(json.dumps({"counter_pass": {"steps": len(its), "pcg_iterations": its, "cells": case.maxele}}))
        model.close()
        return 0
    launch = None
    if not args.no_autotune:
        from gom_release_amd import autotune
        t_tune = time.perf_counter()
        launch = autotune.tuned(sc, log=lambda m: print(m, file=sys.stderr, flush=True))
        print(f"[bench] launch shapes tuned in {time.perf_counter() - t_tune:.1f} s: {launch}",
              file=sys.stderr, flush=True)
    graph = None
    if args.graph:
        # the step submitted as a captured graph: a batch of this one scenario
        graph = host.GpuBatch()
        graph.add(model)

    def run(n):
        if graph is None:
            sc.run(n)
            return
        for _ in range(n):
            sc.stage_forcing()
            graph.step()

    run(args.warmup)
    model.sync()
    print(f"[bench] warmup done ({args.warmup} steps)", file=sys.stderr, flush=True)
    a0 = sc.active_cell_layers()
    model.elm_near_edge(reset=True)
    model.sync()
    t0 = time.perf_counter()
    run(args.steps)
    if graph is not None:
        graph.sync()
    model.sync()
    t1 = time.perf_counter()
    if graph is not None:
        graph.close()
    near = model.elm_near_edge()
    a1 = sc.active_cell_layers()
    elapsed = t1 - t0
    active = 0.5 * (a0 + a1) * args.steps
    value = active / elapsed
    print(f"[bench] timed {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)
    ref_fields = sc.fields(MIXED_FIELDS) if (prec == 0 and not args.no_mixed) else None

    # measurement pass: per-stage and per-kernel HIP event times (synchronising)
    model.timings(True)
    iters = []
    stage_tot = {}
    for _ in range(args.profile_steps):
        sc.step()
        iters.append(model.pcg_iterations())
        for k, v in model.timings(True).items():
            stage_tot[k] = stage_tot.get(k, 0.0) + v
    tim = model.timings(False)
    kernels = {k: tim[k] for k in ("k_pcg", "k_elm", "k_momentum", "k_tvd_column")}
    stages = {k: stage_tot[k] / args.profile_steps for k in stage_tot if not k.startswith("k_")}
    it_mean = sum(iters) / len(iters)
    pcg_bytes = pcg_bytes_per_iteration(sc.mesh, sc.params) * it_mean
    pcg_ms = kernels["k_pcg"]
    achieved = pcg_bytes / (pcg_ms * 1e-3) / 1e9
    traffic = None
    if sc.params.maxele == PCG_TRAFFIC_CELLS and prec == 0:
        traffic = round(PCG_TRAFFIC_LAUNCH + PCG_TRAFFIC_PER_ITERATION * it_mean)
    kname = "k_pcg_cg1m" if sc.params.maxele > 524288 else "k_pcg_cg1"
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": f"{kname} (PCG solve; timer k_pcg)", "pcg_iterations": round(it_mean, 1),
            "kernel_ms": {k: round(v, 4) for k, v in kernels.items()},
            "stage_ms": {k: round(v, 4) for k, v in stages.items()}}
    F, Cl, N = sc.active_layers()
    zbar = pcg_bytes_per_iteration(sc.mesh, sc.params) / sc.params.maxele / 12.0 - 9.0
    bstep = step_bytes(F, Cl, N, sc.params.maxface, sc.params.maxele, zbar, max(case.trans_sub_iter, 1),
                       int(sc.params.maxtran2), it_mean, closure=case.turbulence_flag != 0)
    step_ms = 1e3 * elapsed / args.steps
    roof["step"] = {"algorithmic_bytes": round(bstep), "achieved": round(bstep / (step_ms * 1e-3) / 1e9, 1),
                    "frac": round(bstep / (step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                    "active_face_cell_node_layers": [F, Cl, N]}
    if args.transport == "elm":     # the byte model counts the TVD transport kernels
        roof["step"] = None
    vol, salt = model.budgets()
    out = {
        "metric": "cell-layer-steps/s at 1 GPU + %HBM roofline; p50/p99 scenario latency",
        "value": value, "unit": "cell-layer-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if prec == 0 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": (f"baroclinic storm surge, {case.maxele} mixed tri/quad cells (100 m) x "
                                f"{case.maxlayer} z-layers, seeded fractal bathymetry 1-40 m, Holland cyclone "
                                f"(950 hPa, r_max 40 km, B 1.5, 5 m/s; fields every 10 steps), M2+K1, wet/dry, "
                                f"stratified T/S (10 C, 3 psu) by TVD minmod x{case.trans_sub_iter} sub-cycles, "
                                f"dt={case.dt:g} s, "
                                + ("GLS k-kl turbulence closure (Galperin)" if args.turbulence == "gls"
                                   else "constant Av/Kv")
                                + f", PCG tolerance {case.error_tolerance:g}") if storm_cfg else (
                               f"synthetic estuary, {case.maxele} mixed tri/quad cells x {case.maxlayer} "
                               f"z-layers, dt={case.dt:g} s, "
                               + ("ELM scalar transport, uniform 20 C, 5 psu river" if args.transport == "elm"
                                  else "TVD minmod x2 sub-cycles")
                               + ", M2+K1, 500 m3/s river, "
                               + ("Holland cyclone (950 hPa, r_max 40 km, B 1.5, 5 m/s; fields every 10 steps), "
                                  "stratified T/S (10 C, 3 psu) transported, " if args.storm else "10 m/s wind, ")
                               + ("GLS k-kl turbulence closure (Galperin)"
                                                   if args.turbulence == "gls" else "constant Av/Kv")
                               + f", PCG tolerance {case.error_tolerance:g}"),
                   "turbulence": args.turbulence, "transport": args.transport,
                   "cells": case.maxele, "faces": sc.params.maxface, "nodes": case.maxnod,
                   "layers": case.maxlayer, "active_cell_layers": int(0.5 * (a0 + a1)),
                   "nominal_cell_layer_steps_per_s": case.maxele * case.maxlayer * args.steps / elapsed,
                   "pcg_tolerance": case.error_tolerance, "precision": args.precision,
                   "volume_m3": vol, "salt_psu_m3": salt},
        "roofline": roof,
        "launch": launch,
        # ELM departure points within 1e-12 of a cell edge over the timed steps (their
        # departure cell rests on the last bit of a coordinate), of the active face-layers
        "elm_near_edge": {"momentum": near[0], "tracer": near[1], "face_layer_steps": F * args.steps},
    }
    model.close()
    if ref_fields is not None:
        print("[bench] mixed-precision sample", file=sys.stderr, flush=True)
        out["mixed"] = mixed_sample(case, wind, args.warmup, args.steps, ref_fields, 1e3 * elapsed / args.steps,
                                    launch)
        ref_fields = None
    if not args.no_ensemble:
        print("[bench] batched ensemble sample", file=sys.stderr, flush=True)
        out["ensemble"] = ensemble_sample()
    if not args.no_serving:
        print("[bench] serving-loop sample", file=sys.stderr, flush=True)
        out["serving"] = serving_sample()
        out["config"]["p50_latency_s"] = out["serving"]["p50_latency_s"]
        out["config"]["p99_latency_s"] = out["serving"]["p99_latency_s"]
    if not args.no_cpu_baseline:
        print("[bench] timing the CPU restatement", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(case, wind, label=f"{case.maxele} cells x {case.maxlayer} layers")
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
