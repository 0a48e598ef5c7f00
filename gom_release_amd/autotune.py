"""Autotuner of the launch shapes (threads per block) of the step's kernels.

Coordinate descent over the tunable launches of gomgpu_set_launch: for each kernel in turn
every candidate block size is timed over a few whole steps (the median step time, the
others held at their current best), and the fastest kept. No launch shape changes a result
(each thread computes its own outputs; the PCG partition is not tunable), so tuning runs on
the live scenario: its state is saved before and restored after, and the tuned scenario
continues exactly as an untuned one would. Results are cached per device and problem size
in a JSON file, so later scenarios of the same size class start tuned.
"""
from __future__ import annotations

import json
import os
import time
from typing import Dict, Optional, Sequence

import numpy as np

from . import abi, host

CANDIDATES = (64, 128, 256, 512, 1024)
DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "gom_release_amd", "launch.json")


def size_key(params: "abi.Params") -> str:
    """Size class: cells rounded to a power of two, layers, closure on/off."""
    e = 1 << int(round(np.log2(max(params.maxele, 1))))
    return f"E{e}_L{params.maxlayer}_T{params.turbulence_flag}"


def _device_name() -> str:
    try:
        import ctypes as C
        L = host.lib()
        # hipDeviceProp_t is large; only the name (first 256 bytes) is needed
        buf = C.create_string_buffer(4096)
        if L.hipGetDevicePropertiesR0600(buf, 0) == 0:
            return buf.value.decode(errors="ignore") or "gpu"
    except Exception:
        pass
    return "gpu"


def _step_ms(sc, n: int) -> float:
    """Median wall time of one device step with the forcing held fixed (the scenario's
    host clock and its hurricane snapshots do not advance)."""
    times = []
    for _ in range(n):
        t0 = time.perf_counter()
        sc.model.step()
        sc.model.sync(allow_pcg_warning=True)
        times.append(time.perf_counter() - t0)
    return 1e3 * float(np.median(times))


def tune(sc, steps: int = 3, candidates: Sequence[int] = CANDIDATES, log=None) -> Dict[str, int]:
    """Tune the launch shapes of scenario `sc` in place; returns the chosen configuration."""
    model = sc.model
    saved = abi.state_arrays(sc.params)
    model.get_state(saved)
    saved_f = abi.ForcingArrays(sc.params)    # the wind stage's fields evolve on the device
    model.get_forcing(saved_f)
    saved_f.scalars = dict(sc.forcing.scalars)
    _step_ms(sc, 1)  # warm
    best_ms = _step_ms(sc, steps)
    for which in host.LAUNCH:
        cur = model.get_launch(which)
        best_bs = cur
        for bs in candidates:
            if bs == cur or bs > host.LAUNCH_MAX[which]:
                continue
            model.set_launch(which, bs)
            ms = _step_ms(sc, steps)
            if ms < best_ms * 0.99:
                best_ms, best_bs = ms, bs
        model.set_launch(which, best_bs)
        if log:
            log(f"[tune] {which}: {best_bs} ({best_ms:.3f} ms/step)")
    model.set_state(saved)
    model.set_forcing(saved_f)
    return model.launch_config()


def load(params: "abi.Params", path: Optional[str] = None) -> Optional[Dict[str, int]]:
    path = path or os.environ.get("GOMGPU_TUNE_CACHE", DEFAULT_CACHE)
    try:
        db = json.load(open(path))
    except (OSError, ValueError):
        return None
    return db.get(_device_name(), {}).get(size_key(params))


def save(params: "abi.Params", cfg: Dict[str, int], path: Optional[str] = None) -> None:
    path = path or os.environ.get("GOMGPU_TUNE_CACHE", DEFAULT_CACHE)
    try:
        db = json.load(open(path))
    except (OSError, ValueError):
        db = {}
    db.setdefault(_device_name(), {})[size_key(params)] = cfg
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(db, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def apply(model: "host.GpuModel", cfg: Dict[str, int]) -> None:
    for k, v in cfg.items():
        if k in host.LAUNCH:
            model.set_launch(k, v)


def tuned(sc, cache: Optional[str] = None, steps: int = 3, log=None) -> Dict[str, int]:
    """The cached configuration of the scenario's size class, tuning it first if absent."""
    cfg = load(sc.params, cache)
    if cfg is None:
        cfg = tune(sc, steps=steps, log=log)
        save(sc.params, cfg, cache)
    else:
        apply(sc.model, cfg)
    return cfg
