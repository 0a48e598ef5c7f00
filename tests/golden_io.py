"""Loading of the golden fixtures (tests/golden/*.npz) into the ABI arrays."""
from __future__ import annotations

import json
import os
from typing import Dict, List

import numpy as np

from gom_release_amd import abi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STAGES = [("friction", "pre"), ("elm", "elm"), ("momentum", "mom"), ("freesurf", "fs"),
          ("velocity", "vel"), ("layers", "lay"), ("nodevel", "nodev"), ("transport", "tvd"),
          ("eos", "eos"), ("update", "post")]
FIXTURES = ["channel_16x8x5", "estuary_wetdry"]


def _flat(a):
    a = np.asarray(a)
    return np.asfortranarray(a).ravel(order="F") if a.ndim == 2 else a.ravel()


class Golden:
    def __init__(self, name: str):
        self.name = name
        z = np.load(os.path.join(GOLDEN, name + ".npz"))
        self.data = {k: z[k] for k in z.files}
        self.meta = json.load(open(os.path.join(GOLDEN, name + ".json")))
        self.steps: List[int] = self.meta["steps"]
        g = {k[5:]: v for k, v in self.data.items() if k.startswith("geom/")}
        self.geom = g
        import refdump  # oracle helper: dump dict -> ABI structs
        self.params = refdump.params_from_geom(g)
        self.mesh = refdump.mesh_from_geom(g, self.params)

    def group(self, tag: str) -> Dict[str, np.ndarray]:
        pre = tag + "/"
        return {k[len(pre):]: v for k, v in self.data.items() if k.startswith(pre)}

    def state_in(self, it: int) -> "abi.Arrays":
        s = abi.state_arrays(self.params)
        for k, v in self.group(f"s{it}_in").items():
            if k in s.arr:
                s[k] = _flat(v)
        return s

    def forcing(self, it: int, case=None) -> "abi.ForcingArrays":
        import refdump
        pre = self.group(f"s{it}_pre")
        ob = qb = None
        if case is not None:
            from gom_release_amd import forcing as fz
            p = self.params
            ob, qb = fz.tvd_boundary_concentrations(case, float(pre["julian_day"]),
                                                    float(pre["spinup_function_baroclinic"]),
                                                    p.maxtran2, [p.tran_id1, p.tran_id2])
        return refdump.forcing_from_dump(pre, self.params, ob, qb)

    def expected(self, it: int, tag: str) -> Dict[str, np.ndarray]:
        out = {}
        for k, v in self.group(f"s{it}_{tag}").items():
            out[k] = _flat(v)
        return out

    def case(self):
        from gom_release_amd import meshgen
        m = self.meta
        return meshgen.channel_case(**m["args"]) if m["kind"] == "channel" else meshgen.estuary_case(**m["args"])


def apply(state: "abi.Arrays", values: Dict[str, np.ndarray]) -> None:
    for k, v in values.items():
        if k in state.arr:
            state[k] = v
