"""The CPU oracle behind the interface of gom_release_amd.host.GpuModel (test helper).

Lets the host driver (gom_release_amd.driver.Scenario) run on the oracle, so that the
driver's own logic (forcing, output schedule, restart and dump files) can be checked on
the CPU against files the reference wrote. The river face velocity comes from the host,
which is what the reference does (compute_boundary_velocity.f90:32-46).
"""
from __future__ import annotations

from gom_release_amd import abi
from oracle import Oracle


class OracleModel:
    host_river_velocity = True

    def __init__(self, mesh: "abi.Arrays", params: "abi.Params"):
        self.mesh, self.params = mesh, params
        self.state = abi.state_arrays(params)
        self.forcing = abi.ForcingArrays(params)
        self.o = Oracle(mesh, params, self.state, self.forcing)

    def set_state(self, s, only=None) -> None:
        for f in (only or s.arr):
            self.state[f] = s[f]

    def get_state(self, s, only=None) -> None:
        for f in (only or s.arr):
            s[f] = self.state[f]

    def set_forcing(self, f, only=None) -> None:
        for n in (only or f.arr):
            self.forcing[n] = f[n]
        self.forcing.scalars.update(f.scalars)

    def stage(self, name: str) -> None:
        rc = self.o.run(name)
        if rc != 0:
            raise RuntimeError(f"oracle stage {name} returned {rc}")

    def step(self) -> None:
        self.stage("step")

    def close(self) -> None:
        pass
