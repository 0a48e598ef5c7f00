"""The oracle against the reference's own outputs: every stage, bit for bit.

The fixtures hold the reference's module state at full precision (tests/golden); each
stage of the oracle is run on the reference's input state and must reproduce the fields
the reference wrote exactly.
"""
import numpy as np
import pytest

from golden_io import FIXTURES, STAGES, Golden, apply
from oracle import Oracle


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_stages_bitwise(name):
    g = Golden(name)
    case = g.case()
    for it in g.steps:
        state = g.state_in(it)
        forcing = g.forcing(it, case)
        for stage, tag in STAGES:
            o = Oracle(g.mesh, g.params, state, forcing)
            assert o.run(stage) == 0, (name, it, stage)
            if stage == "freesurf" and g.pcg_iterations(it) is not None:
                # the reference's cgit2 (pre_conj_grad.f90:224-229)
                assert Oracle.pcg_iterations() == g.pcg_iterations(it), (name, it)
            for field, ref in g.expected(it, tag).items():
                if field not in state.arr:
                    continue
                got = state[field]
                bad = np.flatnonzero(got != ref)
                assert bad.size == 0, f"{name} step {it} {stage} {field}: {bad.size} differ, first {bad[:5]}"
            apply(state, g.expected(it, tag))


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_whole_steps_bitwise(name):
    """Chained: the oracle's step from the reference state equals the reference's post state."""
    g = Golden(name)
    case = g.case()
    for it in g.steps:
        state = g.state_in(it)
        o = Oracle(g.mesh, g.params, state, g.step_forcing(it, case))
        assert o.step() == 0
        for field, ref in g.expected(it, "post").items():
            assert np.array_equal(state[field], ref), (name, it, field)
