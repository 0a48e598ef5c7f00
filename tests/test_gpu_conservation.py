"""Volume and salt conservation on the device (closed basin, see test_conservation.py):
the device's FP64 budget reduction (gomgpu_budgets) agrees with the budgets of the state it
holds, and both totals stay fixed over the run to the same bounds as the oracle's."""
from __future__ import annotations

import numpy as np
import pytest

from test_conservation import STEPS, budgets, closed_basin

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu_available():
    from gom_release_amd import host
    if host.device_count() < 1:
        pytest.fail("no GPU visible: the HIP library cannot run")
    return True


def test_gpu_conserves_volume_and_salt(gpu_available):
    from gom_release_amd.driver import Scenario
    print("[gpu-budget] closed basin", flush=True)
    sc = Scenario(closed_basin(), wind=(8.0, 3.0))
    v0, s0 = sc.model.budgets()
    nv0, ns0 = budgets(sc)
    np.testing.assert_allclose([v0, s0], [nv0, ns0], rtol=1e-12)
    sc.run(STEPS)
    v1, s1 = sc.model.budgets()
    nv1, ns1 = budgets(sc)
    np.testing.assert_allclose([v1, s1], [nv1, ns1], rtol=1e-12)
    assert abs(v1 - v0) <= 1e-10 * v0
    assert abs(s1 - s0) <= 1e-12 * s0
    sc.model.close()
