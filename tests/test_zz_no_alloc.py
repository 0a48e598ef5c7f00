"""gomgpu_pcg_solve_noalloc (the solve on the library's fixed device workspace) against
gomgpu_pcg_solve (the same kernels on a workspace allocated per call), through the C ABI.

Both run the same kernels with the same partition on the same rows, so x and the iteration
count must be identical bit for bit. Two calls in a row, the second smaller than the first,
so the second runs on a workspace still holding the first call's values; a size above the
fixed capacity is refused without running anything.
"""
from __future__ import annotations

import numpy as np
import pytest

from test_gpu_pcg_solve import grid_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu_available():
    from gom_release_amd import host
    if host.device_count() < 1:
        pytest.fail("no GPU visible: the HIP library cannot run")
    return True


def test_noalloc_matches_allocating_solve_twice(gpu_available):
    from gom_release_amd import host
    for nx, ny, tol in ((500, 500, 1e-10), (300, 200, 5e-6)):      # 250,000 then 60,000 rows
        sysm = grid_system(nx, ny, 11 * nx + ny)
        print(f"[no-alloc] {nx}x{ny} rows, tol {tol}", flush=True)
        rc_a, it_a, x_a = host.pcg_solve(*sysm, 1000, tol)
        rc_n, it_n, x_n = host.pcg_solve(*sysm, 1000, tol, noalloc=True)
        assert rc_a == 0 and rc_n == 0
        assert it_n == it_a > 0
        assert np.array_equal(x_n.view(np.uint64), x_a.view(np.uint64))


def test_noalloc_refuses_more_rows_than_its_workspace(gpu_available):
    from gom_release_amd import host
    vals, cols, rows, b, x0 = grid_system(600, 500, 5)              # 300,000 rows
    with pytest.raises(host.GomGpuError) as e:
        host.pcg_solve(vals, cols, rows, b, x0, 100, 1e-8, noalloc=True)
    assert e.value.code == 1 and "exceed" in str(e.value)
