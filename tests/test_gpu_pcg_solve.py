"""The standalone free-surface solve (gomgpu_pcg_solve, replacing pre_conj_grad.f90:9)
against the oracle's restatement of pre_conj_grad on the same CSR systems.

The device runs the single-reduction form of the same Jacobi PCG, so its iterates equal the
reference's in exact arithmetic; the dot products are associated differently, so x is held
to 1e-9 relative (the free-surface tolerance of the stage tests) and the iteration count,
which the stop test sqrt(delta^2/delta_0^2) < tol fixes, must be the oracle's. Sizes cover
the register-resident solve (up to 524,288 rows) and the memory-resident one above it.
Malformed matrices are refused before any solve runs, and non-convergence is reported.
"""
from __future__ import annotations

import numpy as np
import pytest

from test_oracle_kat import _pcg, _spd_csr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu_available():
    from gom_release_amd import host
    if host.device_count() < 1:
        pytest.fail("no GPU visible: the HIP library cannot run")
    return True


def grid_system(nx: int, ny: int, seed: int):
    """A free-surface-shaped system on an nx x ny grid of cells: A_ii = area + sum c_j,
    A_ij = -c_j with c_j > 0 per interior face, absent neighbours left out, 1-based CSR with
    the diagonal first and the neighbours in west, east, south, north order
    (solve_free_surface_equation.f90:233-258). Returns (vals, cols, rowstr, b, x0)."""
    rng = np.random.default_rng(seed)
    n = nx * ny
    cx = rng.uniform(2e5, 2e6, (ny, nx - 1))        # faces between (i, j) and (i+1, j)
    cy = rng.uniform(2e5, 2e6, (ny - 1, nx))
    area = rng.uniform(0.8e4, 1.2e4, (ny, nx))
    diag = area.copy()
    diag[:, :-1] += cx
    diag[:, 1:] += cx
    diag[:-1, :] += cy
    diag[1:, :] += cy
    idx = np.arange(n).reshape(ny, nx)
    nbr = np.zeros((4, ny, nx), dtype=np.int64)      # 1-based neighbour, 0 = none
    cof = np.zeros((4, ny, nx))
    nbr[0][:, 1:] = idx[:, :-1] + 1
    cof[0][:, 1:] = -cx
    nbr[1][:, :-1] = idx[:, 1:] + 1
    cof[1][:, :-1] = -cx
    nbr[2][1:, :] = idx[:-1, :] + 1
    cof[2][1:, :] = -cy
    nbr[3][:-1, :] = idx[1:, :] + 1
    cof[3][:-1, :] = -cy
    nbr = nbr.reshape(4, n)
    cof = cof.reshape(4, n)
    present = nbr > 0
    cnt = 1 + present.sum(axis=0)
    rowstr = np.concatenate([[1], 1 + np.cumsum(cnt)]).astype(np.int32)
    vals = np.empty(int(cnt.sum()))
    cols = np.empty(int(cnt.sum()), dtype=np.int32)
    start = rowstr[:-1] - 1
    vals[start] = diag.ravel()
    cols[start] = np.arange(1, n + 1)
    pos = start + 1
    for l in range(4):
        m = present[l]
        vals[pos[m]] = cof[l][m]
        cols[pos[m]] = nbr[l][m]
        pos = pos + m
    eta = 0.3 * np.sin(np.linspace(0, 6, n)) + 0.01 * rng.normal(size=n)
    b = area.ravel() * eta
    x0 = eta + 0.02 * rng.normal(size=n)
    return vals, cols, rowstr, b, x0


def _check(vals, cols, rows, b, x0, maxit, tol):
    from gom_release_amd import host
    it_ref, x_ref = _pcg(b.size, vals, cols, rows, b, x0, maxit, tol)
    rc, it, x = host.pcg_solve(vals, cols, rows, b, x0, maxit, tol)
    print(f"[pcg-solve] n={b.size} tol={tol} iterations gpu {it} oracle {it_ref}", flush=True)
    assert rc == 0 and it > 0
    assert it == it_ref
    np.testing.assert_allclose(x, x_ref, rtol=1e-9, atol=1e-12 * np.max(np.abs(x_ref)))


@pytest.mark.parametrize("n,seed", [(50, 1), (400, 2)])
def test_pcg_solve_random_spd(gpu_available, n, seed):
    _, vals, cols, rows, = _spd_csr(n, seed)[0:4]
    b = np.random.default_rng(seed).normal(size=n)
    _check(vals, cols, rows, b, np.zeros(n), 1000, 1e-10)


@pytest.mark.parametrize("nx,ny,tol", [(40, 25, 5e-6), (500, 400, 5e-6), (1000, 600, 5e-6), (1000, 600, 1e-10)])
def test_pcg_solve_grid(gpu_available, nx, ny, tol):
    """1,000 and 200,000 rows (register-resident) and 600,000 rows (memory-resident)."""
    _check(*grid_system(nx, ny, nx + ny), 1000, tol)


def test_pcg_solve_not_converged_is_reported(gpu_available):
    from gom_release_amd import host
    vals, cols, rows, b, x0 = grid_system(30, 20, 7)
    rc, it, x = host.pcg_solve(vals, cols, rows, b, x0, 1, 1e-14)
    assert rc == 5 and it == 0
    assert np.all(np.isfinite(x))


@pytest.mark.parametrize("bad", ["no_diagonal", "six_entries", "column_out_of_range", "two_diagonals"])
def test_pcg_solve_refuses_malformed_rows(gpu_available, bad):
    from gom_release_amd import host
    vals, cols, rows, b, x0 = grid_system(12, 10, 3)
    vals, cols, rows = vals.copy(), cols.copy(), rows.copy()
    r = 30                                         # an interior row: diagonal + 4 neighbours
    k0 = rows[r] - 1
    if bad == "no_diagonal":
        cols[k0] = cols[k0 + 1]
    elif bad == "six_entries":
        rows[r + 1:] += 1
        vals = np.insert(vals, k0 + 5, -1.0)
        cols = np.insert(cols, k0 + 5, 1)
    elif bad == "column_out_of_range":
        cols[k0 + 2] = b.size + 5
    else:
        cols[k0 + 1] = r + 1
    with pytest.raises(host.GomGpuError) as e:
        host.pcg_solve(vals, cols, rows, b, x0, 100, 1e-8)
    assert e.value.code == 1 and f"row {r + 1}" in str(e.value)
