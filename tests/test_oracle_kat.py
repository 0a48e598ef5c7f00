"""Known-answer tests of the oracle's building blocks, independent of any reference run.

Each check is chosen so that a plausible slip in the restatement (a dropped term, a swapped
operand, a wrong stop rule) fails it:
  * tridiagonal_solver.f90 (Thomas, several right-hand sides) against dense LU solves;
  * pre_conj_grad.f90 against scipy's sparse direct solve, the one-iteration answer for a
    diagonal matrix, and the stop rule sqrt(delta^2/delta_0^2) < tol on r.M^-1 r probed one
    iteration before and at convergence;
  * check_intersection.f90 and bilinear_interpolation.f90 against closed forms and the
    forward bilinear map;
  * the UNESCO-1980 equation of state of calculate_density_full.f90 against the published
    check values (Millero & Poisson 1981; UNESCO Tech. Pap. Mar. Sci. 44, 1983).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle import lib

P = C.POINTER(C.c_double)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# --------------------------------------------------------------------------- tridiagonal
@pytest.mark.parametrize("n,nc", [(1, 1), (2, 3), (7, 2), (30, 3), (64, 1)])
def test_tridiag_matches_dense_solve(n, nc):
    rng = np.random.default_rng(100 + n)
    a = rng.uniform(-1, 1, n)
    c = rng.uniform(-1, 1, n)
    b = 2.5 + rng.uniform(0, 1, n)  # diagonally dominant: no pivoting needed
    ld = n + 3
    rhs = np.zeros((ld, nc), order="F")
    rhs[:n] = rng.normal(size=(n, nc))
    sol = np.zeros_like(rhs)
    rc = lib().oracle_tridiag(n, nc, ld, _ptr(a), _ptr(b), _ptr(c), _ptr(rhs), _ptr(sol))
    assert rc == 0
    # a(k) multiplies x(k-1), c(k) multiplies x(k+1) (tridiagonal_solver.f90:37-55)
    A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    ref = np.linalg.solve(A, rhs[:n])
    np.testing.assert_allclose(sol[:n], ref, rtol=1e-12, atol=1e-13)


def test_tridiag_reports_zero_pivot():
    a = np.array([0.0, 1.0])
    b = np.array([1.0, 1.0])
    c = np.array([1.0, 0.0])  # second pivot: 1 - 1*1/1 = 0
    rhs = np.ones(2)
    sol = np.zeros(2)
    assert lib().oracle_tridiag(2, 1, 2, _ptr(a), _ptr(b), _ptr(c), _ptr(rhs), _ptr(sol)) == 1
    b0 = np.array([0.0, 1.0])
    assert lib().oracle_tridiag(2, 1, 2, _ptr(a), _ptr(b0), _ptr(c), _ptr(rhs), _ptr(sol)) == 1


# --------------------------------------------------------------------------- PCG
def _spd_csr(n: int, seed: int):
    """Random symmetric, strictly diagonally dominant matrix with <= 4 off-diagonals per
    row (the free-surface pattern), as 1-based CSR with the diagonal first."""
    rng = np.random.default_rng(seed)
    A = np.zeros((n, n))
    for i in range(n):
        for j in rng.choice(n, size=2, replace=False):
            if j != i and np.count_nonzero(A[i]) < 4 and np.count_nonzero(A[j]) < 4:
                v = -rng.uniform(0.1, 1.0)
                A[i, j] = A[j, i] = v
    for i in range(n):
        A[i, i] = -A[i].sum() + rng.uniform(0.05, 2.0)
    vals, cols, rows = [], [], [1]
    for i in range(n):
        vals.append(A[i, i])
        cols.append(i + 1)
        for j in np.nonzero(A[i])[0]:
            if j != i:
                vals.append(A[i, j])
                cols.append(j + 1)
        rows.append(len(vals) + 1)
    return A, np.array(vals), np.array(cols, dtype=np.int32), np.array(rows, dtype=np.int32)


def _pcg(n, vals, cols, rows, b, x0, maxit, tol):
    x = x0.copy()
    it = lib().oracle_pcg(n, _ptr(vals), _ptr(cols), _ptr(rows), _ptr(b), _ptr(x), maxit,
                          C.c_double(tol))
    return it, x


def _precond_ratio(A, b, x0, x):
    d = np.diag(A)
    r0 = b - A @ x0
    r = b - A @ x
    return (r @ (r / d)) / (r0 @ (r0 / d))


@pytest.mark.parametrize("n,seed", [(50, 1), (400, 2)])
def test_pcg_matches_direct_solve_and_stop_rule(n, seed):
    from scipy.sparse import csr_matrix
    from scipy.sparse.linalg import spsolve
    A, vals, cols, rows = _spd_csr(n, seed)
    rng = np.random.default_rng(seed)
    b = rng.normal(size=n)
    x0 = rng.normal(size=n) * 0.1
    tol = 1e-10
    it, x = _pcg(n, vals, cols, rows, b, x0, 1000, tol)
    assert it > 0
    ref = spsolve(csr_matrix(A), b)
    # stop rule: r.M^-1 r has dropped below tol times its initial value (pre_conj_grad.f90:199)
    ratio = _precond_ratio(A, b, x0, x)
    assert ratio < tol * 1.001
    err = np.linalg.norm(x - ref) / np.linalg.norm(ref)
    assert err < 1e-3
    # one iteration earlier the rule must not yet hold: pins "ratio of r.M^-1 r", not its root
    it_short, x_short = _pcg(n, vals, cols, rows, b, x0, it - 1, tol)
    assert it_short == 0
    assert _precond_ratio(A, b, x0, x_short) >= tol * 0.999


def test_pcg_tighter_tolerance_converges_further():
    n = 200
    A, vals, cols, rows = _spd_csr(n, 3)
    b = np.ones(n)
    it1, x1 = _pcg(n, vals, cols, rows, b, np.zeros(n), 1000, 1e-6)
    it2, x2 = _pcg(n, vals, cols, rows, b, np.zeros(n), 1000, 1e-14)
    ref = np.linalg.solve(A, b)
    assert 0 < it1 < it2
    assert np.linalg.norm(x2 - ref) < np.linalg.norm(x1 - ref)
    assert np.linalg.norm(x2 - ref) / np.linalg.norm(ref) < 1e-6


def test_pcg_diagonal_matrix_one_iteration():
    n = 17
    d = np.linspace(1.0, 9.0, n)
    vals = d.copy()
    cols = np.arange(1, n + 1, dtype=np.int32)
    rows = np.arange(1, n + 2, dtype=np.int32)
    b = np.linspace(-3, 5, n)
    it, x = _pcg(n, vals, cols, rows, b, np.zeros(n), 50, 1e-12)
    assert it == 1
    np.testing.assert_allclose(x, b / d, rtol=1e-14)


def test_pcg_not_converged_returns_zero():
    n = 400
    A, vals, cols, rows = _spd_csr(n, 4)
    it, _ = _pcg(n, vals, cols, rows, np.ones(n), np.zeros(n), 2, 1e-12)
    assert it == 0


# --------------------------------------------------------------------------- geometry
def _isect(p1, p2, p3, p4):
    xin, yin = C.c_double(), C.c_double()
    hit = lib().oracle_check_intersection(C.c_double(p1[0]), C.c_double(p2[0]), C.c_double(p3[0]),
                                          C.c_double(p4[0]), C.c_double(p1[1]), C.c_double(p2[1]),
                                          C.c_double(p3[1]), C.c_double(p4[1]), C.byref(xin),
                                          C.byref(yin))
    return hit, xin.value, yin.value


def test_check_intersection_closed_forms():
    assert _isect((0, 0), (2, 2), (0, 2), (2, 0)) == (1, 1.0, 1.0)
    hit, x, y = _isect((0, 0), (4, 0), (1, -1), (1, 3))
    assert hit == 1 and (x, y) == (1.0, 0.0)
    # parallel and collinear-disjoint segments never intersect (determinant zero)
    assert _isect((0, 0), (1, 0), (0, 1), (1, 1))[0] == 0
    # the crossing point lies beyond the end of the first segment
    assert _isect((0, 0), (1, 0), (2, -1), (2, 1))[0] == 0
    # touching counts (inclusive bounds, check_intersection.f90:34-48)
    assert _isect((0, 0), (1, 0), (1, 0), (1, 5))[0] == 1
    hit, x, y = _isect((0, 0), (2, 0), (1, 0), (1, 5))
    assert hit == 1 and (x, y) == (1.0, 0.0)


def test_check_intersection_random_against_parametric_form():
    rng = np.random.default_rng(7)
    for _ in range(500):
        p = rng.uniform(-10, 10, size=(4, 2))
        hit, x, y = _isect(*p)
        d1, d2 = p[1] - p[0], p[3] - p[2]
        M = np.array([[d1[0], -d2[0]], [d1[1], -d2[1]]])
        if abs(np.linalg.det(M)) < 1e-9:
            continue
        t, u = np.linalg.solve(M, p[2] - p[0])
        inside = (-1e-12 <= t <= 1 + 1e-12) and (-1e-12 <= u <= 1 + 1e-12)
        if min(abs(t), abs(t - 1), abs(u), abs(u - 1)) < 1e-9:
            continue  # too close to an end point to judge independently
        assert hit == int(inside)
        if hit:
            q = p[0] + t * d1
            assert abs(x - q[0]) < 1e-9 and abs(y - q[1]) < 1e-9


def _bilinear(quad, x, y):
    xi, eta = C.c_double(), C.c_double()
    sf = np.zeros(4)
    rc = lib().oracle_bilinear(*[C.c_double(v) for v in quad[:, 0]], *[C.c_double(v) for v in quad[:, 1]],
                               C.c_double(x), C.c_double(y), C.byref(xi), C.byref(eta), _ptr(sf))
    return rc, xi.value, eta.value, sf


def _shape(xi, eta):
    return np.array([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta), (1 + xi) * (1 + eta),
                     (1 - xi) * (1 + eta)]) / 4.0


@pytest.mark.parametrize("kind", ["general", "parallelogram", "trapezoid"])
def test_bilinear_inverts_the_forward_map(kind):
    rng = np.random.default_rng({"general": 11, "parallelogram": 12, "trapezoid": 13}[kind])
    for _ in range(200):
        base = np.array([[0.0, 0.0], [100.0, 0.0], [100.0, 100.0], [0.0, 100.0]])
        if kind == "general":
            quad = base + rng.uniform(-20, 20, size=(4, 2))
        elif kind == "parallelogram":
            s = rng.uniform(-30, 30)
            quad = base + np.array([[0, 0], [0, 0], [s, 0], [s, 0]])
        else:  # top edge shorter: dxi or deta vanishes, one linear and one quadratic branch
            s = rng.uniform(5, 30)
            quad = base + np.array([[0, 0], [0, 0], [-s, 0], [s, 0]])
        xi0, eta0 = rng.uniform(-0.95, 0.95, 2)
        N = _shape(xi0, eta0)
        x, y = N @ quad[:, 0], N @ quad[:, 1]
        rc, xi, eta, sf = _bilinear(quad, x, y)
        assert rc == 0
        assert abs(xi - xi0) < 1e-8 and abs(eta - eta0) < 1e-8
        np.testing.assert_allclose(sf, _shape(xi, eta), atol=1e-14)
        assert abs(sf.sum() - 1.0) < 1e-14


def test_bilinear_clamps_points_just_outside():
    quad = np.array([[0.0, 0.0], [100.0, 0.0], [110.0, 100.0], [-5.0, 95.0]])
    N = _shape(1.05, -0.3)  # slightly outside in xi (inside the 1.1 acceptance band)
    rc, xi, eta, sf = _bilinear(quad, N @ quad[:, 0], N @ quad[:, 1])
    assert rc == 0
    assert xi == 1.0 and abs(eta + 0.3) < 1e-8


# --------------------------------------------------------------------------- EOS
def _rho(t, s):
    f = lib().oracle_rho_unesco
    f.argtypes = [C.c_double, C.c_double]
    f.restype = C.c_double
    return f(t, s)


def _rho_p(rho0, t, s, p_bar):
    f = lib().oracle_rho_pressure
    f.argtypes = [C.c_double] * 4
    f.restype = C.c_double
    return f(rho0, t, s, p_bar)


@pytest.mark.parametrize("s,t,expected", [(0.0, 5.0, 999.96675), (35.0, 5.0, 1027.67547),
                                          (35.0, 25.0, 1023.34306)])
def test_unesco_surface_check_values(s, t, expected):
    # the coefficients are single-precision literals in the reference, which moves the
    # result by at most ~1e-4 kg/m3; any dropped or mis-signed term moves it by far more
    assert abs(_rho(t, s) - expected) < 2e-3


def test_unesco_pressure_check_value():
    rho0 = _rho(25.0, 35.0)
    assert abs(_rho_p(rho0, 25.0, 35.0, 1000.0) - 1062.53817) < 5e-3
    assert _rho_p(rho0, 25.0, 35.0, 0.0) == rho0
