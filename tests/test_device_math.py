"""gom_log, the device's natural logarithm (the log-law bottom drag of
calculate_bottom_friction.f90:47-55), run on the host through libgomgpu.so.

The reference takes log from the host C library. The device's own math library is not
correctly rounded and differs from it in the last bit for some arguments, so the device
uses gom_log: a double-double evaluation rounded once. Pins: correctly rounded against
60-digit decimal logarithms; equal to the C library's log on all but a few arguments per
thousand (the C library itself is off the correctly rounded value by a hair over half an
ulp there, which the decimal check shows).
"""
from __future__ import annotations

import ctypes as C
import math
import os
from decimal import Decimal, getcontext

import numpy as np
import pytest

from gom_release_amd import host


def _gom_log(x: np.ndarray) -> np.ndarray:
    if not os.path.exists(host.LIB_PATH):
        pytest.skip("libgomgpu.so not built")
    L = C.CDLL(host.LIB_PATH)
    L.gom_log_array.argtypes = [C.c_void_p, C.c_void_p, C.c_long]
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    L.gom_log_array(x.ctypes.data, y.ctypes.data, x.size)
    return y


def _samples(n, seed):
    rng = np.random.default_rng(seed)
    return np.concatenate([np.exp(rng.uniform(-30, 30, n)), 1.0 + rng.uniform(-1e-4, 1e-4, n // 4),
                           rng.uniform(0.5, 2.0, n // 2),
                           np.array([1.0, 2.0, 0.5, 10.0, 1e-300, 1e300, 1.0 + 2.0 ** -52, 1.0 - 2.0 ** -53])])


def test_gom_log_is_correctly_rounded():
    getcontext().prec = 60
    x = _samples(6000, 1)
    y = _gom_log(x)
    for xi, yi in zip(x, y):
        t = Decimal(float(xi)).ln()
        ulp = Decimal(math.ulp(float(yi))) if yi != 0.0 else Decimal(0)
        if xi == 1.0:
            assert yi == 0.0
            continue
        assert abs(Decimal(float(yi)) - t) <= ulp / 2, (float(xi).hex(), float(yi).hex())


def test_gom_log_agrees_with_the_c_library():
    x = _samples(400000, 2)
    y = _gom_log(x)
    ref = np.array([math.log(v) for v in x])
    diff = y.view(np.int64) - ref.view(np.int64)
    assert np.all(np.abs(diff) <= 1)
    assert np.count_nonzero(diff) <= x.size // 1000


# ---- gom_cos, the device's cosine (the harmonic tide of compute_boundary_eta.f90:72-90)

def _gom_cos(x: np.ndarray) -> np.ndarray:
    if not os.path.exists(host.LIB_PATH):
        pytest.skip("libgomgpu.so not built")
    L = C.CDLL(host.LIB_PATH)
    L.gom_cos_array.argtypes = [C.c_void_p, C.c_void_p, C.c_long]
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    L.gom_cos_array(x.ctypes.data, y.ctypes.data, x.size)
    return y


def _tide_arguments(n, seed):
    """Arguments of the shape 2 pi t / T + V0 - g - shift: up to ~2000 rad for a year of
    tide, with the near-multiples of pi/2 where the reduction cancels most."""
    rng = np.random.default_rng(seed)
    near = np.arange(-1200, 1200) * (math.pi / 2)
    return np.concatenate([rng.uniform(-60.0, 60.0, n), rng.uniform(-2000.0, 2000.0, n // 2),
                           rng.uniform(-1e-3, 1e-3, n // 8), near, np.nextafter(near, 1e9),
                           np.array([0.0, -0.0, 1e-300, 0.5, 1.0, math.pi, 2.0 * math.pi])])


def test_gom_cos_is_correctly_rounded():
    import mpmath
    mpmath.mp.prec = 300
    x = _tide_arguments(3000, 3)
    y = _gom_cos(x)
    for xi, yi in zip(x, y):
        t = mpmath.cos(mpmath.mpf(float(xi)))
        assert abs(mpmath.mpf(float(yi)) - t) <= mpmath.mpf(math.ulp(float(yi))) / 2, (float(xi).hex(), float(yi))


def test_gom_cos_differs_from_the_c_library_only_where_the_library_misrounds():
    import mpmath
    mpmath.mp.prec = 300
    x = _tide_arguments(200000, 4)
    y = _gom_cos(x)
    ref = np.array([math.cos(v) for v in x])
    diff = np.abs(y.view(np.int64) - ref.view(np.int64))
    assert diff.max() <= 1
    bad = np.flatnonzero(diff)
    assert bad.size < 0.005 * x.size          # a few in a thousand
    for i in bad[:200]:
        t = mpmath.cos(mpmath.mpf(float(x[i])))
        err_lib = abs(mpmath.mpf(float(ref[i])) - t) / mpmath.mpf(math.ulp(float(ref[i])))
        assert err_lib > 0.5, float(x[i]).hex()  # the library's value is the misrounded one
