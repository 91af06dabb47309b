"""Device checks of the kernels' rewritten arithmetic against the IEEE
operations they replace (tests/hip/capaz_selftest.hip): the reciprocal used
for Gauss-Jordan pivots and the exact division built on it, over random
operands including all-ones significands and powers of two, and the square
root and square that stand for pow(x, 0.5) and pow(t, 2.) in the van
Genuchten closures."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(__file__), "hip", "libcapaz_selftest.so")


def selftest(which, n, seed, emin, emax):
    from capaz_amd import capaz
    capaz.load()  # one HIP runtime for the process
    L = ctypes.CDLL(LIB)
    L.capaz_selftest.restype = ctypes.c_longlong
    L.capaz_selftest.argtypes = [ctypes.c_int, ctypes.c_ulonglong, ctypes.c_ulonglong, ctypes.c_int, ctypes.c_int]
    return L.capaz_selftest(which, n, seed, emin, emax)


@pytest.mark.parametrize("emin,emax", [(-60, 60), (-1000, 1000), (-3, 3)])
def test_rcp_rn_is_correctly_rounded(emin, emax):
    assert selftest(0, 1 << 30, 12345 + emin, emin, emax) == 0


@pytest.mark.parametrize("emin,emax", [(-60, 60), (-400, 400)])
def test_exact_div_with_rcp_rn_is_ieee_division(emin, emax):
    assert selftest(1, 1 << 29, 777 + emin, emin, emax) == 0


def device_map(which, x):
    from capaz_amd import capaz
    capaz.load()
    L = ctypes.CDLL(LIB)
    L.capaz_selftest_map.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_ulonglong]
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    assert L.capaz_selftest_map(which, x.ctypes.data, y.ctypes.data, x.size) == 0
    return y


def closure_operands(n, seed):
    """Saturation-like operands: uniform in [0, 1], their cubes (small
    values), 1 - u near 1, and the special values."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(0.0, 1.0, n)
    special = np.array([0.0, -0.0, 1.0, 1e-3, 5e-324, 2.0 ** -1022, np.inf, -1.0, np.nan, 4.0, 0.25])
    return np.concatenate([u, u ** 3, 1.0 - u * 1e-6, special])


def test_pow_half_is_the_correctly_rounded_square_root():
    """pow(x, 0.5) of the van Genuchten closures is sqrt(x) on the device:
    numpy's IEEE square root bit for bit (with pow's +0 for -0)."""
    x = closure_operands(1 << 22, 5)
    got = device_map(0, x)
    with np.errstate(invalid="ignore"):
        want = np.sqrt(x) + 0.0
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint64), want[~nan].view(np.uint64))


def test_pow_two_is_the_rounded_square():
    x = closure_operands(1 << 22, 6)
    got = device_map(1, x)
    want = x * x
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint64), want[~nan].view(np.uint64))
