"""Device checks of the kernels' rewritten arithmetic against the IEEE
operations they replace (tests/hip/capaz_selftest.hip): the reciprocal used
for Gauss-Jordan pivots and the exact division built on it, over random
operands including all-ones significands and powers of two, and the square
root and square that stand for pow(x, 0.5) and pow(t, 2.) in the van
Genuchten closures, and exp(y * log|x|), which stands for the closures' other
powers."""
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


def device_pow(x, e):
    from capaz_amd import capaz
    capaz.load()
    L = ctypes.CDLL(LIB)
    L.capaz_selftest_pow.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_ulonglong]
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    assert L.capaz_selftest_pow(x.ctypes.data, e, y.ctypes.data, x.size) == 0
    return y


# The closures' exponents: the reference Three-phase deck's van Genuchten
# a = A/(A-1), b = (A-1)/A, c = 2(A-1)/A with A = 3.25; mu_g(T)'s 0.683; a
# Corey exponent; Brooks-Corey-type exponents; integers.
POW_EXPONENTS = [3.25 / 2.25, 2.25 / 3.25, 4.5 / 3.25, 0.683, 2.5, 3.0 + 2.0 / 2.7, 1.0 / 2.7, 2.0, 3.0, 0.5]


@pytest.mark.parametrize("e", POW_EXPONENTS)
def test_closure_power_is_within_a_few_ulp_of_pow(e):
    """pow_cl(x, e) = exp(e * log|x|) on the device against numpy's pow, on
    the closures' bases: within 2 (|e ln x| + 1) + 1 ulp -- the roundings of
    log x and of e * log x are absolute errors of up to |e ln x| 2^-53 each in
    the exponent, which exp turns into relative ones; exp(e * log(x)) from
    numpy's correctly rounded log and exp reaches 1.8 (|e ln x| + 1) on these
    operands -- and pow's special cases exactly: 1 for
    x = 1, 0 for x = 0, NaN for a NaN or a negative base with a non-integer
    exponent, +-|x|^e for a negative base with an integer one."""
    x = closure_operands(1 << 20, 7)
    x = x[np.isfinite(x) & (x > 0)]
    got = device_pow(x, e)
    want = np.power(x, e)
    ok = want > 0
    ulps = np.abs(got[ok] - want[ok]) / np.spacing(want[ok])
    bound = 2.0 * (np.abs(e * np.log(x[ok])) + 1.0) + 1.0
    assert (ulps <= bound).all(), (e, float(ulps.max()), float(x[ok][np.argmax(ulps - bound)]))
    special = np.array([1.0, 0.0, -0.0, np.nan, -2.0, -0.5, np.inf])
    gs = device_pow(special, e)
    with np.errstate(invalid="ignore", divide="ignore"):
        ws = np.power(special, e)
    assert gs[0] == 1.0 and gs[1] == 0.0 and gs[2] == 0.0 and np.isnan(gs[3]) and gs[6] == np.inf
    for k in (4, 5):
        if e == int(e):
            assert np.sign(gs[k]) == np.sign(ws[k]) and abs(gs[k] - ws[k]) <= 4 * np.spacing(abs(ws[k])), (k, gs[k], ws[k])
        else:
            assert np.isnan(gs[k])


def test_closure_power_with_a_zero_exponent_is_one():
    x = np.array([0.0, 1.0, 0.3, -2.0, np.nan, np.inf])
    assert (device_pow(x, 0.0) == 1.0).all()


def test_closure_log_is_within_two_ulp_of_log():
    """log_cl, the log of the closures' powers, against numpy's log: within 2
    ulp (the series' argument s carries ~2^-52 relative error, which shows
    where ln x is small), on closure operands, normal and subnormal numbers
    across the range, and exactly for the special values."""
    from capaz_amd import capaz
    capaz.load()
    L = ctypes.CDLL(LIB)
    L.capaz_selftest_log.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_ulonglong]
    rng = np.random.default_rng(9)
    wide = np.exp(rng.uniform(np.log(5e-324), np.log(1.7e308), 1 << 20))
    x = np.concatenate([closure_operands(1 << 20, 8), wide])
    x = np.ascontiguousarray(x[np.isfinite(x) & (x > 0)])
    y = np.empty_like(x)
    assert L.capaz_selftest_log(x.ctypes.data, y.ctypes.data, x.size) == 0
    want = np.log(x)
    nz = want != 0
    ulps = np.abs(y[nz] - want[nz]) / np.spacing(np.abs(want[nz]))
    assert ulps.max() <= 2.0, (float(ulps.max()), float(x[nz][np.argmax(ulps)]))
    assert (y[~nz] == 0).all()
    special = np.array([0.0, -0.0, np.inf, -1.0, np.nan])
    ys = np.empty_like(special)
    assert L.capaz_selftest_log(special.ctypes.data, ys.ctypes.data, special.size) == 0
    assert ys[0] == -np.inf and ys[1] == -np.inf and ys[2] == np.inf and np.isnan(ys[3]) and np.isnan(ys[4])
