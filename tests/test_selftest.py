"""Device checks of the kernels' rewritten arithmetic against the IEEE
operations they replace (tests/hip/capaz_selftest.hip): the reciprocal used
for Gauss-Jordan pivots and the exact division built on it, over random
operands including all-ones significands and powers of two."""
import ctypes
import os

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
