"""The C ABI library without a GPU: it loads, exports every function that
include/capaz.h declares, and rejects unsupported configurations with
CAPAZ_ERR_INVALID and a message (validation precedes any device call)."""
import ctypes
import os
import re

import pytest

from capaz_amd import capaz
from capaz_cases import case

HEADER = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "capaz.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(capaz_[A-Za-z_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    L = capaz.load()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing
    assert set(names) == set(capaz.EXPORTS)


def test_abi_version_and_params_defaults():
    L = capaz.load()
    assert L.capaz_abi_version() == 1
    p = capaz.Params()
    L.capaz_params_init(ctypes.byref(p))
    assert p.struct_size == ctypes.sizeof(capaz.Params)
    assert p.model == 3 and p.newton_iterations == 1 and p.corey_n == 2.0


@pytest.mark.parametrize("over,msg", [
    (dict(NY=2), "Ny"),
    (dict(NX=2), "Nx"),
    (dict(NZ=2), "Nz"),
    (dict(MODEL=4), "model"),
    (dict(HX=0.0), "hx"),
    (dict(DT=-1.0), "dt"),
    (dict(ENERGY=1, MODEL=2), "energy"),
    (dict(MODEL=2, NX=1), "Nx"),
])
def test_create_rejects_unsupported(over, msg):
    L = capaz.load()
    keys = case("3ph", 5, 6, 7)
    keys.update(over)
    p = capaz.params_from_keys(keys)
    ctx = ctypes.c_void_p()
    rc = L.capaz_create(ctypes.byref(p), ctypes.byref(ctx))
    assert rc == -1
    assert not ctx.value
    assert msg in L.capaz_last_error(None).decode()


def test_struct_size_mismatch_rejected():
    L = capaz.load()
    p = capaz.params_from_keys(case("3ph", 5, 6, 7))
    p.struct_size = 8
    ctx = ctypes.c_void_p()
    assert L.capaz_create(ctypes.byref(p), ctypes.byref(ctx)) == -1
    assert "struct_size" in L.capaz_last_error(None).decode()


def test_null_context_calls_fail_cleanly():
    L = capaz.load()
    assert L.capaz_S(None) == -1
    assert L.capaz_step(None, 1, 1) == -1
    assert L.capaz_destroy(None) == 0
    assert L.capaz_cells(None) == 0


def test_noalloc_sentinel_validates_before_any_device_call():
    L = capaz.load()
    first = ctypes.c_int64(7)
    ptrs = (ctypes.c_void_p * 9)(*([16] * 9))
    assert L.capaz_check_finite_noalloc(ptrs, 9, 10, None, ctypes.byref(first)) == -1
    assert "outside 1..8" in L.capaz_last_error(None).decode()
    assert L.capaz_check_finite_noalloc(ptrs, 0, 10, None, ctypes.byref(first)) == -1
    assert L.capaz_check_finite_noalloc(None, 1, 10, None, ctypes.byref(first)) == -1
    assert L.capaz_check_finite_noalloc(ptrs, 1, -1, None, ctypes.byref(first)) == -1
    # nothing to scan: finite, without touching a device
    assert L.capaz_check_finite_noalloc(ptrs, 1, 0, None, ctypes.byref(first)) == 0 and first.value == -1
