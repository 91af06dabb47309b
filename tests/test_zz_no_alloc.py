"""capaz_check_finite_noalloc, the sentinel entry point that allocates
nothing, against capaz_check_finite (the same scan with the context's own
device buffer), through the C ABI on device 0: two calls in a row, the second
on a smaller input, give each the first bad index of the context entry point;
a call after a failing one starts from clear slots; a prefix that ends before
the bad cell is finite; more fields than the static scratch holds are
rejected."""
import ctypes

import numpy as np
import pytest

from capaz_amd import capaz
from capaz_cases import case, random_state

pytestmark = pytest.mark.gpu

CAPAZ_ERR_INVALID, CAPAZ_ERR_NONFINITE = -1, -5
STATE = {"2ph": ("P_w", "S_n"), "3ph": ("P_w", "S_n", "S_w"), "3phe": ("P_w", "S_n", "S_w", "T")}


def noalloc(d, names, n=None):
    """capaz_check_finite_noalloc over the domain's device arrays `names`."""
    L = capaz.load()
    ptrs = (ctypes.c_void_p * len(names))(*[d.device_ptr(f) for f in names])
    first = ctypes.c_int64(-2)
    rc = L.capaz_check_finite_noalloc(ptrs, len(names), d.n if n is None else n, ctypes.c_void_p(d.stream()),
                                      ctypes.byref(first))
    return rc, first.value, L.capaz_last_error(None).decode()


def with_context(d):
    """capaz_check_finite of the domain: (status, first bad cell or -1)."""
    L = capaz.load()
    first = ctypes.c_int64(-2)
    rc = L.capaz_check_finite(d._ctx, ctypes.byref(first))
    return rc, first.value


def domain(model, dims, seed, bad):
    over = dict(CLOSURE="corey", S_WR_0=0.1, S_NR_0=0.05) if model != "3phe" else {}
    keys = case(model, *dims, DT=1e-4, **over)
    st = random_state(keys, seed)
    if model == "3phe":
        st["T"] = np.full(st["P_w"].shape, 290.0)
    for f, cell, v in bad:
        st[f] = st[f].copy()
        st[f].flat[cell] = v
    return capaz.Domain(keys, st)


def test_two_calls_in_a_row_match_the_context_entry_point():
    # first call: three phases + energy, NaN and infinity in different fields
    d1 = domain("3phe", (14, 13, 12), 5, [("S_w", 1500, np.nan), ("T", 901, np.inf), ("P_w", 1777, np.nan)])
    rc, first, msg = noalloc(d1, STATE["3phe"])
    want = with_context(d1)
    assert (rc, first) == want == (CAPAZ_ERR_NONFINITE, 901)
    assert "index 901" in msg and "field(s) 3" in msg
    # second call, smaller input: two phases, one NaN at a smaller index than
    # the first call found (the slots must not keep the first call's minimum)
    d2 = domain("2ph", (8, 7, 6), 6, [("S_n", 300, np.nan)])
    rc, first, msg = noalloc(d2, STATE["2ph"])
    assert (rc, first) == with_context(d2) == (CAPAZ_ERR_NONFINITE, 300)
    assert "field(s) 1" in msg
    d1.close()
    d2.close()


def test_finite_state_and_prefix_after_a_failing_call():
    d = domain("3ph", (16, 15, 14), 7, [("P_w", 2500, np.nan)])
    assert noalloc(d, STATE["3ph"])[:2] == with_context(d) == (CAPAZ_ERR_NONFINITE, 2500)
    # a prefix ending before the bad cell is finite, and so is a finite domain
    assert noalloc(d, STATE["3ph"], n=2500)[:2] == (0, -1)
    assert noalloc(d, STATE["3ph"], n=2501)[:2] == (CAPAZ_ERR_NONFINITE, 2500)
    e = domain("2ph", (9, 8, 7), 8, [])
    assert noalloc(e, STATE["2ph"])[:2] == with_context(e) == (0, -1)
    # eight fields (the maximum) over one repeated finite array
    assert noalloc(e, ("P_w",) * 8)[:2] == (0, -1)
    d.close()
    e.close()


def test_rejects_more_fields_than_the_scratch_holds():
    e = domain("2ph", (9, 8, 7), 9, [])
    rc, first, msg = noalloc(e, ("P_w",) * 9)
    assert rc == CAPAZ_ERR_INVALID and "outside 1..8" in msg
    e.close()
