"""Failure detection: the non-finite sentinel (capaz_check_finite and the
check_every key of capaz_tune).

The reference only prints NaNs, under MY_TEST (test_nan, shared_test.cpp:56-70).
The oracle fixes how one NaN spreads: a NaN in one interior cell's P_w reaches
exactly that cell and its six neighbours in one step (each neighbour's upwind
flux reads the cell's coefficients; the boundary cells stay finite when the
cell is three or more cells from every face).  With the energy equation it
also reaches the cells two apart along x and y (their advection term reads u
at the neighbour, whose central difference reads the cell) and the four
cells (k +- 1, j +- 1), whose z advection term reads u_y at (k, j +- 1)
(energy.cpp:300-305) -- but none two apart along z.  On the GPU the sentinel must
name the first such cell, stop capaz_step at the checked step, and leave the
state the oracle has after it.
"""
import numpy as np
import pytest

from capaz_amd import capaz
from capaz_cases import bit_equal, case, norm_rel, random_state
from oracle import oracle

GRID = (12, 11, 10)
CELL = (5, 4, 6)  # (k, j, i): three or more cells from every face


def nan_case(model: str):
    over = dict(CLOSURE="corey", S_WR_0=0.1, S_NR_0=0.05) if model != "3phe" else {}
    keys = case(model, GRID[0], GRID[1], GRID[2], DT=1e-4, SOURCE=1, **over)
    st = random_state(keys, 31)
    st["P_w"] = st["P_w"].copy()
    st["P_w"][CELL] = np.nan
    return keys, st


def lin(keys, k, j, i):
    return i + j * keys["NX"] + k * keys["NX"] * keys["NY"]


STATE = {"2ph": ("P_w", "S_n"), "3ph": ("P_w", "S_w", "S_n"), "3phe": ("P_w", "S_w", "S_n", "T")}


def bad_cells(od, model):
    bad = np.zeros(od.shape, dtype=bool)
    for f in STATE[model]:
        bad |= ~np.isfinite(od.field(f))
    return bad


@pytest.mark.parametrize("model", ["2ph", "3ph", "3phe"])
def test_one_nan_spreads_to_the_seven_point_stencil_on_the_oracle(model):
    keys, st = nan_case(model)
    od = oracle.OracleDomain(keys, st)
    od.run(1)
    bad = bad_cells(od, model)
    k, j, i = CELL
    want = np.zeros_like(bad)
    spread = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    if model == "3phe":
        spread += [(0, 2, 0), (0, -2, 0), (0, 0, 2), (0, 0, -2), (1, 1, 0), (1, -1, 0), (-1, 1, 0), (-1, -1, 0)]
    for dk, dj, di in spread:
        want[k + dk, j + dj, i + di] = True
    assert np.array_equal(bad, want), np.argwhere(bad != want)


@pytest.mark.gpu
@pytest.mark.parametrize("model", ["2ph", "3ph", "3phe"])
def test_sentinel_names_the_first_bad_cell_and_stops_the_step(model):
    keys, st = nan_case(model)
    k, j, i = CELL
    d = capaz.Domain(keys, st)
    # before stepping: the injected cell itself
    with pytest.raises(capaz.NonFiniteState) as e:
        d.check_finite()
    assert e.value.first_bad == lin(keys, k, j, i)
    assert "P_w" in str(e.value)
    # checked every second step: step 1 passes unchecked, step 2 stops the run
    d.tune(check_every=2)
    with pytest.raises(capaz.NonFiniteState) as e:
        d.run(5, first_step=1)
    od = oracle.OracleDomain(keys, st)
    od.run(2)
    bad = bad_cells(od, model)
    assert e.value.first_bad == int(np.flatnonzero(bad.ravel())[0])
    assert "after step 2" in str(e.value)
    # the context holds the state after the failing step: the oracle's (bit
    # for bit with the Corey closures; the energy model's exp/pow are the
    # device library's: the same non-finite cells, the rest within 1e-12)
    for f in od.names:
        g, w = d.field(f), od.field(f)
        if model != "3phe":
            assert bit_equal(g, w), f
        else:
            fin = np.isfinite(w)
            assert np.array_equal(fin, np.isfinite(g)), f
            assert norm_rel(g[fin], w[fin]) <= 1e-12, f
    d.close()


@pytest.mark.gpu
@pytest.mark.parametrize("model", ["2ph", "3ph"])
def test_sentinel_on_a_finite_state_changes_nothing(model):
    keys, st = nan_case(model)
    st["P_w"][CELL] = keys["P_ATM"]
    a, b = capaz.Domain(keys, st), capaz.Domain(keys, st)
    a.tune(check_every=1)
    assert a.check_finite() == -1
    a.run(6)
    b.run(6)
    for f in a.names:
        assert bit_equal(a.field(f), b.field(f)), f
    a.close()
    b.close()
