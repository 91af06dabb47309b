"""The two-phase energy model (the non-isothermal workload's physics: the
configurations' energy closure) on device 0 through the C ABI, against the
oracle (oracle/capaz_oracle.c, pinned by tests/test_oracle_config.py and, bit
for bit, by the host build of the kernels' bodies in tests/test_emul.py).

mu(T) calls exp, the device library's on the GPU and glibc's in the oracle,
so the bar is the workload's tolerance: every field within 1e-12 after one
step and within 1e-9 after the run (element-wise for P_w, S_n and T;
relative to the field's magnitude for the fields that cross zero).  The
fused kernel and the entry-point kernels run the same device arithmetic, so
between them the bar is bit for bit.
"""
import sys

import numpy as np
import pytest

from capaz_amd import capaz, workloads
from capaz_cases import bit_equal, case, max_rel, norm_rel, random_state
from oracle import oracle

pytestmark = pytest.mark.gpu

STATE = ("P_w", "S_n", "T")

def check(got: dict, want: dict, tol: float, what: str) -> None:
    for f, v in want.items():
        r = max_rel(got[f], v) if f in STATE else norm_rel(got[f], v)
        assert r <= tol, (what, f, r)


GRIDS = [(11, 10, 9), (70, 13, 9), (12, 10, 1), (3, 5, 4)]


@pytest.mark.parametrize("dims", GRIDS, ids=lambda d: "x".join(map(str, d)))
@pytest.mark.parametrize("kernel", [1, 0], ids=["fused", "entry_points"])
def test_steps_match_the_oracle(dims, kernel):
    keys = case("2phe", *dims, P_BOTTOM=1.0e5)
    st = random_state(keys, 21)
    d = capaz.Domain(keys, st)
    d.tune(energy_kernel=kernel)
    od = oracle.OracleDomain(keys, st)
    for first, count, tol in ((1, 1, 1e-12), (2, 9, 1e-9)):
        d.run(count, first_step=first)
        od.run(count, first_step=first)
        check(d.fields(), od.fields(), tol, f"after {first + count - 1} steps")
    d.close()


@pytest.mark.parametrize("dims", GRIDS[:3], ids=lambda d: "x".join(map(str, d)))
def test_fused_equals_the_entry_points_bit_for_bit(dims):
    keys = case("2phe", *dims, P_BOTTOM=1.0e5, T_BOTTOM=280.0)
    st = random_state(keys, 22)
    a, b = capaz.Domain(keys, st), capaz.Domain(keys, st)
    a.tune(energy_kernel=1)
    b.tune(energy_kernel=0)
    for first, count in ((1, 1), (2, 4)):
        a.run(count, first_step=first)
        b.run(count, first_step=first)
        for f in a.names:
            assert bit_equal(a.field(f), b.field(f)), (first + count - 1, f)
    a.close()
    b.close()


def test_reference_closure_and_hydrostatic_bottom():
    """The Brooks-Corey closure of the two-phase deck (P_n = P_w + P_k) and the
    reference's hydrostatic bottom rule in the energy model."""
    keys = case("2phe", 14, 12, 10, CLOSURE="reference", LAMBDA_0=2.7, P_D_0=755.0, S_WR_0=0.09)
    keys.pop("P_BOTTOM", None)
    st = random_state(keys, 23)
    d, od = capaz.Domain(keys, st), oracle.OracleDomain(keys, st)
    d.run(1)
    od.run(1)
    check(d.fields(), od.fields(), 1e-12, "1 step")
    d.run(5, first_step=2)
    od.run(5, first_step=2)
    check(d.fields(), od.fields(), 1e-9, "6 steps")
    d.close()


@pytest.mark.parametrize("dims", GRIDS, ids=lambda d: "x".join(map(str, d)))
@pytest.mark.parametrize("kernel", [1, 0], ids=["fused", "entry_points"])
def test_steps_bit_for_bit_without_transcendentals(dims, kernel):
    """With mu(T) constant (MU_T_A 0: exp(-0) = 1 exactly in every library)
    the step calls no transcendental function, so the device and the oracle
    must agree bit for bit, every field, over several steps."""
    keys = case("2phe", *dims, P_BOTTOM=1.0e5, MU_T_A=0.0)
    st = random_state(keys, 25)
    d = capaz.Domain(keys, st)
    d.tune(energy_kernel=kernel)
    od = oracle.OracleDomain(keys, st)
    d.run(6)
    od.run(6)
    for f in od.names:
        assert bit_equal(d.field(f), od.field(f)), (f, max_rel(d.field(f), od.field(f)))
    d.close()


def test_pivot_search_cells_equal_the_oracle():
    """Incompressible fluids (beta 0) make the Newton's pressure pivot 0 in
    every cell, so the fused kernel takes its out-of-line general routine
    (the signed pivot search of gauss.cpp:39-58, which finds no row and
    leaves the matrix as it is) and the entry points their second pass.  The
    un-inverted matrix then scales the residual by ~1e6, so a last-bit
    difference of exp would dominate: with mu(T) constant both paths must
    give the oracle's bits.  Some cells overflow (T ~ 1e300): the kernels'
    exact division rewrite q = a*r, e = fma(-q, b, a), fma(e, r, q) is the
    IEEE quotient for finite operands, and for an infinite dividend it gives
    NaN where the division gives an infinity, so there the bar is the same
    non-finite cells and the oracle's bits everywhere else."""
    keys = case("2phe", 12, 11, 7, BETA_W=0.0, BETA_N=0.0, P_BOTTOM=1.0e5, MU_T_A=0.0)
    st = random_state(keys, 24)
    a, b = capaz.Domain(keys, st), capaz.Domain(keys, st)
    a.tune(energy_kernel=1)
    b.tune(energy_kernel=0)
    a.run(1)
    b.run(1)
    od = oracle.OracleDomain(keys, st)
    od.run(1)
    for f in od.names:
        w = od.field(f)
        fin = np.isfinite(w)
        for d in (a, b):
            g = d.field(f)
            assert np.array_equal(np.isfinite(g), fin), f
            assert bit_equal(g[fin], w[fin]), f
    assert not np.isfinite(od.field("T")).all()  # the case does reach the overflow
    a.close()
    b.close()


def test_whole_run_40_within_1e9():
    """The non-isothermal physics at the Buckley-Leverett workload's 40^3
    size over 1000 steps: every field within 1e-9 of the oracle."""
    keys, st = workloads.nonisothermal_2ph(40, 40, 40)
    d, od = capaz.Domain(keys, st), oracle.OracleDomain(keys, st)
    d.run(1000)
    od.run(1000)
    check(d.fields(), od.fields(), 1e-9, "1000 steps")
    d.close()


def test_workload_full_size_20_steps_within_bars():
    """The non-isothermal workload at its full size (256 x 256 x 128): within
    1e-12 of the oracle after the first fused step and 1e-9 after 20, and the
    primary fields after 3 more steps equal to the entry-point kernels'."""
    keys, st = workloads.nonisothermal_2ph()
    od = oracle.OracleDomain(keys, st)
    d = capaz.Domain(keys, st)
    for first, count, tol in ((1, 1, 1e-12), (2, 19, 1e-9)):
        od.run(count, first_step=first)
        d.run(count, first_step=first)
        got = {f: d.field(f) for f in od.names}
        check(got, od.fields(), tol, f"{first + count - 1} steps")
        print(f"non-isothermal 256x256x128: {first + count - 1} steps within {tol:g}", file=sys.stderr, flush=True)
    del od
    e = capaz.Domain(keys, {f: d.field(f) for f in ("P_w", "S_n", "T", "m", "K")})
    e.tune(energy_kernel=0)
    d.run(3, first_step=21)
    e.run(3, first_step=21)
    for f in STATE:
        assert bit_equal(d.field(f), e.field(f)), f
    d.close()
    e.close()
