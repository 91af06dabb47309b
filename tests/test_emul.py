"""The kernels' dense solves, built for the host, against the oracle's
gauss.cpp restatement bit for bit (the per-entry-point and fused kernel
bodies are covered by tests/test_golden.py)."""
import numpy as np
import pytest

from capaz_cases import bit_equal
from emul import emul
from oracle import oracle


def random_matrices(n, count, seed):
    rng = np.random.default_rng(seed)
    for t in range(count):
        a = rng.standard_normal((n, n)) * 10.0 ** rng.uniform(-6, 6, (n, 1))
        kind = t % 5
        if kind == 1:
            a[0, 0] = rng.uniform(-1e-8, 1e-8)          # first pivot needs the search
        elif kind == 2:
            a[1] = a[0] * rng.uniform(0.5, 2) + rng.uniform(-1e-9, 1e-9, n)  # later pivot tiny
        elif kind == 3:
            a[:, 0] = -np.abs(a[:, 0]) * 1e-9            # no admissible pivot (signed search)
        elif kind == 4 and n == 3:
            a[0, 2] = 0.0                                 # the three-phase Jacobian's zero (dF[2])
        yield a


@pytest.mark.parametrize("n", [3, 4, 5])
def test_reverse_matrix_matches_oracle(n):
    for a in random_matrices(n, 3000, 100 + n):
        rc_o, out_o = oracle.reverse_matrix(a)
        rc_e, out_e = emul.reverse_matrix(a)
        assert rc_o == rc_e
        assert bit_equal(out_o, out_e)


def test_reverse3_fast_path_matches_oracle():
    hits = 0
    for a in random_matrices(3, 20000, 7):
        rc_o, out_o = oracle.reverse_matrix(a)
        rc_e, out_e = emul.reverse_matrix(a, fast=True)
        assert rc_o == rc_e
        assert bit_equal(out_o, out_e), a
        hits += rc_o == 0
    assert hits > 10000


def test_reverse3_a02zero_path_matches_oracle():
    rng = np.random.default_rng(8)
    hits = 0
    for a in random_matrices(3, 20000, 9):
        a = a.copy()
        a[0, 2] = 0.0
        if rng.uniform() < 0.1:
            a[1, 2] = 0.0  # more structural zeros
        rc_o, out_o = oracle.reverse_matrix(a)
        rc_e, out_e = emul.reverse_matrix(a, a02zero=True)
        assert rc_o == rc_e
        assert bit_equal(out_o, out_e), a
        hits += rc_o == 0
    assert hits > 10000


@pytest.mark.parametrize("n", [4, 5])
def test_reverse_nopivot_matches_oracle_or_refuses(n):
    """The energy Newtons' 5x5 and 4x4 inverses without the pivot search: the
    oracle's bits whenever every pivot has |f| >= 1e-8, otherwise they refuse
    and leave the matrix as it was (the kernel then redoes the cell with the
    search)."""
    hits = refused = 0
    for a in random_matrices(n, 5000, 11 + n):
        rc_o, out_o = oracle.reverse_matrix(a)
        rc_e, out_e = emul.reverse_matrix(a, fast=True)
        if rc_e == 1:
            refused += 1
            assert bit_equal(out_e, a)
        else:
            assert rc_o == 0 and bit_equal(out_o, out_e), a
            hits += 1
    assert hits > 2000 and refused > 1000


@pytest.mark.parametrize("closure,gas_free", [("reference", False), ("corey", False), ("reference", True)])
def test_energy_newton_two_passes_equal_the_general_routine(closure, gas_free):
    """The energy model's Newton as the kernels run it (every interior cell
    without the pivot search, then the refused cells again with it) leaves
    every field with the bits of the general routine.  gas_free: water and
    oil incompressible and a few cells with S_w + S_n = 1, whose pressure
    pivot is then 0, so the second pass runs."""
    from capaz_amd import workloads
    keys, st = workloads.energy_3ph(12, 10, 8, closure=closure)
    st = {f: np.array(v, copy=True) for f, v in st.items()}
    if gas_free:
        keys = dict(keys, BETA_W=0.0, BETA_N=0.0)
        for c in ((3, 4, 5), (4, 5, 6), (5, 2, 8)):
            st["S_w"][c], st["S_n"][c] = 0.625, 0.375
    a, b = emul.EmulDomain(keys, st), emul.EmulDomain(keys, st)
    redone = 0
    for _ in range(6):
        for d in (a, b):
            for e in ("S_calculation", "ro_P_Xi_calculation", "u_calculation", "roS_calculation",
                      "H_E_current_calculation", "E_calculation"):
                d.call(e)
        a.call("P_S_calculation")
        b.call("newton_energy_two_pass")
        redone += b.redone
        for d in (a, b):
            d.call("boundary_conditions")
    assert redone >= 3 if gas_free else redone == 0
    for f, v in a.fields().items():
        assert bit_equal(b.field(f), v), f


def test_pivot_search_state_host_build_matches_oracle():
    """The kernels' bodies built for the host on the state whose Newton
    matrices all need the signed pivot search (interior S_w in
    [0.005, 0.02] at beta_w 4.5e-10): three entry-point steps and three
    fused steps equal the oracle bit for bit, every primary field."""
    from test_gpu_parity import pivot_search_state

    keys, st = pivot_search_state(12)
    od = oracle.OracleDomain(keys, st)
    od.run(3)
    seq = emul.EmulDomain(keys, st)
    seq.run(3)
    fused = emul.EmulDomain(keys, st)
    for _ in range(3):
        fused.fused_step()
    for f in ("P_w", "S_w", "S_n"):
        assert bit_equal(seq.field(f), od.field(f)), f
        assert bit_equal(fused.field(f), od.field(f)), f
    for f in od.names:
        assert bit_equal(seq.field(f), od.field(f)), f


@pytest.mark.parametrize("closure", ["corey", "reference"])
def test_two_phase_energy_host_build_matches_oracle(closure):
    """The two-phase energy model (the configurations' energy closure): the
    kernels' bodies built for the host equal the oracle bit for bit, every
    field after every entry point over three steps (both use glibc's exp and
    pow here), and the Newton's two passes equal its general routine."""
    from capaz_amd import workloads
    keys, st = workloads.nonisothermal_2ph(11, 10, 9)
    keys["P_BOTTOM"] = 1.0e5
    if closure == "reference":  # Brooks-Corey k_r and P_c of the two-phase deck
        keys.update(CLOSURE="reference", LAMBDA_0=2.7, P_D_0=755.0, S_WR_0=0.09)
    rng = np.random.default_rng(12)
    st = {f: np.array(v, copy=True) for f, v in st.items()}
    st["T"] += 20 * rng.random(st["T"].shape)
    st["S_n"] = rng.uniform(0.2, 0.8, st["S_n"].shape)
    od, ed, e2 = oracle.OracleDomain(keys, st), emul.EmulDomain(keys, st), emul.EmulDomain(keys, st)
    entries = ("S_calculation", "ro_P_Xi_calculation", "u_calculation", "roS_calculation",
               "H_E_current_calculation", "E_calculation", "P_S_calculation", "boundary_conditions")
    for s in range(1, 4):
        for e in entries:
            od.call(e, s * keys["DT"])
            ed.call(e, s * keys["DT"])
            e2.call("newton_energy_two_pass" if e == "P_S_calculation" else e, s * keys["DT"])
            for f in od.names:
                assert bit_equal(ed.field(f), od.field(f)), (s, e, f)
                assert bit_equal(e2.field(f), od.field(f)), (s, e, f)
