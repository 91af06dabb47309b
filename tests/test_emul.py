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


@pytest.mark.parametrize("n", [3, 5])
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
