"""Pins the oracle (oracle/capaz_oracle.c) to the reference.

* against the compiled reference CPU path (oracle/_ref, built from the
  reference's own cpu.cpp, gauss.cpp, model .cpp and energy.cpp) bit for bit,
  entry point by entry point, for every model and for the reference's own
  input decks (skipped where oracle/_ref is not built);
* gauss.cpp's reverse_matrix against the known-answer inverses of
  Three-phase/Test/test_main.cpp:48-67 and against a long-double solve;
* closures against closed forms (the Corey closure has no reference
  counterpart).
"""
import numpy as np
import pytest

from capaz_cases import bit_equal, case, max_rel, random_state, ref_keys
from oracle import oracle, ref

ENTRIES = ["S_calculation", "ro_P_Xi_calculation", "u_calculation", "roS_calculation",
           "H_E_current_calculation", "E_calculation", "P_S_calculation", "boundary_conditions"]

CASES = [
    ("2ph", 15, 15, 1, 1, {}),                 # Two-phase/defines.ini grid
    ("2ph", 10, 12, 9, 2, {"SOURCE": 2}),
    ("2ph", 3, 5, 4, 3, {"SOURCE": 1, "P_D_0": 0.0}),
    ("3ph", 1, 15, 1, 4, {}),                  # Three-phase/defines.ini grid
    ("3ph", 10, 12, 9, 5, {}),
    ("3ph", 10, 12, 9, 6, {"SOURCE": 1, "S_W_GR": 0.45, "S_N_GR": 0.25}),
    ("3ph", 3, 3, 3, 7, {}),
    ("3ph", 7, 3, 1, 8, {"SOURCE": 2}),
    ("3phe", 8, 9, 7, 9, {}),
    ("3phe", 1, 9, 1, 10, {}),
]

needs_ref = pytest.mark.skipif(not ref.available("3ph"), reason="oracle/_ref not built (needs the reference sources)")


@needs_ref
@pytest.mark.parametrize("model,nx,ny,nz,seed,over", CASES)
def test_entry_points_match_reference(model, nx, ny, nz, seed, over):
    keys = case(model, nx, ny, nz, **over)
    st = random_state(keys, seed)
    rd = ref.RefDomain(model, ref_keys(keys))
    for k, v in st.items():
        rd.field(k)[...] = v
    od = oracle.OracleDomain(keys, st)
    for s in range(1, 4):
        for e in ENTRIES:
            if e in ("H_E_current_calculation", "E_calculation") and model != "3phe":
                continue
            rd.call(e, s * keys["DT"])
            od.call(e, s * keys["DT"])
            for f in od.names:
                assert bit_equal(rd.field(f), od.field(f)), (s, e, f, max_rel(rd.field(f), od.field(f)))


@needs_ref
@pytest.mark.parametrize("model", ["2ph", "3ph", "3phe"])
def test_reference_data_initialization_run(model):
    """The reference's own data_initialization and deck, many steps."""
    dims = {"2ph": (15, 15, 1), "3ph": (1, 15, 1), "3phe": (1, 9, 1)}[model]
    keys = case(model, *dims)
    rd = ref.RefDomain(model, ref_keys(keys), data_initialization=True)
    od = oracle.OracleDomain(keys, rd.fields())
    rd.run(200)
    od.run(200)
    for f in od.names:
        assert bit_equal(rd.field(f), od.field(f)), (f, max_rel(rd.field(f), od.field(f)))


# Three-phase/Test/test_main.cpp:53-66 (inputs, expected inverses).
KATS = [
    (np.eye(3), np.eye(3)),
    (np.array([[-4 / 19, 7 / 38, 5 / 38], [3 / 19, 9 / 38, 1 / 38], [2 / 19, -13 / 38, 7 / 38]]),
     np.array([[-2, 3, 1], [1, 2, -1], [3, 2, 3]], dtype=float)),
    (np.array([[3, 0, 0], [2, 2, 1], [1, 0, 1]], dtype=float),
     np.array([[1 / 3, 0, 0], [-1 / 6, 0.5, -0.5], [-1 / 3, 0, 1]])),
]


@pytest.mark.parametrize("a,inv", KATS)
def test_reverse_matrix_known_answers(a, inv):
    rc, out = oracle.reverse_matrix(a)
    assert rc == 0
    assert np.all(np.abs(out - inv) < 1e-10)  # the test's own tolerance (test_main.cpp:36)


def exact_inverse(a: np.ndarray) -> np.ndarray:
    """Inverse of the double matrix `a` in exact rational arithmetic."""
    from fractions import Fraction
    n = a.shape[0]
    m = [[Fraction(float(a[i, j])) for j in range(n)] + [Fraction(int(i == j)) for j in range(n)] for i in range(n)]
    for k in range(n):
        piv = max(range(k, n), key=lambda r: abs(m[r][k]))
        m[k], m[piv] = m[piv], m[k]
        f = m[k][k]
        m[k] = [x / f for x in m[k]]
        for i in range(n):
            if i != k and m[i][k] != 0:
                f1 = m[i][k]
                m[i] = [x - y * f1 for x, y in zip(m[i], m[k])]
    return np.array([[float(m[i][n + j]) for j in range(n)] for i in range(n)])


def test_reverse_matrix_random_against_exact_inverse():
    rng = np.random.default_rng(11)
    for n in (3, 5):
        for _ in range(100):
            a = rng.standard_normal((n, n)) * 10.0 ** rng.uniform(-3, 3, (n, 1))
            a += np.eye(n) * n * np.abs(a).max(axis=1, keepdims=True)
            rc, out = oracle.reverse_matrix(a)
            assert rc == 0
            exact = exact_inverse(a)
            assert np.max(np.abs(out - exact) / np.abs(exact).max()) < 1e-12


def test_reverse_matrix_pivot_rules():
    # |a00| < 1e-8 with a later positive entry: rows swap and the inverse is right.
    a = np.array([[1e-9, 1.0, 2.0], [3.0, 1.0, 0.5], [0.2, 4.0, 1.0]])
    rc, out = oracle.reverse_matrix(a)
    assert rc == 0
    np.testing.assert_allclose(out @ a, np.eye(3), atol=1e-12)
    # no later entry > 1e-8 (signed test): -1 and the matrix is left as it was.
    b = np.array([[0.0, 1.0, 2.0], [-3.0, 1.0, 0.5], [-0.2, 4.0, 1.0]])
    rc, out = oracle.reverse_matrix(b)
    assert rc == -1
    assert np.array_equal(out, b)


@needs_ref
def test_reverse_matrix_matches_reference_gauss():
    rng = np.random.default_rng(12)
    for n in (3, 5):
        for _ in range(300):
            a = rng.standard_normal((n, n))
            if rng.uniform() < 0.3:
                a[0, 0] = rng.uniform(-1e-8, 1e-8)  # exercise the pivot search
            rc_r, out_r = ref.reverse_matrix(a)
            rc_o, out_o = oracle.reverse_matrix(a)
            assert rc_r == rc_o
            assert bit_equal(out_r, out_o)


def test_corey_closure_closed_form():
    keys = dict(MODEL=3, CLOSURE="corey", COREY_N=2, NX=3, NY=3, NZ=3, S_WR_0=0.1, S_NR_0=0.05, S_GR_0=0.02)
    den = 1 - 0.1 - 0.05 - 0.02
    for Sw, Sn in [(0.2, 0.7), (0.5, 0.3), (0.05, 0.9), (0.9, 0.02), (0.12, 0.06)]:
        swe, sne = (Sw - 0.1) / den, (Sn - 0.05) / den
        sge = 1 - swe - sne
        want = [min(max(s, 0), 1) ** 2 for s in (swe, sne, sge)]
        np.testing.assert_allclose(oracle.relperm(keys, Sw, Sn), want, rtol=1e-15, atol=0)
    keys2 = dict(MODEL=2, CLOSURE="corey", COREY_N=3, NX=3, NY=3, NZ=3, S_WR_0=0.1, S_NR_0=0.05)
    for Sn in (0.1, 0.5, 0.8):
        se = (1 - Sn - 0.1) / 0.85
        np.testing.assert_allclose(oracle.relperm(keys2, 1 - Sn, Sn)[:2], [se ** 3, (1 - se) ** 3], rtol=1e-14)


def test_reference_closures_closed_form():
    # Brooks-Corey (two-phase.cpp:6-12) and the van Genuchten-type k_r
    # (three-phase.cpp:146-186) evaluated independently in long double.
    k2 = dict(MODEL=2, NX=3, NY=3, NZ=3, S_WR_0=0.09, LAMBDA_0=2.7)
    for Sn in (0.1, 0.4, 0.7):
        se = (1 - Sn - 0.09) / (1 - 0.09)
        lam = 2.7
        want = [se ** ((2 + 3 * lam) / lam), (1 - se) ** 2 * (1 - se ** ((2 + lam) / lam))]
        np.testing.assert_allclose(oracle.relperm(k2, 1 - Sn, Sn)[:2], want, rtol=1e-13)
    k3 = dict(MODEL=3, NX=3, NY=3, NZ=3, LAMBDA_0=3.25)
    A = 3.25
    for Sw, Sn in [(0.4, 0.3), (0.2, 0.5), (0.6, 0.0005)]:
        sge = 1 - Sw - Sn
        kw = Sw ** 0.5 * (1 - (1 - Sw ** (A / (A - 1))) ** ((A - 1) / A)) ** 2
        kg = sge ** 0.5 * (1 - (1 - sge) ** (A / (A - 1))) ** (2 * (A - 1) / A)
        if Sn >= 1e-3:
            knw = (1 - Sw) ** 0.5 * (1 - Sw ** (A / (A - 1))) ** (2 * (A - 1) / A)
            kng = Sn ** 0.5 * (1 - (1 - Sn ** (A / (A - 1))) ** ((A - 1) / A)) ** 2
            kn = Sn * knw * kng / (1 - Sw) / (1 - sge)
        else:
            kn = 0.0
        np.testing.assert_allclose(oracle.relperm(k3, Sw, Sn), [kw, kn, kg], rtol=1e-12)


def test_uniform_state_without_gravity_is_stationary():
    """g = 0 and a uniform state: every flux is exactly zero, so one step
    returns the state bit for bit (Newton starts at the root)."""
    keys = case("3ph", 6, 7, 5, G_CONST=0.0)
    shape = (5, 7, 6)
    st = dict(P_w=np.full(shape, 1.0e5), P_n=np.full(shape, 1.0e5), P_g=np.full(shape, 1.0e5),
              S_w=np.full(shape, 0.4), S_n=np.full(shape, 0.3), m=np.full(shape, 0.4),
              K=np.full(shape, keys["K_0"]))
    od = oracle.OracleDomain(keys, st)
    od.run(3)
    for f in ("P_w", "S_w", "S_n"):
        assert np.array_equal(od.field(f), st[f])


def mass_change(d, keys):
    """Applies boundary_conditions (with g = 0 every boundary cell takes its
    inward neighbour's pressure, so no flux crosses the domain faces), then
    the flux update, and returns (sum of m*roS after - before, sum of |change|)
    over the interior cells."""
    for e in ("boundary_conditions", "S_calculation", "ro_P_Xi_calculation", "u_calculation"):
        d.call(e)
    m = d.field("m")[1:-1, 1:-1, 1:-1].astype(np.longdouble)
    d.call("roS_calculation")
    out = []
    for ph in ("w", "n", "g"):
        new = d.field("roS_" + ph)[1:-1, 1:-1, 1:-1].astype(np.longdouble)
        old = d.field("roS_" + ph + "_old")[1:-1, 1:-1, 1:-1].astype(np.longdouble)
        # rounding of the stored roS_new is the only loss: half an ulp per cell
        noise = 4 * new.size * np.abs(m * new).max() * 2.0 ** -53
        out.append(((m * new).sum() - (m * old).sum(), np.abs(m * (new - old)).sum(), noise))
    return out


def test_mass_conserved_without_gravity():
    """With g = 0 the upwind face flux of cell c toward c+1 is the negative of
    c+1's flux toward c (cpu.cpp:236-264), so the divergence telescopes and
    sum(m * roS) over the interior is conserved to rounding, although every
    cell changes."""
    keys = case("3ph", 8, 9, 7, G_CONST=0.0, DT=1e-4)
    od = oracle.OracleDomain(keys, random_state(keys, 3))
    for total, moved, noise in mass_change(od, keys):
        assert moved > 1e6 * noise
        assert abs(total) <= noise
