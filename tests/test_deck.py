"""Input decks, initial state, time loop and Tecplot plots of the reference
driver (capaz_amd.deck).

  * the deck reader follows read_defines' rules (main.cpp:1280-1639): '#'
    lines, key up to '=', value up to the first space, atof/atoi prefixes,
    model-specific keys;
  * run on the oracle, the deck runner writes plot files byte-identical to
    the ones the reference CPU program (built from the reference's own
    main.cpp) writes for the same deck -- committed as fixtures
    tests/golden/plots_<deck>.npz by tests/golden/make_golden.py: reader,
    initial state, loop, file names, headers, row order and number formats
    all agree;
  * on the GPU the same decks give the reference program's plots to the
    printed precision, and the fused mode writes the same files as the
    entry-point sequence.
"""
import math
import os

import numpy as np
import pytest

from capaz_amd import deck, io
from capaz_cases import PLOT_DECKS, REF_2PH, REF_3PH, case, plot_deck_id, ref_keys, write_deck
from golden.make_golden import load_plots

DECKS = PLOT_DECKS
IDS = [plot_deck_id(m, d) for m, d in DECKS]
MODEL = {"2ph": (2, 0), "3ph": (3, 0), "3phe": (3, 1)}


def reference_plots(model, keys):
    """The reference CPU program's plot files and printed text for a deck."""
    return load_plots(model, keys)


def read_plots(workdir):
    d = os.path.join(workdir, "plots")
    return {n: open(os.path.join(d, n)).read() for n in sorted(os.listdir(d))}


def silent(*a, **k):
    pass


# ---- reader ------------------------------------------------------------------


def test_atof_atoi_read_numeric_prefixes():
    assert deck.atof("1e-3\n") == 1e-3
    assert deck.atof("  -2.5e+2abc") == -250.0
    assert deck.atof(".5") == 0.5 and deck.atof("5.") == 5.0
    assert deck.atof("abc") == 0.0 and deck.atof("") == 0.0
    assert deck.atof("1e") == 1.0          # strtod stops before a bare exponent
    assert deck.atof("0x1p-2") == 0.25 and deck.atof("0x10") == 16.0
    assert deck.atof("inf") == math.inf and math.isnan(deck.atof("nan"))
    assert deck.atoi("12\r\n") == 12 and deck.atoi(" -3x") == -3 and deck.atoi("x3") == 0
    assert deck.atoi("7.9") == 7


def test_parse_follows_the_reference_rules():
    text = ("# comment line with KEY=1\n"
            "NX=12      trailing words\n"
            "NY=7.9\n"                 # atoi
            "DT=1e-4\r\n"              # Windows line end
            " HX=3\n"                  # the name keeps its leading space: unknown
            "HY =4\n"                  # and its trailing one
            "HZ=abc\n"                 # atof of no number
            "MU_W=1\tMU_N=2\n"         # a tab does not end the value
            "DT=2e-4\n"                # the last assignment wins
            "no equals sign here\n"
            "P_D_0=755\n"              # a two-phase key
            "RO_G=1.4\n")              # a three-phase key
    k3 = deck.parse_defines(text, 3)
    assert k3["NX"] == 12 and k3["NY"] == 7 and k3["DT"] == 2e-4
    assert "HX" not in k3 and "HY" not in k3 and k3["HZ"] == 0.0
    assert k3["MU_W"] == 1.0 and "MU_N" not in k3
    assert k3["RO_G"] == 1.4 and "P_D_0" not in k3
    k2 = deck.parse_defines(text, 2)
    assert k2["P_D_0"] == 755.0 and "RO_G" not in k2
    assert isinstance(k3["NX"], int) and isinstance(k3["DT"], float)


def test_read_defines_and_find(tmp_path):
    keys = case("3ph", 4, 5, 6, TIMEX=1.0, SAVE_PLOTS=3, PRINT_SCREEN=2)
    write_deck(keys, str(tmp_path / "run"))
    path = deck.find_defines(str(tmp_path / "run"))
    got = deck.read_defines(path, 3)
    for k, v in ref_keys(keys).items():
        if k in deck.THREE_PHASE_KEYS or k in deck.COMMON_KEYS:
            assert got[k] == v, k
    assert got["MODEL"] == 3 and "ENERGY" not in got
    # DEFINES_FILE "..//defines.ini" is tried first
    write_deck(dict(keys, NX=9), str(tmp_path))
    assert deck.read_defines(deck.find_defines(str(tmp_path / "run")), 3)["NX"] == 9
    with pytest.raises(ValueError):
        deck.read_defines(path, 2, energy=1)
    with pytest.raises(FileNotFoundError):
        deck.find_defines(str(tmp_path / "nowhere"))


@pytest.mark.parametrize("model,want", [(2, REF_2PH), (3, REF_3PH)])
def test_reference_decks_read_as_their_values(model, want):
    # the reference's own decks, where its sources are present
    path = os.path.join("/root/reference", "Two-phase" if model == 2 else "Three-phase", "defines.ini")
    if not os.path.exists(path):
        pytest.skip("reference decks absent")
    got = deck.read_defines(path, model)
    for k, v in want.items():
        if k != "MODEL":
            assert got[k] == v, k


# ---- initial state -----------------------------------------------------------


def test_initial_state_two_phase_is_hydrostatic():
    keys = case("2ph", 7, 6, 5, SOURCE=1)
    st = deck.initial_state(keys, "/nonexistent", log=silent)
    S_n, P = st["S_n"], st["P_w"]
    # source square on the top face: |i - 3| <= 1 and |k - 2| <= 1
    top = np.zeros((5, 7))
    top[1:4, 2:5] = keys["S_N_GR"]
    assert np.array_equal(S_n[:, 0, :], top) and not S_n[:, 1:, :].any()
    assert (P[:, 0, :] == keys["P_ATM"]).all()
    # below the top: P rises by the effective density of the cell above
    ro_w = keys["RO_W"] * (1 + keys["BETA_W"] * (P[:, :-1] - keys["P_ATM"]))
    ro_n = keys["RO_N"] * (1 + keys["BETA_N"] * (P[:, :-1] - keys["P_ATM"]))
    dP = (ro_n * S_n[:, :-1] + ro_w * (1 - S_n[:, :-1])) * keys["M_0"] * keys["G_CONST"] * keys["HY"]
    assert np.array_equal(P[:, 1:], P[:, :-1] + dP)
    assert (st["K"] == keys["K_0"]).all() and (st["m"] == keys["M_0"]).all()


def test_initial_state_three_phase_and_noise_dat(tmp_path):
    keys = case("3phe", 4, 3, 2)
    ln = np.arange(12, dtype=float).reshape(3, 4) * 0.1
    io.write_noise_dat(str(tmp_path / "noise.dat"), ln, nz=2)
    st = deck.initial_state(keys, str(tmp_path), log=silent)
    assert (st["S_w"] == keys["S_W_INIT"]).all() and (st["S_g"] == 1 - keys["S_W_INIT"] - keys["S_N_INIT"]).all()
    assert (st["P_w"] == keys["P_ATM"]).all() and (st["P_g"] == keys["P_ATM"]).all() and (st["T"] == 285).all()
    assert np.allclose(st["K"], 1e-10 * np.exp(ln)[None], rtol=0, atol=0)


# ---- plots -------------------------------------------------------------------


def test_plot_layout_matches_print_plot_row():
    keys = case("3ph", 2, 3, 2, HX=0.5, HY=0.25, HZ=0.125)
    shape = (2, 3, 2)
    f = {n: np.zeros(shape) for n in deck.PLOT_FIELDS if n != "T"}
    f["S_w"] = np.arange(12, dtype=float).reshape(shape)
    f["uy_w"] = np.ones(shape)
    text = deck.plot_text(f, keys, 1.5)
    lines = text.splitlines()
    assert lines[0] == 'TITLE =  "Filtration in time= 1.50" '
    assert lines[2] == 'ZONE T = "BIG ZONE", K=2,J=3,I=2, F = POINT'
    # first rows: i = 0, j = 0, k = 0 then k = 1 (k fastest); y = (Ny-1-j) hy
    r0, r1 = lines[3].split(), lines[4].split()
    assert r0[:3] == ["0.00e+00", "0.00e+00", "5.00e-01"] and r1[:3] == ["0.00e+00", "1.25e-01", "5.00e-01"]
    assert r0[3] == "0.00000e+00" and r1[3] == "6.00000e+00"
    assert r0[9] == "-1.00000e+00" and r0[10] == "0.00000e+00"   # uw_y negated; un_x
    assert r0[12] == "-0.00000e+00"                              # (-1) * 0
    assert len(lines) == 3 + 12 and deck.plot_name(1.5) == "S=0000001.5000.dat"


def test_c_row_writer_matches_python_formatting_and_checks_formats(tmp_path):
    import ctypes

    from capaz_amd import capaz

    keys = case("3phe", 3, 4, 5)
    rng = np.random.default_rng(5)
    shape = (5, 4, 3)
    f = {n: rng.standard_normal(shape) * 10.0 ** rng.integers(-8, 8, shape) for n in deck.PLOT_FIELDS}
    f["uy_w"][0, 0, 0] = 0.0
    path = str(tmp_path / "p.dat")
    deck.write_plot(path, f, keys, 0.25)
    assert open(path).read() == deck.plot_text(f, keys, 0.25)
    L = capaz.load()
    one = np.zeros(1)
    for bad in (b"%d", b"%s", b"%.2e %e", b"x", b"%*e"):
        fm = (ctypes.c_char_p * 1)(bad)
        assert L.capaz_write_plot_rows(path.encode(), one.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1, 1,
                                       fm) == -1, bad


@pytest.mark.parametrize("model,keys", DECKS, ids=IDS)
def test_deck_on_the_oracle_writes_the_reference_programs_plots(model, keys, tmp_path):
    from oracle import oracle

    want, out = reference_plots(model, keys)
    write_deck(keys, str(tmp_path / "ours"))
    lines = []
    m, e = MODEL[model]
    r = deck.run_deck(str(tmp_path / "ours"), m, e, backend=oracle.OracleDomain, log=lines.append)
    got = read_plots(str(tmp_path / "ours"))
    assert list(got) == list(want) and len(got) >= 2
    for name in want:
        assert got[name] == want[name], name
    # the progress lines the reference prints
    ref_t = [l for l in out.splitlines() if l.startswith("t=")]
    assert [l for l in lines if l.startswith("t=")] == ref_t
    assert r["steps"] == int(math.floor(keys["TIMEX"] / keys["DT"]))


# ---- GPU ---------------------------------------------------------------------


def numbers(text):
    rows = [l.split() for l in text.splitlines()[3:] if l.strip()]
    return np.array([[float(x) for x in r] for r in rows])


@pytest.mark.gpu
@pytest.mark.parametrize("model,keys", DECKS, ids=IDS)
def test_deck_on_the_gpu_matches_the_reference_program(model, keys, tmp_path):
    want, _ = reference_plots(model, keys)
    m, e = MODEL[model]
    write_deck(keys, str(tmp_path / "seq"))
    r = deck.run_deck(str(tmp_path / "seq"), m, e, log=silent)
    r["domain"].close()
    got = read_plots(str(tmp_path / "seq"))
    assert list(got) == list(want)
    for name in want:
        assert got[name].splitlines()[:3] == want[name].splitlines()[:3]
        x, y = numbers(want[name]), numbers(got[name])
        assert x.shape == y.shape, name
        scale = np.maximum(np.abs(x).max(axis=0), 1e-300)
        # one unit in the printed last digit
        assert (np.abs(x - y) / scale).max() <= 2e-5, name
    if not e:
        write_deck(keys, str(tmp_path / "fused"))
        r = deck.run_deck(str(tmp_path / "fused"), m, e, fused=True, log=silent)
        r["domain"].close()
        assert read_plots(str(tmp_path / "fused")) == got
