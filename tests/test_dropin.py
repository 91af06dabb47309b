"""Drop-in integration: the reference's own driver (main.cpp, the model .cpp,
gauss.cpp, no_communication.cpp, shared_test.cpp) linked against the
capaz_amd adapter and libcapaz_amd.so in place of cpu.cpp
(capaz_amd/adapter/_bin/capaz_dropin_<m>, built by `make -C oracle dropin`, which
__graft_entry__.build() runs where the reference sources are present), run on
an input deck, writes the Tecplot plot files the reference CPU program writes
for it (committed fixtures tests/golden/plots_<deck>.npz).

Plots print six significant digits (main.cpp:700-780); every column is
compared relative to its largest magnitude, since near-zero velocities are
printed in full exponent range."""
import os
import subprocess

import numpy as np
import pytest

from capaz_cases import PLOT_DECKS, plot_deck_id, write_deck
from golden.make_golden import load_plots

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(__file__)), "capaz_amd", "adapter", "_bin")

DECKS = [PLOT_DECKS[i] for i in (0, 2, 3, 5)]


def run(binary, deck, workdir, fused=True):
    write_deck(deck, workdir)
    env = dict(os.environ, CAPAZ_DROPIN_FUSED="1" if fused else "0")
    r = subprocess.run([binary], cwd=workdir, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120, env=env)
    assert r.returncode == 0, r.stdout[-2000:]
    plots = os.path.join(workdir, "plots")
    return {n: open(os.path.join(plots, n)).read() for n in sorted(os.listdir(plots))}


def numbers(text):
    rows = [l.split() for l in text.splitlines()[3:] if l.strip()]
    return np.array([[float(x) for x in r] for r in rows])


@pytest.mark.parametrize("model,deck", DECKS, ids=[plot_deck_id(m, d) for m, d in DECKS])
def test_dropin_driver_reproduces_reference_plots(model, deck, tmp_path):
    gpu = os.path.join(BIN, f"capaz_dropin_{model}")
    assert os.path.exists(gpu), "drop-in driver not built: __graft_entry__.build() builds it (make -C oracle dropin)"
    a, _ = load_plots(model, deck)
    b = run(gpu, deck, str(tmp_path / "gpu"))
    assert list(a) == list(b) and len(a) >= 2
    for name in a:
        x, y = numbers(a[name]), numbers(b[name])
        assert x.shape == y.shape, name
        scale = np.maximum(np.abs(x).max(axis=0), 1e-300)
        rel = np.abs(x - y) / scale
        r, c = np.unravel_index(np.argmax(rel), rel.shape)
        # one unit in the printed 6th digit
        assert rel.max() <= 2e-5, (name, f"row {r} column {c}: cpu {x[r, c]!r} gpu {y[r, c]!r}", rel.max(),
                                   a[name].splitlines()[1])


@pytest.mark.parametrize("model,deck", [PLOT_DECKS[i] for i in (1, 3, 6)],
                         ids=[plot_deck_id(*PLOT_DECKS[i]) for i in (1, 3, 6)])
def test_dropin_fused_steps_write_the_entry_points_plots(model, deck, tmp_path):
    """The adapter running each time step as one fused capaz_step writes the
    same plot files, character for character, as with every entry point run
    as the driver calls it (CAPAZ_DROPIN_FUSED=0): the same bits."""
    gpu = os.path.join(BIN, f"capaz_dropin_{model}")
    assert os.path.exists(gpu), "drop-in driver not built: __graft_entry__.build() builds it (make -C oracle dropin)"
    a = run(gpu, deck, str(tmp_path / "fused"), fused=True)
    b = run(gpu, deck, str(tmp_path / "entry"), fused=False)
    assert list(a) == list(b) and len(a) >= 2
    for name in a:
        assert a[name] == b[name], name
