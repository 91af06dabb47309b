"""Drop-in integration: the reference's own driver (main.cpp, the model .cpp,
gauss.cpp, no_communication.cpp, shared_test.cpp) linked against the
capaz_amd adapter and libcapaz_amd.so in place of cpu.cpp
(oracle/_ref/capaz_dropin_<m>), run on an input deck, writes the same Tecplot
plot files as the reference CPU program built from the same sources
(oracle/_ref/capaz_cpu_<m>).  Both binaries are built in this container by
`make -C oracle dropin` (they need the reference sources); the test is
skipped where they are absent.

Plots print six significant digits (main.cpp:700-780); every column is
compared relative to its largest magnitude, since near-zero velocities are
printed in full exponent range."""
import os
import subprocess

import numpy as np
import pytest

from capaz_cases import case, ref_keys

pytestmark = pytest.mark.gpu

REF = os.path.join(os.path.dirname(os.path.dirname(__file__)), "oracle", "_ref")

DECKS = [
    ("2ph", case("2ph", 15, 15, 1, TIMEX=0.2, SAVE_PLOTS=100, PRINT_SCREEN=100)),
    ("3ph", case("3ph", 1, 15, 1, TIMEX=0.2, SAVE_PLOTS=100, PRINT_SCREEN=100)),
    ("3ph", case("3ph", 10, 12, 9, DT=1e-4, TIMEX=0.005, SAVE_PLOTS=25, PRINT_SCREEN=25, SOURCE=1)),
    ("3phe", case("3phe", 1, 9, 1, TIMEX=0.02, SAVE_PLOTS=10, PRINT_SCREEN=10)),
]


def run(binary, deck, workdir):
    os.makedirs(workdir, exist_ok=True)
    with open(os.path.join(workdir, "defines.ini"), "w") as f:
        for k, v in ref_keys(deck).items():
            f.write(f"{k}={v}\n")
    r = subprocess.run([binary], cwd=workdir, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout[-2000:]
    plots = os.path.join(workdir, "plots")
    return {n: open(os.path.join(plots, n)).read() for n in sorted(os.listdir(plots))}


def numbers(text):
    rows = [l.split() for l in text.splitlines()[3:] if l.strip()]
    return np.array([[float(x) for x in r] for r in rows])


@pytest.mark.parametrize("model,deck", DECKS, ids=[f"{m}-{d['NX']}x{d['NY']}x{d['NZ']}" for m, d in DECKS])
def test_dropin_driver_reproduces_reference_plots(model, deck, tmp_path):
    cpu, gpu = os.path.join(REF, f"capaz_cpu_{model}"), os.path.join(REF, f"capaz_dropin_{model}")
    if not (os.path.exists(cpu) and os.path.exists(gpu)):
        pytest.skip("drop-in drivers not built (make -C oracle dropin needs the reference sources)")
    a = run(cpu, deck, str(tmp_path / "cpu"))
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
