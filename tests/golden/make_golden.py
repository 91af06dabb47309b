"""Writes the golden vectors of tests/golden/*.npz from the compiled
reference CPU path (oracle/_ref, built by oracle/Makefile from the reference's
own cpu.cpp, gauss.cpp, model .cpp and energy.cpp).

Each fixture holds, for one deck and seeded initial state:
  in/<field>    every field before the first step,
  ent/<entry>/<field>  every field after each entry point of step 1,
  out<N>/<field> every field after N time steps (N = 1 and NSTEPS),
plus the deck as JSON.

``plots_<deck>.npz`` holds, for each input deck of capaz_cases.PLOT_DECKS,
the Tecplot plot files the reference CPU program (oracle/_ref/capaz_cpu_<m>,
the reference's own main.cpp and cpu.cpp) writes for it, keyed by file name,
plus the lines it prints (``__stdout__``).

The reference itself is not run anywhere else (in particular not next to the
GPU tests); these arrays are data only.

    python tests/golden/make_golden.py   # needs `make -C oracle ref refprog`
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from capaz_cases import PLOT_DECKS, case, plot_deck_id, random_state, ref_keys, write_deck  # noqa: E402
from oracle import ref  # noqa: E402

ENTRIES = ["S_calculation", "ro_P_Xi_calculation", "u_calculation", "roS_calculation",
           "H_E_current_calculation", "E_calculation", "P_S_calculation", "boundary_conditions"]

# name, model, grid, seed (None: the model's data_initialization), steps, deck overrides
FIXTURES = [
    ("ref2ph_deck_15x15x1", "2ph", (15, 15, 1), None, 50, {}),
    ("ref3ph_deck_1x15x1", "3ph", (1, 15, 1), None, 50, {}),
    ("ref2ph_8x9x7", "2ph", (8, 9, 7), 101, 10, {"SOURCE": 2, "DT": 1e-4}),
    ("ref3ph_8x9x7", "3ph", (8, 9, 7), 102, 10, {"DT": 1e-4}),
    ("ref3ph_src_7x6x5", "3ph", (7, 6, 5), 103, 10, {"SOURCE": 1, "S_W_GR": 0.45, "S_N_GR": 0.25, "DT": 1e-4}),
    ("ref3phe_6x7x5", "3phe", (6, 7, 5), 104, 5, {"DT": 1e-4}),
]


def make(name, model, dims, seed, nsteps, over):
    keys = case(model, *dims, **over)
    out = {}
    # entry by entry, step 1
    d = ref.RefDomain(model, ref_keys(keys), data_initialization=seed is None)
    if seed is not None:
        for k, v in random_state(keys, seed).items():
            d.field(k)[...] = v
    for f, a in d.fields().items():
        out[f"in/{f}"] = a
    for e in ENTRIES:
        if e in ("H_E_current_calculation", "E_calculation") and model != "3phe":
            continue
        d.call(e, keys["DT"])
        for f, a in d.fields().items():
            out[f"ent/{e}/{f}"] = a
    for f, a in d.fields().items():
        out[f"out1/{f}"] = a
    d.run(nsteps - 1, first_step=2)
    for f, a in d.fields().items():
        out[f"out{nsteps}/{f}"] = a
    d.close()
    meta = dict(name=name, model=model, keys=keys, seed=seed, nsteps=nsteps,
                generator="tests/golden/make_golden.py (compiled reference CPU path)")
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, __meta__=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8), **out)
    return path


def load(path):
    z = np.load(path)
    meta = json.loads(bytes(z["__meta__"]).decode())
    return meta, z


def plots_path(model: str, keys: dict) -> str:
    return os.path.join(HERE, f"plots_{plot_deck_id(model, keys)}.npz")


def make_plots(model: str, keys: dict) -> str:
    """Runs the reference CPU program on `keys` in a scratch directory and
    stores its plot files and printed lines."""
    import subprocess
    import tempfile

    binary = os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle", "_ref", f"capaz_cpu_{model}")
    if not os.path.exists(binary):
        raise SystemExit(f"{binary} is not built: run `make -C oracle refprog` first")
    with tempfile.TemporaryDirectory() as work:
        write_deck(keys, work)
        r = subprocess.run([binary], cwd=work, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                           timeout=300, check=True)
        plots = os.path.join(work, "plots")
        files = {n: np.frombuffer(open(os.path.join(plots, n), "rb").read(), dtype=np.uint8)
                 for n in sorted(os.listdir(plots))}
    meta = dict(model=model, keys=keys, generator="tests/golden/make_golden.py (reference CPU program)")
    path = plots_path(model, keys)
    np.savez_compressed(path, __meta__=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8),
                        __stdout__=np.frombuffer(r.stdout.encode(), dtype=np.uint8), **files)
    return path


def load_plots(model: str, keys: dict):
    """(plot files {name: text}, printed text) of the reference program for a
    deck of PLOT_DECKS, from its fixture."""
    z = np.load(plots_path(model, keys))
    meta = json.loads(bytes(z["__meta__"]).decode())
    if meta["keys"] != json.loads(json.dumps(keys)):
        raise ValueError(f"fixture {plots_path(model, keys)} was written for another deck")
    plots = {n: bytes(z[n]).decode() for n in sorted(z.files) if not n.startswith("__")}
    return plots, bytes(z["__stdout__"]).decode()


def main():
    if not ref.available("3ph"):
        raise SystemExit("oracle/_ref is not built: run `make -C oracle ref` first")
    for fx in FIXTURES:
        p = make(*fx)
        print(p, os.path.getsize(p))
    for model, keys in PLOT_DECKS:
        p = make_plots(model, keys)
        print(p, os.path.getsize(p))


if __name__ == "__main__":
    if sys.argv[1:] == ["plots"]:
        for model, keys in PLOT_DECKS:
            p = make_plots(model, keys)
            print(p, os.path.getsize(p))
    else:
        main()
