"""Writes the golden vectors of tests/golden/*.npz from the compiled
reference CPU path (oracle/_ref, built by oracle/Makefile from the reference's
own cpu.cpp, gauss.cpp, model .cpp and energy.cpp).

Each fixture holds, for one deck and seeded initial state:
  in/<field>    every field before the first step,
  ent/<entry>/<field>  every field after each entry point of step 1,
  out<N>/<field> every field after N time steps (N = 1 and NSTEPS),
plus the deck as JSON.  The reference itself is not run anywhere else (in
particular not next to the GPU tests); these arrays are data only.

    python tests/golden/make_golden.py        # needs oracle/_ref built
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from capaz_cases import case, random_state, ref_keys  # noqa: E402
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


def main():
    if not ref.available("3ph"):
        raise SystemExit("oracle/_ref is not built: run `make -C oracle ref` first")
    for fx in FIXTURES:
        p = make(*fx)
        print(p, os.path.getsize(p))


if __name__ == "__main__":
    main()
