"""The oracle and the host build of the kernel bodies against the golden
vectors of the compiled reference (tests/golden/*.npz, written by
tests/golden/make_golden.py), bit for bit: every field after every entry
point of step 1 and after the last step."""
import glob
import os

import pytest

from capaz_cases import bit_equal, max_rel
from emul import emul
from golden.make_golden import ENTRIES, load
from oracle import oracle

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "ref*.npz")))
assert GOLDEN, "no golden fixtures"


def _initial(z, names):
    return {f: z[f"in/{f}"] for f in names if f"in/{f}" in z}


def _entries(meta):
    return [e for e in ENTRIES if meta["model"] == "3phe" or e not in ("H_E_current_calculation", "E_calculation")]


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
@pytest.mark.parametrize("impl", ["oracle", "emul"])
def test_golden(path, impl):
    meta, z = load(path)
    keys = meta["keys"]
    names = oracle.model_fields(keys)
    cls = oracle.OracleDomain if impl == "oracle" else emul.EmulDomain
    d = cls(keys, _initial(z, names))
    for e in _entries(meta):
        d.call(e, keys["DT"])
        for f in names:
            want = z[f"ent/{e}/{f}"]
            assert bit_equal(d.field(f), want), (e, f, max_rel(d.field(f), want))
    d.run(meta["nsteps"] - 1, first_step=2)
    n = meta["nsteps"]
    for f in names:
        want = z[f"out{n}/{f}"]
        assert bit_equal(d.field(f), want), (f, max_rel(d.field(f), want))


@pytest.mark.parametrize("path", [p for p in GOLDEN if "3phe" not in p],
                         ids=[os.path.basename(p)[:-4] for p in GOLDEN if "3phe" not in p])
def test_golden_fused_emulated(path):
    """The fused kernel's body, run on the CPU, reproduces the primary state
    of the reference after every step."""
    meta, z = load(path)
    keys = meta["keys"]
    names = oracle.model_fields(keys)
    d = emul.EmulDomain(keys, _initial(z, names))
    for _ in range(meta["nsteps"]):
        d.fused_step()
    n = meta["nsteps"]
    prim = ["P_w", "S_n"] + (["S_w"] if keys["MODEL"] == 3 else [])
    for f in prim:
        assert bit_equal(d.field(f), z[f"out{n}/{f}"]), (f, max_rel(d.field(f), z[f"out{n}/{f}"]))
