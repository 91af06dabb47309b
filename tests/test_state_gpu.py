"""Checkpoint / restart through the C ABI (capaz_save_state,
capaz_load_state): a restored context continues bit for bit, including the
boundary-persistent P_n / P_g the reference's save.dat leaves out, and the
saved state, read on the host (io.read_state) and continued by the oracle,
gives the GPU continuation's values."""
import numpy as np
import pytest

from capaz_amd import capaz, io
from capaz_cases import bit_equal, case, max_rel, norm_rel, random_state
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("model", ["2ph", "3ph", "3phe"])
def test_restart_continues_bit_for_bit(tmp_path, model):
    over = dict(CLOSURE="corey", S_WR_0=0.1, S_NR_0=0.05) if model != "3phe" else {}
    keys = case(model, 13, 11, 9, **over)
    st = random_state(keys, 5)
    a = capaz.Domain(keys, st)
    a.run(4, first_step=1)
    path = str(tmp_path / "state.bin")
    a.save_state(path, step=4)
    f = io.read_state(path)
    assert f["step"] == 4 and f["grid"] == (13, 11, 9)
    for fid, arr in f["fields"].items():
        assert bit_equal(arr, a.field(capaz.FIELDS[fid]))
    a.run(3, first_step=5)
    b = capaz.Domain(keys, None)
    assert b.load_state(path) == 4
    b.run(3, first_step=5)
    for name in ("P_w", "S_n", "S_w") + (("P_n", "P_g") if model != "2ph" else ()):
        assert bit_equal(b.field(name), a.field(name)), name
    a.close()
    b.close()


def test_load_rejects_another_grid(tmp_path):
    keys = case("3ph", 9, 9, 9, CLOSURE="corey", S_WR_0=0.1, S_NR_0=0.05)
    a = capaz.Domain(keys, random_state(keys, 6))
    path = str(tmp_path / "s.bin")
    a.save_state(path)
    b = capaz.Domain(dict(keys, NX=10), None)
    with pytest.raises(capaz.CapazError, match="grid"):
        b.load_state(path)
    a.close()
    b.close()


@pytest.mark.parametrize("model", ["2ph", "3ph", "3phe"])
def test_restart_continues_like_the_oracle(tmp_path, model):
    """A checkpoint written after 4 GPU steps, read back on the host and
    loaded into the oracle, continues for 3 steps to the values of the GPU
    context restored from the same file: bit for bit with the Corey
    closures (two and three phases), within 1e-9 for the energy model, whose
    mu(T) and lambda_g(T) call exp/pow (the device library's on the GPU)."""
    over = dict(CLOSURE="corey", S_WR_0=0.1, S_NR_0=0.05) if model != "3phe" else {}
    keys = case(model, 13, 11, 9, SOURCE=1, DT=1e-4, **over)
    a = capaz.Domain(keys, random_state(keys, 17))
    a.run(4, first_step=1)
    path = str(tmp_path / "state.bin")
    a.save_state(path, step=4)
    a.close()
    saved = io.read_state(path)
    od = oracle.OracleDomain(keys, {capaz.FIELDS[fid]: arr for fid, arr in saved["fields"].items()})
    od.run(3, first_step=saved["step"] + 1)
    b = capaz.Domain(keys, None)
    step = b.load_state(path)
    b.run(3, first_step=step + 1)
    for f in od.names:
        got, want = b.field(f), od.field(f)
        if model == "3phe":
            r = max_rel(got, want) if f in ("P_w", "S_w", "S_n", "T") else norm_rel(got, want)
            assert r <= 1e-9, (f, r)
        else:
            assert bit_equal(got, want), (f, max_rel(got, want))
    b.close()
