"""Checkpoint / restart through the C ABI (capaz_save_state,
capaz_load_state): a restored context continues bit for bit, including the
boundary-persistent P_n / P_g the reference's save.dat leaves out."""
import numpy as np
import pytest

from capaz_amd import capaz, io
from capaz_cases import bit_equal, case, random_state

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
