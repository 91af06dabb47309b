"""Host-side file formats (capaz_amd/io.py): the reference's noise.dat
permeability input (main.cpp:918-1032) and the state checkpoint format of
capaz_save_state."""
import math

import numpy as np
import pytest

from capaz_amd import io


def test_noise_dat_in_the_reference_number_format(tmp_path):
    # the reference's noise.dat: "Nx Ny", then rows of E-notation values with
    # leading / trailing blanks and CRLF line ends
    rng = np.random.default_rng(3)
    v = rng.uniform(-4, 3, (4, 6))
    lines = ["6 4"] + [" " + " ".join("%.14E" % x for x in row) + " " for row in v]
    p = tmp_path / "noise.dat"
    p.write_bytes("\r\n".join(lines).encode() + b"\r\n")
    K = io.load_noise_dat(str(p), 6, 4, 1)
    want = np.array([[1e-10 * math.exp(float("%.14E" % x)) for x in row] for row in v])
    assert K.shape == (1, 4, 6) and np.array_equal(K[0], want)


def test_noise_dat_three_dimensional_header_replicates_the_plane(tmp_path):
    v = np.arange(12.0).reshape(3, 4) / 10
    p = tmp_path / "noise.dat"
    io.write_noise_dat(str(p), v, nz=5)
    K = io.load_noise_dat(str(p), 4, 3, 5, upscale_l=0.1)
    assert K.shape == (5, 3, 4)
    assert np.array_equal(K[0], K[4])
    assert np.allclose(K[2], 1e-10 * np.exp(v) * 0.01, rtol=1e-15)


def test_noise_dat_grid_mismatch_warns_and_short_rows_fail(tmp_path):
    p = tmp_path / "noise.dat"
    io.write_noise_dat(str(p), np.zeros((3, 4)))
    with pytest.warns(UserWarning, match="differ"):
        io.load_noise_dat(str(p), 4, 2, 1)
    with pytest.warns(UserWarning, match="differ"), pytest.raises(ValueError, match="row 0"):
        io.load_noise_dat(str(p), 5, 3, 1)


def test_state_file_round_trip(tmp_path):
    rng = np.random.default_rng(4)
    fields = {0: rng.standard_normal((3, 4, 5)), 1: rng.standard_normal((3, 4, 5))}
    p = tmp_path / "s.bin"
    io.write_state(str(p), 3, 0, (5, 4, 3), 123, fields)
    st = io.read_state(str(p))
    assert st["model"] == 3 and st["grid"] == (5, 4, 3) and st["step"] == 123
    assert all(np.array_equal(st["fields"][k], fields[k]) for k in fields)
    p.write_bytes(p.read_bytes()[:-8])
    with pytest.raises(ValueError, match="truncated"):
        io.read_state(str(p))
