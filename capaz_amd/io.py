"""File formats around the step: the reference's permeability input
(`noise.dat`) and this backend's state checkpoints (capaz_save_state).

``load_noise_dat`` reads what ``load_permeability`` reads under
LOAD_K_FROM_FILE (main.cpp:918-1032): a header "Nx Ny" (two-dimensional
decks, Nz < 2) or "Nx Ny Nz", then Ny rows of Nx ln-multipliers v, giving
K = 1e-10 * exp(v) * upscale_l**2 in every plane k.

``read_state`` / ``write_state`` read and write the checkpoint format of
capaz_save_state / capaz_load_state (include/capaz.h) on the host, so a
checkpoint can be inspected, edited or produced without a GPU.
"""
from __future__ import annotations

import struct
import warnings

import numpy as np

MAGIC = b"CAPAZST1"


def load_noise_dat(path: str, nx: int, ny: int, nz: int, upscale_l: float = 1.0) -> np.ndarray:
    """K of shape (nz, ny, nx) from a noise.dat file (main.cpp:948-1025)."""
    with open(path, "rt") as f:
        head = f.readline().split()
        dims = [int(t) for t in head]
        if len(dims) < (2 if nz < 2 else 3):
            raise ValueError(f"{path}: header must hold {'Nx Ny' if nz < 2 else 'Nx Ny Nz'}")
        fnx, fny = dims[0], dims[1]
        fnz = dims[2] if nz >= 2 else 1
        if (fnx, fny, fnz) != (nx, ny, nz):
            # the reference warns and reads on (main.cpp:957-961)
            warnings.warn(f"{path}: Nx/Ny/Nz {fnx}x{fny}x{fnz} differ from the grid {nx}x{ny}x{nz}")
        rows = []
        for j in range(ny):
            vals = f.readline().split()
            if len(vals) < nx:
                raise ValueError(f"{path}: row {j} holds {len(vals)} values, needs {nx}")
            rows.append([float(v) for v in vals[:nx]])
    plane = 1e-10 * np.exp(np.array(rows)) * upscale_l ** 2
    K = np.broadcast_to(plane, (nz, ny, nx)).copy()
    if not (K > 0).all():  # test_positive (main.cpp:1027-1031)
        raise ValueError(f"{path}: non-positive permeability")
    return K


def write_noise_dat(path: str, lnmult: np.ndarray, nz: int = 1) -> None:
    """A noise.dat file for an (ny, nx) field of ln-multipliers."""
    ny, nx = lnmult.shape
    with open(path, "wt") as f:
        f.write(f"{nx} {ny}\n" if nz < 2 else f"{nx} {ny} {nz}\n")
        for row in lnmult:
            f.write(" ".join(repr(float(v)) for v in row) + " \n")


def read_state(path: str) -> dict:
    """A capaz_save_state file: {"model", "energy", "grid", "step", "fields": {id: array (Nz, Ny, Nx)}}."""
    with open(path, "rb") as f:
        if f.read(8) != MAGIC:
            raise ValueError(f"{path}: not a capaz state file")
        version, model, energy, nx, ny, nz = struct.unpack("<6i", f.read(24))
        if version != 1:
            raise ValueError(f"{path}: state format version {version}")
        (step,) = struct.unpack("<q", f.read(8))
        (nf,) = struct.unpack("<i", f.read(4))
        ids = struct.unpack(f"<{nf}i", f.read(4 * nf))
        n = nx * ny * nz
        fields = {}
        for fid in ids:
            a = np.frombuffer(f.read(8 * n), dtype="<f8")
            if a.size != n:
                raise ValueError(f"{path}: truncated")
            fields[fid] = a.reshape(nz, ny, nx).copy()
    return {"model": model, "energy": energy, "grid": (nx, ny, nz), "step": step, "fields": fields}


def write_state(path: str, model: int, energy: int, grid: tuple, step: int, fields: dict) -> None:
    """The inverse of read_state (fields: {field id: array of the grid's size})."""
    nx, ny, nz = grid
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(struct.pack("<6i", 1, model, energy, nx, ny, nz))
        f.write(struct.pack("<q", step))
        f.write(struct.pack("<i", len(fields)))
        f.write(struct.pack(f"<{len(fields)}i", *fields.keys()))
        for a in fields.values():
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
