"""ctypes access to the compiled reference CPU path (test infrastructure).

Loads ``oracle/_ref/libcapaz_ref_<model>.so`` -- the reference's own cpu.cpp,
gauss.cpp, model .cpp and energy.cpp compiled by ``oracle/Makefile`` around
``ref_harness.cpp`` -- and exposes one single-domain reference context whose
``consts`` block is set from defines.ini keys (main.cpp:1327-1626) and whose
host arrays are numpy views.  Only tests, the golden-vector script and the
CPU-baseline leg of bench.py use this module, to compare against.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

MODELS = ("2ph", "3ph", "3phe")

_libs: dict[str, ctypes.CDLL] = {}


def lib_path(model: str) -> str:
    return os.path.join(REF_DIR, f"libcapaz_ref_{model}.so")


def available(model: str = "3ph") -> bool:
    return os.path.exists(lib_path(model))


def _lib(model: str) -> ctypes.CDLL:
    if model not in _libs:
        lib = ctypes.CDLL(lib_path(model))
        lib.capaz_ref_create.restype = ctypes.c_void_p
        lib.capaz_ref_set.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_double]
        lib.capaz_ref_init.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.capaz_ref_cells.argtypes = [ctypes.c_void_p]
        lib.capaz_ref_cells.restype = ctypes.c_long
        lib.capaz_ref_field.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        lib.capaz_ref_field.restype = ctypes.POINTER(ctypes.c_double)
        lib.capaz_ref_call.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_double]
        lib.capaz_ref_run.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_long]
        lib.capaz_ref_destroy.argtypes = [ctypes.c_void_p]
        lib.capaz_ref_field_name.restype = ctypes.c_char_p
        lib.capaz_ref_reverse_matrix.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        lib.capaz_ref_mult_matrix_vector.argtypes = [ctypes.POINTER(ctypes.c_double)] * 3 + [ctypes.c_int]
        _libs[model] = lib
    return _libs[model]


def field_names(model: str) -> list[str]:
    lib = _lib(model)
    return [lib.capaz_ref_field_name(i).decode() for i in range(lib.capaz_ref_field_count())]


def reverse_matrix(a: np.ndarray, model: str = "3ph") -> tuple[int, np.ndarray]:
    """The reference's gauss.cpp:6-92 applied to a copy of row-major ``a``."""
    n = a.shape[0]
    buf = np.ascontiguousarray(a, dtype=np.float64).copy()
    rc = _lib(model).capaz_ref_reverse_matrix(buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n)
    return rc, buf


class RefDomain:
    """One reference single-domain context (consts + ptr_Arrays)."""

    def __init__(self, model: str, keys: dict[str, float], data_initialization: bool = False):
        self.model = model
        self._lib = _lib(model)
        self._ctx = self._lib.capaz_ref_create()
        for k, v in keys.items():
            if self._lib.capaz_ref_set(self._ctx, k.encode(), float(v)) != 0:
                raise KeyError(f"reference has no defines.ini key {k!r} for model {model}")
        self._lib.capaz_ref_init(self._ctx, 1 if data_initialization else 0)
        self.shape = (int(keys["NZ"]), int(keys["NY"]), int(keys["NX"]))  # k, j, i (x fastest)
        self.n = int(self._lib.capaz_ref_cells(self._ctx))
        self.dt = float(keys["DT"])

    def field(self, name: str) -> np.ndarray:
        """Writable numpy view (k, j, i) of one reference host array."""
        p = self._lib.capaz_ref_field(self._ctx, name.encode())
        if not p:
            raise KeyError(name)
        return np.ctypeslib.as_array(p, shape=(self.n,)).reshape(self.shape)

    def fields(self) -> dict[str, np.ndarray]:
        return {f: self.field(f).copy() for f in field_names(self.model)}

    def call(self, entry: str, t: float = 0.0) -> None:
        if self._lib.capaz_ref_call(self._ctx, entry.encode(), float(t)) != 0:
            raise KeyError(entry)

    def run(self, nsteps: int, first_step: int = 1) -> None:
        self._lib.capaz_ref_run(self._ctx, first_step, nsteps)

    def close(self) -> None:
        if self._ctx:
            self._lib.capaz_ref_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
