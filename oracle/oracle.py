"""ctypes access to the oracle's C restatement (test infrastructure).

``oracle/liboracle.so`` is built from ``capaz_oracle.c`` by ``oracle/Makefile``
(and by ``__graft_entry__.build()``).  Parameters are given as defines.ini
keys (main.cpp:1327-1626) plus keys of the backend's own: ``MODEL``
(2 or 3), ``ENERGY`` (0/1), ``CLOSURE`` ("reference" or "corey"),
``COREY_N``, ``P_BOTTOM`` and the energy closure of the configurations
(``ENERGY_CLOSURE`` "config" with ``LAMBDA_EFF``, ``C_R``, ``RO_R``,
``T_TOP``, ``T_BOTTOM``, ``MU_T_A``, ``MU_T_REF``).  Only tests, smoke()
and the CPU-baseline leg of bench.py use this module, to compare against.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

FIELDS = (
    "S_n", "S_w", "P_w", "P_n", "ro_w", "ro_n", "ux_w", "uy_w", "uz_w", "ux_n", "uy_n", "uz_n",
    "Xi_w", "Xi_n", "roS_w", "roS_w_old", "roS_n", "roS_n_old", "m", "K",
    "S_g", "P_g", "ro_g", "ux_g", "uy_g", "uz_g", "Xi_g", "roS_g", "roS_g_old",
    "T", "H_w", "H_n", "H_g", "H_r", "E", "E_new",
)
THREE_PHASE_FIELDS = FIELDS[20:29]
ENERGY_FIELDS = FIELDS[29:]


class _Params(ctypes.Structure):
    _fields_ = [
        ("model", ctypes.c_int), ("energy", ctypes.c_int), ("closure", ctypes.c_int),
        ("Nx", ctypes.c_int), ("Ny", ctypes.c_int), ("Nz", ctypes.c_int),
        ("source", ctypes.c_int), ("iterations", ctypes.c_int),
        ("hx", ctypes.c_double), ("hy", ctypes.c_double), ("hz", ctypes.c_double), ("dt", ctypes.c_double),
        ("P_atm", ctypes.c_double), ("g_const", ctypes.c_double),
        ("mu_w", ctypes.c_double), ("mu_n", ctypes.c_double), ("mu_g", ctypes.c_double),
        ("ro0_w", ctypes.c_double), ("ro0_n", ctypes.c_double), ("ro0_g", ctypes.c_double),
        ("beta_w", ctypes.c_double), ("beta_n", ctypes.c_double),
        ("lambda_", ctypes.c_double),
        ("S_wr", ctypes.c_double), ("S_nr", ctypes.c_double), ("S_gr", ctypes.c_double),
        ("P_d", ctypes.c_double),
        ("S_w_gr", ctypes.c_double), ("S_n_gr", ctypes.c_double),
        ("corey_n", ctypes.c_double),
        ("P_bottom", ctypes.c_double), ("energy_closure", ctypes.c_int),
        ("lambda_eff", ctypes.c_double), ("c_r", ctypes.c_double), ("ro_r", ctypes.c_double),
        ("T_top", ctypes.c_double), ("T_bottom", ctypes.c_double),
        ("mu_T_a", ctypes.c_double), ("mu_T_ref", ctypes.c_double),
    ]


class _Fields(ctypes.Structure):
    _fields_ = [(f, ctypes.POINTER(ctypes.c_double)) for f in FIELDS]


_lib = None


def available() -> bool:
    return os.path.exists(LIB)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built (run `make -C oracle oracle`)")
        L = ctypes.CDLL(LIB)
        P, F = ctypes.POINTER(_Params), ctypes.POINTER(_Fields)
        for name in ("S_calculation", "ro_P_Xi_calculation", "u_calculation", "H_E_current_calculation",
                     "E_calculation", "P_S_calculation", "boundary_conditions"):
            getattr(L, "oracle_" + name).argtypes = [P, F]
        L.oracle_roS_calculation.argtypes = [P, F, ctypes.c_double]
        L.oracle_time_step.argtypes = [P, F, ctypes.c_double]
        L.oracle_run.argtypes = [P, F, ctypes.c_long, ctypes.c_long]
        L.oracle_reverse_matrix.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        L.oracle_relperm.argtypes = [P, ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
        L.oracle_threads.restype = ctypes.c_int
        L.oracle_threads.argtypes = []
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def params_from_keys(keys: dict) -> _Params:
    g = lambda k, d=0.0: float(keys.get(k, d))
    closure = keys.get("CLOSURE", "reference")
    p = _Params()
    p.model = int(keys.get("MODEL", 3))
    p.energy = int(keys.get("ENERGY", 0))
    p.closure = 1 if closure == "corey" else 0
    p.Nx, p.Ny, p.Nz = int(keys["NX"]), int(keys["NY"]), int(keys["NZ"])
    p.source = int(keys.get("SOURCE", 0))
    p.iterations = int(keys.get("ITERATIONS", 1))
    p.hx, p.hy, p.hz, p.dt = g("HX"), g("HY"), g("HZ"), g("DT")
    p.P_atm, p.g_const = g("P_ATM"), g("G_CONST")
    p.mu_w, p.mu_n, p.mu_g = g("MU_W"), g("MU_N"), g("MU_G")
    p.ro0_w, p.ro0_n, p.ro0_g = g("RO_W"), g("RO_N"), g("RO_G")
    p.beta_w, p.beta_n = g("BETA_W"), g("BETA_N")
    p.lambda_ = g("LAMBDA_0")
    p.S_wr, p.S_nr, p.S_gr = g("S_WR_0"), g("S_NR_0"), g("S_GR_0")
    p.P_d = g("P_D_0")
    p.S_w_gr, p.S_n_gr = g("S_W_GR"), g("S_N_GR")
    p.corey_n = g("COREY_N", 2.0)
    p.P_bottom = g("P_BOTTOM")
    p.energy_closure = 1 if keys.get("ENERGY_CLOSURE", "reference") == "config" else 0
    p.lambda_eff, p.c_r, p.ro_r = g("LAMBDA_EFF"), g("C_R"), g("RO_R")
    p.T_top, p.T_bottom = g("T_TOP"), g("T_BOTTOM")
    p.mu_T_a, p.mu_T_ref = g("MU_T_A"), g("MU_T_REF")
    return p


def model_fields(keys: dict) -> list[str]:
    model, energy = int(keys.get("MODEL", 3)), int(keys.get("ENERGY", 0))
    out = [f for f in FIELDS[:20]]
    if model == 3:
        out += list(THREE_PHASE_FIELDS)
    if energy:
        out += [f for f in ENERGY_FIELDS if model == 3 or f != "H_g"]
    return out


class OracleDomain:
    """One oracle domain: parameters plus every field as a (Nz, Ny, Nx) array."""

    def __init__(self, keys: dict, fields: dict[str, np.ndarray] | None = None):
        self.keys = dict(keys)
        self.p = params_from_keys(keys)
        self.shape = (self.p.Nz, self.p.Ny, self.p.Nx)
        self.names = model_fields(keys)
        self.arrays = {n: np.zeros(self.shape) for n in self.names}
        if fields:
            for n, v in fields.items():
                if n in self.arrays:
                    self.arrays[n][...] = v
        self._f = _Fields()
        for n in self.names:
            setattr(self._f, n, self.arrays[n].ctypes.data_as(ctypes.POINTER(ctypes.c_double)))

    def field(self, name: str) -> np.ndarray:
        return self.arrays[name]

    def fields(self) -> dict[str, np.ndarray]:
        return {n: a.copy() for n, a in self.arrays.items()}

    def call(self, entry: str, t: float = 0.0) -> None:
        L = lib()
        if entry == "roS_calculation":
            L.oracle_roS_calculation(ctypes.byref(self.p), ctypes.byref(self._f), t)
        elif entry == "time_step_function":
            L.oracle_time_step(ctypes.byref(self.p), ctypes.byref(self._f), t)
        else:
            getattr(L, "oracle_" + entry)(ctypes.byref(self.p), ctypes.byref(self._f))

    def run(self, nsteps: int, first_step: int = 1) -> None:
        lib().oracle_run(ctypes.byref(self.p), ctypes.byref(self._f), first_step, nsteps)


def threads() -> int:
    """Threads the oracle's per-cell passes use."""
    return lib().oracle_threads()


def set_threads(n: int) -> None:
    lib().oracle_set_threads(n)


def reverse_matrix(a: np.ndarray) -> tuple[int, np.ndarray]:
    buf = np.ascontiguousarray(a, dtype=np.float64).copy()
    rc = lib().oracle_reverse_matrix(buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.shape[0])
    return rc, buf


def relperm(keys: dict, S_w: float, S_n: float) -> np.ndarray:
    out = np.zeros(3)
    p = params_from_keys(keys)
    lib().oracle_relperm(ctypes.byref(p), S_w, S_n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out
