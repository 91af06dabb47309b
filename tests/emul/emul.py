"""ctypes access to the host build of the kernel bodies (test infrastructure).

``EmulDomain`` runs the per-cell bodies of the HIP kernels on numpy arrays on
the CPU, with the same interface as ``oracle.OracleDomain``, so the kernels'
arithmetic can be compared with the oracle bit for bit without a GPU.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from capaz_amd.capaz import FIELDS, Params, params_from_keys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libcapaz_emul.so")

ENTRY = {"S_calculation": 0, "ro_P_Xi_calculation": 1, "u_calculation": 2, "roS_calculation": 3,
         "H_E_current_calculation": 4, "E_calculation": 5, "P_S_calculation": 6, "boundary_conditions": 7,
         "newton_energy_two_pass": 8}

_lib = None


def available() -> bool:
    return os.path.exists(LIB)


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(LIB)
        PP = ctypes.POINTER(ctypes.POINTER(ctypes.c_double))
        L.capaz_emul_entry.argtypes = [ctypes.POINTER(Params), PP, ctypes.c_int]
        L.capaz_emul_fused_step.argtypes = [ctypes.POINTER(Params), PP, PP]
        L.capaz_emul_reverse_matrix.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        L.capaz_emul_adjoint_step.argtypes = [ctypes.POINTER(Params), PP, PP, PP, PP, PP]
        L.capaz_emul_tangent_step.argtypes = [ctypes.POINTER(Params), PP, PP, PP, PP]
        L.capaz_emul_vg_d.argtypes = [ctypes.POINTER(Params), ctypes.c_double, ctypes.c_double,
                                      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.capaz_emul_coef_d.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_double)]
        L.capaz_emul_validate.argtypes = [ctypes.POINTER(Params)]
        L.capaz_emul_validate.restype = ctypes.c_char_p
        _lib = L
    return _lib


def validate(keys: dict) -> str:
    p = params_from_keys(keys)
    return lib().capaz_emul_validate(ctypes.byref(p)).decode()


def vg_d(keys: dict, x: float, y: float):
    """The adjoint's van Genuchten k_r at effective saturations (x, y):
    ((k_w, k_n, k_g), (dk_w/dx, dk_n/dx, dk_n/dy, dk_g/dz))."""
    k, d = (ctypes.c_double * 3)(), (ctypes.c_double * 4)()
    p = params_from_keys(keys)
    if lib().capaz_emul_vg_d(ctypes.byref(p), x, y, k, d):
        raise ValueError("invalid parameters")
    return tuple(k), tuple(d)


def coef_d(keys: dict, Pw: float, Sw: float, Sn: float, K: float):
    """The adjoint's coefficients of one interior cell: (Xi[3], dXi/dS_w[3],
    dXi/dS_n[3])."""
    x = (ctypes.c_double * 6)(Pw, Sw, Sn, Pw, Pw, K)
    out = (ctypes.c_double * 9)()
    p = params_from_keys(keys)
    if lib().capaz_emul_coef_d(ctypes.byref(p), x, 1, out):
        raise ValueError("invalid parameters")
    v = list(out)
    return v[0:3], v[3:6], v[6:9]


def reverse_matrix(a: np.ndarray, fast: bool = False, a02zero: bool = False) -> tuple[int, np.ndarray]:
    """The kernels' reverse_matrix<n>; fast=True runs the 3x3 no-pivot path
    with its fallback (for n = 4, 5 the no-pivot routine alone, rc 1 when it
    refuses), a02zero=True its variant for a[0,2] == +0 (the one newton3
    uses)."""
    buf = np.ascontiguousarray(a, dtype=np.float64).copy()
    n = a.shape[0]
    if a02zero:
        n = -4
    elif fast:
        n = -14 if n == 4 else -n
    rc = lib().capaz_emul_reverse_matrix(buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n)
    return rc, buf


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None


class EmulDomain:
    def __init__(self, keys: dict, fields: dict | None = None):
        self.keys = dict(keys)
        self.p = params_from_keys(keys)
        why = lib().capaz_emul_validate(ctypes.byref(self.p)).decode()
        if why:
            raise ValueError(why)
        self.shape = (self.p.Nz, self.p.Ny, self.p.Nx)
        model, energy = self.p.model, self.p.energy
        self.names = [f for i, f in enumerate(FIELDS)
                      if i < 20 or (20 <= i < 29 and model == 3) or (i >= 29 and energy)]
        self.arrays = {n: np.zeros(self.shape) for n in self.names}
        if fields:
            for n, v in fields.items():
                if n in self.arrays:
                    self.arrays[n][...] = v

    def field(self, name):
        return self.arrays[name]

    def fields(self):
        return {n: a.copy() for n, a in self.arrays.items()}

    def _table(self):
        arr = (ctypes.POINTER(ctypes.c_double) * len(FIELDS))()
        for i, f in enumerate(FIELDS):
            arr[i] = _ptr(self.arrays.get(f))
        return arr

    def call(self, entry: str, t: float = 0.0) -> None:
        if entry == "time_step_function":
            for e in ("S_calculation", "ro_P_Xi_calculation", "u_calculation", "roS_calculation"):
                self.call(e, t)
            if self.p.energy:
                self.call("H_E_current_calculation")
                self.call("E_calculation")
            self.call("P_S_calculation")
            self.call("boundary_conditions")
            return
        rc = lib().capaz_emul_entry(ctypes.byref(self.p), self._table(), ENTRY[entry])
        if rc < 0 or (rc and entry != "newton_energy_two_pass"):
            raise RuntimeError(f"emul entry {entry} failed")
        self.redone = rc

    def run(self, nsteps: int, first_step: int = 1) -> None:
        for s in range(first_step, first_step + nsteps):
            self.call("time_step_function", s * self.p.dt)

    def fused_step(self) -> None:
        """One fused-kernel step on the primary state (P_w, S_w, S_n, P_n, P_g)."""
        names = ("P_w", "S_w", "S_n", "P_n", "P_g")
        cur = [self.arrays.get(n) for n in names] + [self.arrays["m"], self.arrays["K"]]
        nxt = [a.copy() if a is not None else None for a in cur[:5]]
        C = (ctypes.POINTER(ctypes.c_double) * 7)(*[_ptr(a) for a in cur])
        N = (ctypes.POINTER(ctypes.c_double) * 5)(*[_ptr(a) for a in nxt])
        if lib().capaz_emul_fused_step(ctypes.byref(self.p), C, N):
            raise RuntimeError("emul fused step failed")
        # Primary state only: P_w, S_n (+ S_w, P_n, P_g for three phases; of
        # P_n/P_g the kernel writes the boundary cells, the interior keeps the
        # copied previous values).
        primary = names if self.p.model == 3 else ("P_w", "S_n")
        for n, a in zip(names, nxt):
            if n in primary:
                self.arrays[n][...] = a


# ---- discrete adjoint (capaz_adjoint.h) on the host -----------------------------

ADJ_STATE = ("P_w", "S_w", "S_n", "P_n", "P_g")


def adjoint_step(keys: dict, state: dict, lam1: dict, grad: dict) -> dict:
    """One adjoint step: ``state`` holds x_t (P_w, S_w, S_n, P_n, P_g) and the
    parameters m, K; ``lam1`` the adjoint of x_{t+1}.  Returns lambda_t and
    adds the step's share to grad["lnK"], grad["m"]."""
    p = params_from_keys(keys)
    shape = state["P_w"].shape
    dtm = p.dt / state["m"]
    st = [np.ascontiguousarray(state[n]) for n in ADJ_STATE] + [np.ascontiguousarray(state["m"]),
                                                              np.ascontiguousarray(state["K"]), dtm]
    l1 = [np.ascontiguousarray(lam1[n]) for n in ADJ_STATE]
    work = [np.zeros(shape) for _ in range(6)]
    lam = {n: np.zeros(shape) for n in ADJ_STATE}
    P = ctypes.POINTER(ctypes.c_double)
    tab = lambda arrs: (P * len(arrs))(*[_ptr(a) for a in arrs])
    rc = lib().capaz_emul_adjoint_step(ctypes.byref(p), tab(st), tab(l1), tab(work), tab([lam[n] for n in ADJ_STATE]),
                                       tab([grad["lnK"], grad["m"]]))
    if rc:
        raise RuntimeError("emul adjoint step failed")
    return lam


def objective(states_at: dict, obs: dict) -> float:
    """0.5 * sum over the observed steps and cells of the squared S_w and
    S_n mismatch; states_at[step] = (S_w, S_n) flattened.  Summed sample by
    sample, cell by cell, as capaz_objective does (the same bits)."""
    total = 0.0
    for s_i, step in enumerate(obs["steps"]):
        Sw, Sn = states_at[step]
        dw = Sw[obs["cells"]] - obs["S_w"][s_i]
        dn = Sn[obs["cells"]] - obs["S_n"][s_i]
        for a, b in zip(dw.tolist(), dn.tolist()):
            total += a * a + b * b
    return 0.5 * total


def gradient(keys: dict, fields: dict, nsteps: int, obs: dict):
    """J and dJ/d lnK, dJ/dm of `nsteps` fused steps from `fields`, by the
    host build of the forward step and of the adjoint (all states kept)."""
    d = EmulDomain(keys, fields)
    two = d.p.model == 2  # state (P_w, S_n); S_w = 1 - S_n
    states, seen = [], {}
    for t in range(nsteps):
        states.append({n: d.arrays[n].copy() if n in d.arrays else np.zeros(d.shape)
                       for n in ADJ_STATE + ("m", "K")})
        d.fused_step()
        if t + 1 in obs["steps"]:
            Sn = d.arrays["S_n"].ravel().copy()
            seen[t + 1] = (1. - Sn if two else d.arrays["S_w"].ravel().copy(), Sn)
    J = objective(seen, obs)
    shape = d.shape
    grad = {"lnK": np.zeros(shape), "m": np.zeros(shape)}

    def add_obs(lam, step):
        if step in obs["steps"]:
            s_i = list(obs["steps"]).index(step)
            Sw, Sn = seen[step]
            lw, ln = lam["S_w"].reshape(-1), lam["S_n"].reshape(-1)
            dw = Sw[obs["cells"]] - obs["S_w"][s_i]
            dn = Sn[obs["cells"]] - obs["S_n"][s_i]
            if two:
                np.add.at(ln, obs["cells"], dn - dw)
            else:
                np.add.at(lw, obs["cells"], dw)
                np.add.at(ln, obs["cells"], dn)

    lam = {n: np.zeros(shape) for n in ADJ_STATE}
    add_obs(lam, nsteps)
    for t in range(nsteps - 1, -1, -1):
        lam = adjoint_step(keys, states[t], lam, grad)
        if t >= 1:
            add_obs(lam, t)
    return J, grad["lnK"], grad["m"]


def tangent(keys: dict, fields: dict, nsteps: int, obs: dict, dlnK: np.ndarray, dm: np.ndarray):
    """J and its directional derivative dJ along (dlnK, dm) by the host
    build of the tangent-linear step, run alongside the forward step."""
    p = params_from_keys(keys)
    d = EmulDomain(keys, fields)
    shape = d.shape
    dx = {n: np.zeros(shape) for n in ADJ_STATE}
    theta = [np.ascontiguousarray(dlnK, dtype=np.float64), np.ascontiguousarray(dm, dtype=np.float64)]
    P = ctypes.POINTER(ctypes.c_double)
    tab = lambda arrs: (P * len(arrs))(*[_ptr(a) for a in arrs])
    seen, dseen = {}, {}
    for t in range(nsteps):
        st = [np.ascontiguousarray(d.arrays[n]) for n in ADJ_STATE] + [d.arrays["m"], d.arrays["K"],
                                                                      p.dt / d.arrays["m"]]
        nxt = {n: np.zeros(shape) for n in ADJ_STATE}
        if lib().capaz_emul_tangent_step(ctypes.byref(p), tab(st), tab([dx[n] for n in ADJ_STATE]), tab(theta),
                                         tab([nxt[n] for n in ADJ_STATE])):
            raise RuntimeError("emul tangent step failed")
        d.fused_step()
        dx = nxt
        if t + 1 in obs["steps"]:
            seen[t + 1] = (d.arrays["S_w"].ravel().copy(), d.arrays["S_n"].ravel().copy())
            dseen[t + 1] = (dx["S_w"].ravel().copy(), dx["S_n"].ravel().copy())
    J = objective(seen, obs)
    dJ = 0.0
    for s_i, step in enumerate(obs["steps"]):
        Sw, Sn = seen[step]
        dSw, dSn = dseen[step]
        c = obs["cells"]
        dJ += float(np.sum((Sw[c] - obs["S_w"][s_i]) * dSw[c] + (Sn[c] - obs["S_n"][s_i]) * dSn[c]))
    return J, dJ
