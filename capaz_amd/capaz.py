"""Python host side of the backend: ctypes binding of the C ABI
(include/capaz.h) and a ``Domain`` class whose methods carry the names of the
reference entry points they stand in for (defines.h:115-224), so driver code
and tests read like capaz's own ``time_step_function`` (main.cpp:93-108).

The library is ``capaz_amd/libcapaz_amd.so`` (built in-tree by
``capaz_amd/build.py``).  There is no fallback: if the library or a GPU is
missing, creating a Domain raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libcapaz_amd.so")

# Field names in ptr_Arrays member order (defines.h:65-76) == capaz_field_id.
FIELDS = (
    "S_n", "S_w", "P_w", "P_n", "ro_w", "ro_n", "ux_w", "uy_w", "uz_w", "ux_n", "uy_n", "uz_n",
    "Xi_w", "Xi_n", "roS_w", "roS_w_old", "roS_n", "roS_n_old", "m", "K",
    "S_g", "P_g", "ro_g", "ux_g", "uy_g", "uz_g", "Xi_g", "roS_g", "roS_g_old",
    "T", "H_w", "H_n", "H_g", "H_r", "E", "E_new",
)
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}

# The C entry points of include/capaz.h with their ctypes signatures.
EXPORTS = {
    "capaz_params_init": (None, ["params_p"]),
    "capaz_abi_version": (ctypes.c_int, []),
    "capaz_create": (ctypes.c_int, ["params_p", "ctx_pp"]),
    "capaz_destroy": (ctypes.c_int, ["ctx"]),
    "capaz_last_error": (ctypes.c_char_p, ["ctx"]),
    "capaz_cells": (ctypes.c_size_t, ["ctx"]),
    "capaz_has_field": (ctypes.c_int, ["ctx", ctypes.c_int]),
    "capaz_set_field": (ctypes.c_int, ["ctx", ctypes.c_int, "dbl_p", ctypes.c_size_t]),
    "capaz_get_field": (ctypes.c_int, ["ctx", ctypes.c_int, "dbl_p", ctypes.c_size_t]),
    "capaz_field_device_ptr": (ctypes.c_int, ["ctx", ctypes.c_int, "dbl_pp"]),
    "capaz_set_field_device": (ctypes.c_int, ["ctx", ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t]),
    "capaz_S": (ctypes.c_int, ["ctx"]),
    "capaz_P_Xi_ro": (ctypes.c_int, ["ctx"]),
    "capaz_u": (ctypes.c_int, ["ctx"]),
    "capaz_roS": (ctypes.c_int, ["ctx", ctypes.c_double]),
    "capaz_energy_HE": (ctypes.c_int, ["ctx"]),
    "capaz_energy_E": (ctypes.c_int, ["ctx"]),
    "capaz_newton": (ctypes.c_int, ["ctx"]),
    "capaz_boundary": (ctypes.c_int, ["ctx"]),
    "capaz_step_sequence": (ctypes.c_int, ["ctx", ctypes.c_double]),
    "capaz_step": (ctypes.c_int, ["ctx", ctypes.c_long, ctypes.c_long]),
    "capaz_tune": (ctypes.c_int, ["ctx", ctypes.c_char_p, ctypes.c_int]),
    "capaz_get_tuning": (ctypes.c_int, ["ctx", ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]),
    "capaz_kernel_time": (ctypes.c_int, ["ctx", ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long)]),
    "capaz_check_finite": (ctypes.c_int, ["ctx", ctypes.POINTER(ctypes.c_int64)]),
    "capaz_check_finite_noalloc": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int64,
                                                  ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "capaz_sync": (ctypes.c_int, ["ctx"]),
    "capaz_stream": (ctypes.c_int, ["ctx", ctypes.POINTER(ctypes.c_void_p)]),
    "capaz_load_face": (ctypes.c_int, ["ctx", ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    "capaz_save_face": (ctypes.c_int, ["ctx", ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    "capaz_load_faces": (ctypes.c_int, ["ctx", ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int,
                                        ctypes.c_void_p]),
    "capaz_save_faces": (ctypes.c_int, ["ctx", ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int,
                                        ctypes.c_void_p]),
    "capaz_set_halo_pack": (ctypes.c_int, ["ctx", ctypes.c_void_p]),
    "capaz_ghost_planes": (ctypes.c_int, ["ctx", ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_int)]),
    "capaz_save_state": (ctypes.c_int, ["ctx", ctypes.c_char_p, ctypes.c_long]),
    "capaz_write_plot_rows": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.c_long,
                                             ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]),
    "capaz_load_state": (ctypes.c_int, ["ctx", ctypes.c_char_p, ctypes.POINTER(ctypes.c_long)]),
    "capaz_objective": (ctypes.c_int, ["ctx", ctypes.c_long, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "capaz_adjoint": (ctypes.c_int, ["ctx", ctypes.c_long, ctypes.c_void_p, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_double), "dbl_p", "dbl_p"]),
    "capaz_tangent": (ctypes.c_int, ["ctx", ctypes.c_long, ctypes.c_void_p, "dbl_p", "dbl_p",
                                     ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "capaz_batch_create": (ctypes.c_int, ["params_p", ctypes.c_int, ctypes.POINTER(ctypes.c_int), "ctx_pp"]),
    "capaz_batch_destroy": (ctypes.c_int, ["ctx"]),
    "capaz_batch_last_error": (ctypes.c_char_p, ["ctx"]),
    "capaz_batch_size": (ctypes.c_int, ["ctx"]),
    "capaz_batch_domain": (ctypes.c_int, ["ctx", ctypes.c_int, "ctx_pp"]),
    "capaz_batch_step": (ctypes.c_int, ["ctx", ctypes.c_long, ctypes.c_long]),
    "capaz_batch_tune": (ctypes.c_int, ["ctx", ctypes.c_char_p, ctypes.c_int]),
    "capaz_batch_get_tuning": (ctypes.c_int, ["ctx", ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]),
    "capaz_batch_autotune": (ctypes.c_int, ["ctx", ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    "capaz_batch_sync": (ctypes.c_int, ["ctx"]),
    "capaz_batch_stream": (ctypes.c_int, ["ctx", ctypes.POINTER(ctypes.c_void_p)]),
}


class Params(ctypes.Structure):
    """capaz_params (include/capaz.h)."""
    _fields_ = [
        ("struct_size", ctypes.c_uint), ("model", ctypes.c_int), ("energy", ctypes.c_int),
        ("closure", ctypes.c_int), ("Nx", ctypes.c_int), ("Ny", ctypes.c_int), ("Nz", ctypes.c_int),
        ("source", ctypes.c_int), ("newton_iterations", ctypes.c_int), ("device", ctypes.c_int),
        ("hx", ctypes.c_double), ("hy", ctypes.c_double), ("hz", ctypes.c_double), ("dt", ctypes.c_double),
        ("P_atm", ctypes.c_double), ("g_const", ctypes.c_double),
        ("mu_w", ctypes.c_double), ("mu_n", ctypes.c_double), ("mu_g", ctypes.c_double),
        ("ro0_w", ctypes.c_double), ("ro0_n", ctypes.c_double), ("ro0_g", ctypes.c_double),
        ("beta_w", ctypes.c_double), ("beta_n", ctypes.c_double), ("lambda_", ctypes.c_double),
        ("S_wr", ctypes.c_double), ("S_nr", ctypes.c_double), ("S_gr", ctypes.c_double),
        ("P_d", ctypes.c_double), ("S_w_gr", ctypes.c_double), ("S_n_gr", ctypes.c_double),
        ("corey_n", ctypes.c_double),
        ("P_bottom", ctypes.c_double), ("energy_closure", ctypes.c_int),
        ("lambda_eff", ctypes.c_double), ("c_r", ctypes.c_double), ("ro_r", ctypes.c_double),
        ("T_top", ctypes.c_double), ("T_bottom", ctypes.c_double),
        ("mu_T_a", ctypes.c_double), ("mu_T_ref", ctypes.c_double),
    ]


_TYPES = {
    "params_p": ctypes.POINTER(Params),
    "ctx": ctypes.c_void_p,
    "ctx_pp": ctypes.POINTER(ctypes.c_void_p),
    "dbl_p": ctypes.POINTER(ctypes.c_double),
    "dbl_pp": ctypes.POINTER(ctypes.POINTER(ctypes.c_double)),
}

_lib = None


class CapazError(RuntimeError):
    """A C ABI call failed; ``code`` is its capaz_status."""

    def __init__(self, msg: str, code: int = 0):
        super().__init__(msg)
        self.code = code


class NonFiniteState(CapazError):
    """CAPAZ_ERR_NONFINITE: the state holds a NaN or an infinity at cell ``first_bad``."""

    def __init__(self, msg: str, first_bad: int):
        super().__init__(msg, CAPAZ_ERR_NONFINITE)
        self.first_bad = first_bad


CAPAZ_ERR_NONFINITE = -5


def library_path() -> str:
    return LIB_PATH


def load() -> ctypes.CDLL:
    """Loads libcapaz_amd.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CapazError(f"{LIB_PATH} is missing: build it with `python -m capaz_amd.build` "
                             "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = [_TYPES.get(a, a) if isinstance(a, str) else a for a in args]
        _lib = L
    return _lib


def params_from_keys(keys: dict, device: int = 0) -> Params:
    """capaz_params from defines.ini keys (main.cpp:1327-1626) plus MODEL,
    ENERGY, CLOSURE ("reference"/"corey"), COREY_N, P_BOTTOM and the
    configurations' energy closure (ENERGY_CLOSURE "config", LAMBDA_EFF, C_R,
    RO_R, T_TOP, T_BOTTOM, MU_T_A, MU_T_REF)."""
    p = Params()
    p.struct_size = ctypes.sizeof(Params)
    g = lambda k, d=0.0: float(keys.get(k, d))
    p.model = int(keys.get("MODEL", 3))
    p.energy = int(keys.get("ENERGY", 0))
    p.closure = 1 if keys.get("CLOSURE", "reference") == "corey" else 0
    p.Nx, p.Ny, p.Nz = int(keys["NX"]), int(keys["NY"]), int(keys["NZ"])
    p.source = int(keys.get("SOURCE", 0))
    p.newton_iterations = int(keys.get("ITERATIONS", 1))
    p.device = device
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


class Domain:
    """One single-GPU domain: a capaz_ctx and every field of its model."""

    def __init__(self, keys: dict, fields: dict[str, np.ndarray] | None = None, device: int = 0,
                 _ctx: ctypes.c_void_p | None = None):
        L = load()
        self.keys = dict(keys)
        self.params = params_from_keys(keys, device)
        self.shape = (self.params.Nz, self.params.Ny, self.params.Nx)
        self.n = int(np.prod(self.shape))
        self._owned = _ctx is None
        if _ctx is None:
            ctx = ctypes.c_void_p()
            rc = L.capaz_create(ctypes.byref(self.params), ctypes.byref(ctx))
            if rc != 0:
                raise CapazError(f"capaz_create failed ({rc}): {L.capaz_last_error(None).decode()}")
        else:
            ctx = _ctx  # a batch member: the batch owns the context
        self._ctx = ctx
        self.names = [f for f in FIELDS if L.capaz_has_field(ctx, FIELD_ID[f])]
        if fields:
            for name, v in fields.items():
                self.set_field(name, v)

    # -- errors -------------------------------------------------------------
    def _check(self, rc: int, what: str) -> None:
        if rc != 0:
            msg = load().capaz_last_error(self._ctx).decode()
            if rc == CAPAZ_ERR_NONFINITE:
                cell = int(msg.split("cell ", 1)[1].split(" ", 1)[0]) if "cell " in msg else -1
                raise NonFiniteState(f"{what} failed ({rc}): {msg}", cell)
            raise CapazError(f"{what} failed ({rc}): {msg}", rc)

    # -- data (load_data_to_device / load_data_to_host) ---------------------
    def set_field(self, name: str, value) -> None:
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(value, dtype=np.float64), self.shape))
        self._check(load().capaz_set_field(self._ctx, FIELD_ID[name],
                                           a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), self.n),
                    f"capaz_set_field({name})")

    def field(self, name: str) -> np.ndarray:
        out = np.empty(self.shape)
        self._check(load().capaz_get_field(self._ctx, FIELD_ID[name],
                                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), self.n),
                    f"capaz_get_field({name})")
        return out

    def fields(self) -> dict[str, np.ndarray]:
        return {n: self.field(n) for n in self.names}

    def device_ptr(self, name: str) -> int:
        p = ctypes.POINTER(ctypes.c_double)()
        self._check(load().capaz_field_device_ptr(self._ctx, FIELD_ID[name], ctypes.byref(p)),
                    "capaz_field_device_ptr")
        return ctypes.cast(p, ctypes.c_void_p).value or 0

    def stream(self) -> int:
        s = ctypes.c_void_p()
        self._check(load().capaz_stream(self._ctx, ctypes.byref(s)), "capaz_stream")
        return s.value or 0

    def set_halo_pack(self, address: int) -> None:
        """Device buffer (address; 0 turns it off) that every step fills with
        the z exchange planes of its new state, capaz_load_faces' layout."""
        self._check(load().capaz_set_halo_pack(self._ctx, ctypes.c_void_p(address or None)),
                    "capaz_set_halo_pack")

    def ghost_planes(self) -> list[int]:
        """Device addresses of the z ghost planes of the exchange fields in the
        arrays the next step reads ([field][side], capaz_ghost_planes)."""
        out, n = (ctypes.c_void_p * 10)(), ctypes.c_int()
        self._check(load().capaz_ghost_planes(self._ctx, out, 10, ctypes.byref(n)), "capaz_ghost_planes")
        return [out[q] or 0 for q in range(n.value)]

    # -- entry points (defines.h:186-193, 220-222) ----------------------------
    def S_calculation(self):
        self._check(load().capaz_S(self._ctx), "capaz_S")

    def ro_P_Xi_calculation(self):
        self._check(load().capaz_P_Xi_ro(self._ctx), "capaz_P_Xi_ro")

    def u_calculation(self):
        self._check(load().capaz_u(self._ctx), "capaz_u")

    def roS_calculation(self, t: float = 0.0):
        self._check(load().capaz_roS(self._ctx, t), "capaz_roS")

    def H_E_current_calculation(self):
        self._check(load().capaz_energy_HE(self._ctx), "capaz_energy_HE")

    def E_calculation(self):
        self._check(load().capaz_energy_E(self._ctx), "capaz_energy_E")

    def P_S_calculation(self):
        self._check(load().capaz_newton(self._ctx), "capaz_newton")

    def boundary_conditions(self):
        self._check(load().capaz_boundary(self._ctx), "capaz_boundary")

    def time_step_function(self, t: float):
        self._check(load().capaz_step_sequence(self._ctx, t), "capaz_step_sequence")

    def call(self, entry: str, t: float = 0.0) -> None:
        fn = getattr(self, entry)
        fn(t) if entry in ("roS_calculation", "time_step_function") else fn()

    # -- fused path ----------------------------------------------------------
    def run(self, nsteps: int, first_step: int = 1) -> None:
        """nsteps fused time steps (capaz_step)."""
        self._check(load().capaz_step(self._ctx, first_step, nsteps), "capaz_step")

    def tune(self, **kw) -> None:
        """Launch shape of the fused step (capaz_tune): fused_kernel, tile_y,
        kchunk, adjoint_kernel, adjoint_streams, energy_kernel, energy_newton; timing=1 records events around
        the main kernel of each step; check_every=K runs the non-finite
        sentinel after every K-th step."""
        for k, v in kw.items():
            self._check(load().capaz_tune(self._ctx, k.encode(), int(v)), f"capaz_tune({k}={v})")

    def tuning(self) -> dict:
        out = {}
        for k in ("check_every", "fused_kernel", "tile_y", "kchunk", "adjoint_kernel", "adjoint_streams", "energy_kernel",
                  "energy_newton", "timing", "adjoint_segment"):
            v = ctypes.c_int()
            self._check(load().capaz_get_tuning(self._ctx, k.encode(), ctypes.byref(v)), "capaz_get_tuning")
            out[k] = v.value
        return out

    def save_state(self, path: str, step: int = 0) -> None:
        """capaz_save_state: the fields the next step reads, to `path`."""
        self._check(load().capaz_save_state(self._ctx, os.fsencode(path), step), "capaz_save_state")

    def load_state(self, path: str) -> int:
        """capaz_load_state: restores a saved state; returns its step number."""
        st = ctypes.c_long()
        self._check(load().capaz_load_state(self._ctx, os.fsencode(path), ctypes.byref(st)), "capaz_load_state")
        return st.value

    def objective(self, nsteps: int, obs: dict) -> float:
        """capaz_objective: J after nsteps fused steps from the current state
        (obs: cells, steps, S_w, S_n as in include/capaz.h)."""
        o, keep = make_observations(obs["cells"], obs["steps"], obs["S_w"], obs["S_n"])
        J = ctypes.c_double()
        self._check(load().capaz_objective(self._ctx, nsteps, ctypes.byref(o), ctypes.byref(J)), "capaz_objective")
        return J.value

    def adjoint(self, nsteps: int, obs: dict, checkpoint_every: int = 0):
        """capaz_adjoint: (J, dJ/d lnK, dJ/dm) after nsteps fused steps from the
        current state.  checkpoint_every: > 0 the segment length, < 0
        round(sqrt(nsteps)), 0 the whole trajectory when it fits in a quarter
        of the device memory, else round(sqrt(nsteps)) (tuning()
        ["adjoint_segment"] reports the one used)."""
        o, keep = make_observations(obs["cells"], obs["steps"], obs["S_w"], obs["S_n"])
        J = ctypes.c_double()
        gK, gm = np.empty(self.shape), np.empty(self.shape)
        self._check(load().capaz_adjoint(self._ctx, nsteps, ctypes.byref(o), checkpoint_every, ctypes.byref(J),
                                         gK.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                         gm.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "capaz_adjoint")
        return J.value, gK, gm

    def tangent(self, nsteps: int, obs: dict, dlnK, dm):
        """capaz_tangent: (J, dJ along (dlnK, dm)) after nsteps fused steps from the current state."""
        o, keep = make_observations(obs["cells"], obs["steps"], obs["S_w"], obs["S_n"])
        arr = lambda v: None if v is None else np.ascontiguousarray(np.broadcast_to(v, self.shape), dtype=np.float64)
        vK, vm = arr(dlnK), arr(dm)
        ptr = lambda v: v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if v is not None else None
        J, dJ = ctypes.c_double(), ctypes.c_double()
        self._check(load().capaz_tangent(self._ctx, nsteps, ctypes.byref(o), ptr(vK), ptr(vm), ctypes.byref(J),
                                         ctypes.byref(dJ)), "capaz_tangent")
        return J.value, dJ.value

    def kernel_time(self) -> tuple[float, int]:
        """(total ms, launches) of the fused step's main kernel since timing was set."""
        ms, n = ctypes.c_double(), ctypes.c_long()
        self._check(load().capaz_kernel_time(self._ctx, ctypes.byref(ms), ctypes.byref(n)), "capaz_kernel_time")
        return ms.value, n.value

    def check_finite(self) -> int:
        """capaz_check_finite: -1 if the primary state is finite, else raises
        NonFiniteState with the first bad cell."""
        first = ctypes.c_int64(-1)
        self._check(load().capaz_check_finite(self._ctx, ctypes.byref(first)), "capaz_check_finite")
        return first.value

    def sync(self) -> None:
        self._check(load().capaz_sync(self._ctx), "capaz_sync")

    def close(self) -> None:
        if getattr(self, "_ctx", None):
            if self._owned:
                load().capaz_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Observations(ctypes.Structure):
    """capaz_observations (include/capaz.h)."""
    _fields_ = [("ncells", ctypes.c_size_t), ("cells", ctypes.POINTER(ctypes.c_int64)),
                ("nsamples", ctypes.c_int), ("steps", ctypes.POINTER(ctypes.c_long)),
                ("S_w", ctypes.POINTER(ctypes.c_double)), ("S_n", ctypes.POINTER(ctypes.c_double))]


def make_observations(cells, steps, S_w, S_n):
    """An Observations struct and the arrays it points into (keep both alive)."""
    keep = [np.ascontiguousarray(cells, dtype=np.int64), np.ascontiguousarray(steps, dtype=np.int64),
            np.ascontiguousarray(S_w, dtype=np.float64), np.ascontiguousarray(S_n, dtype=np.float64)]
    o = Observations()
    o.ncells = keep[0].size
    o.cells = keep[0].ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    o.nsamples = keep[1].size
    o.steps = keep[1].ctypes.data_as(ctypes.POINTER(ctypes.c_long))
    o.S_w = keep[2].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    o.S_n = keep[3].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    return o, keep


class Batch:
    """An ensemble of realizations advanced together (capaz_batch): the
    physics of `keys`, one member per dict of `members` with its grid under
    "dims" = (Nx, Ny, Nz) and its initial fields under their names.
    ``domains[r]`` is realization r as a Domain (owned by the batch)."""

    def __init__(self, keys: dict, members: list[dict], device: int = 0):
        L = load()
        if not members:
            raise CapazError("a batch needs at least one realization")
        dims = [int(d) for m in members for d in m["dims"]]
        nx, ny, nz = dims[:3]
        self.keys = dict(keys)
        self.params = params_from_keys(dict(keys, NX=nx, NY=ny, NZ=nz), device)
        arr = (ctypes.c_int * len(dims))(*dims)
        b = ctypes.c_void_p()
        rc = L.capaz_batch_create(ctypes.byref(self.params), len(members), arr, ctypes.byref(b))
        if rc != 0:
            raise CapazError(f"capaz_batch_create failed ({rc}): {L.capaz_batch_last_error(None).decode()}")
        self._b = b
        self.domains = []
        for r, m in enumerate(members):
            c = ctypes.c_void_p()
            self._check(L.capaz_batch_domain(b, r, ctypes.byref(c)), "capaz_batch_domain")
            x, y, z = m["dims"]
            d = Domain(dict(keys, NX=x, NY=y, NZ=z), None, device, _ctx=c)
            for name, v in m.items():
                if name != "dims":
                    d.set_field(name, v)
            self.domains.append(d)

    def _check(self, rc: int, what: str) -> None:
        if rc != 0:
            raise CapazError(f"{what} failed ({rc}): {load().capaz_batch_last_error(self._b).decode()}")

    def __len__(self) -> int:
        return len(self.domains)

    def run(self, nsteps: int, first_step: int = 1) -> None:
        """capaz_batch_step: nsteps fused steps of every realization."""
        self._check(load().capaz_batch_step(self._b, first_step, nsteps), "capaz_batch_step")

    def tune(self, **kw) -> None:
        """capaz_batch_tune: tile_x, tile_y, kchunk, persistent, graph."""
        for k, v in kw.items():
            self._check(load().capaz_batch_tune(self._b, k.encode(), int(v)), f"capaz_batch_tune({k}={v})")

    def tuning(self) -> dict:
        out = {}
        for k in ("tile_x", "tile_y", "kchunk", "persistent", "graph"):
            v = ctypes.c_int()
            self._check(load().capaz_batch_get_tuning(self._b, k.encode(), ctypes.byref(v)), "capaz_batch_get_tuning")
            out[k] = v.value
        return out

    def autotune(self, steps: int = 4) -> float:
        """capaz_batch_autotune: keeps the fastest launch shape; returns its ms per step."""
        ms = ctypes.c_double()
        self._check(load().capaz_batch_autotune(self._b, steps, ctypes.byref(ms)), "capaz_batch_autotune")
        return ms.value

    def stream(self) -> int:
        s = ctypes.c_void_p()
        self._check(load().capaz_batch_stream(self._b, ctypes.byref(s)), "capaz_batch_stream")
        return s.value or 0

    def sync(self) -> None:
        self._check(load().capaz_batch_sync(self._b), "capaz_batch_sync")

    def close(self) -> None:
        if getattr(self, "_b", None):
            for d in self.domains:
                d.close()
            load().capaz_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
