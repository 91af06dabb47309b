"""Seeded synthetic inputs of the benchmark workloads.

Each builder returns ``(keys, fields)``: a deck of defines.ini keys
(main.cpp:1327-1626) plus the backend's own keys (MODEL, ENERGY, CLOSURE,
COREY_N, P_BOTTOM and the energy closure's), and the initial primary fields as (Nz, Ny, Nx) float64 arrays, x
fastest.  Random fields come from numpy's PCG64 (two-phase Buckley-Leverett
seed 1000, three-phase injection seed 1001, non-isothermal two-phase seed
1002).  "Top" is j = 0 (gravity along
+j, cpu.cpp:392).

Parameter substitutions (the decks as specified are unstable for this
explicit scheme; the stable values used are):
  * three-phase (128^3 and the 256^3 headline): dt = 1e-5 s instead of
    1e-4 s.  At h = 0.01 m and beta = 4.5e-10 1/Pa the scheme diverges within
    20 steps at dt = 1e-4 (24^3, oracle and reference alike); the threshold
    measured at 24^3 is ~6e-5 s and lognormal K tails at 256^3 tighten it.
  * two-phase Buckley-Leverett (40^3): beta_w = beta_n = 1e-6 1/Pa instead of 4.5e-10.
    Without gas the only compressibility is beta, and the pressure-diffusion
    limit h^2*phi*beta*mu/(6*K*k) is ~2.5e-7 s at 4.5e-10, 400x below
    dt = 1e-4; at 1e-6 it is ~5.6e-4 s.
"""
from __future__ import annotations

import numpy as np

P_ATM = 1.0e5
K_MEAN = 6.64e-11


def _porosity_noise(rng, shape):
    return np.clip(0.4 * (1.0 + 0.05 * rng.standard_normal(shape)), 0.3, 0.5)


def buckley_leverett_2ph(n: int = 40):
    """Two-phase water/NAPL, n^3, Corey-2 (S_wr 0.1, S_nr 0.05), scalar K,
    S_w0 = 0.15, top face injects S_w = 0.9 at P = 1.2e5 Pa, bottom face
    held at P = 1e5 Pa, side faces no-flow.

    The reference's two-phase boundary sets the top pressure to P_atm
    (two-phase.cpp:200-211), so the top pressure is carried by
    P_ATM = 1.2e5; the bottom takes the backend's fixed bottom pressure
    (P_BOTTOM) instead of the reference's hydrostatic increment."""
    rng = np.random.default_rng(1000)
    shape = (n, n, n)
    keys = dict(MODEL=2, CLOSURE="corey", COREY_N=2, NX=n, NY=n, NZ=n, HX=0.01, HY=0.01, HZ=0.01, DT=1e-4,
                P_ATM=1.2e5, P_BOTTOM=1.0e5, G_CONST=9.807, MU_W=1e-3, MU_N=5e-3, RO_W=1000.0, RO_N=850.0,
                BETA_W=1e-6, BETA_N=1e-6, S_WR_0=0.1, S_NR_0=0.05, SOURCE=n, S_N_GR=0.1,
                ITERATIONS=4, K_0=K_MEAN)
    P = np.full(shape, keys["P_ATM"])
    fields = dict(P_w=P, S_n=np.full(shape, 0.85), m=_porosity_noise(rng, shape), K=np.full(shape, K_MEAN))
    return keys, fields


def nonisothermal_2ph(nx: int = 256, ny: int = 256, nz: int = 128, seed: int = 1002):
    """Non-isothermal two-phase water/NAPL on an nx x ny x nz grid (256 x 256 x
    128 as specified), 1000 steps: the Buckley-Leverett physics and initial
    state of buckley_leverett_2ph (Corey-2, S_wr 0.1, S_nr 0.05, scalar K
    6.64e-11, porosity 0.4(1 + 0.05 N(0,1)) clipped to [0.3, 0.5], S_w0 0.15,
    the whole top face injecting S_w 0.9 at P 1.2e5 Pa, bottom at 1e5 Pa,
    no-flow sides, beta 1e-6 1/Pa, dt 1e-4 s, 4 Newton iterations) with the
    energy equation: rock heat capacity 800 J/(kg K) and density 2000 kg/m^3,
    thermal conductivity 1.5 W/(m K), T0 = 293 K, inflow at 353 K on the top
    face (the other faces insulated), mu_a(T) = mu_a exp(-0.02 (T - 293)).
    The fluids' enthalpies, heat capacities and thermal expansion are
    energy.cpp's (T_0 = 273 K their base temperature)."""
    keys, _ = buckley_leverett_2ph(4)
    keys = dict(keys, NX=nx, NY=ny, NZ=nz, SOURCE=max(nx, nz), ENERGY=1, ENERGY_CLOSURE="config",
                LAMBDA_EFF=1.5, C_R=800.0, RO_R=2000.0, T_TOP=353.0, T_BOTTOM=0.0, MU_T_A=0.02,
                MU_T_REF=293.0)
    rng = np.random.default_rng(seed)
    shape = (nz, ny, nx)
    fields = dict(P_w=np.full(shape, keys["P_ATM"]), S_n=np.full(shape, 0.85), T=np.full(shape, 293.0),
                  m=_porosity_noise(rng, shape), K=np.full(shape, K_MEAN))
    return keys, fields


def three_phase_injection(n: int = 128, seed: int = 1001):
    """Three-phase water/NAPL/gas on an n^3 grid (128^3 as specified; the
    headline measures the same physics at 256^3): 4 Newton iterations,
    ideal-gas density with ro0_g = 1.2, mu_g = 1.8e-5, Corey-2 relperms
    (S_wr 0.1, S_nr 0.05, S_gr 0), S0 = (0.2, 0.7, 0.1), the top face injects
    (0.6, 0.1, 0.3), porosity 0.4(1 + 0.05 N(0,1)) clipped to [0.3, 0.5],
    K lognormal with geometric mean 6.64e-11 m^2 and ln-sigma 0.5 (white
    noise per cell).  h = 0.01 m, beta = 4.5e-10 1/Pa, mu_w 1e-3, mu_n 5e-3,
    ro0 1000 / 850 as in the two-phase workload; P = P_atm everywhere initially (the
    reference's three-phase data_initialization, three-phase.cpp:470-472)."""
    rng = np.random.default_rng(seed)
    shape = (n, n, n)
    keys = dict(MODEL=3, CLOSURE="corey", COREY_N=2, NX=n, NY=n, NZ=n, HX=0.01, HY=0.01, HZ=0.01, DT=1e-5,
                P_ATM=P_ATM, G_CONST=9.807, MU_W=1e-3, MU_N=5e-3, MU_G=1.8e-5, RO_W=1000.0, RO_N=850.0,
                RO_G=1.2, BETA_W=4.5e-10, BETA_N=4.5e-10, S_WR_0=0.1, S_NR_0=0.05, S_GR_0=0.0,
                SOURCE=1, S_W_GR=0.6, S_N_GR=0.1, ITERATIONS=4, K_0=K_MEAN)
    m = _porosity_noise(rng, shape)
    K = np.exp(np.log(K_MEAN) + 0.5 * rng.standard_normal(shape))
    P = np.full(shape, P_ATM)
    fields = dict(P_w=P, P_n=P.copy(), P_g=P.copy(), S_w=np.full(shape, 0.2), S_n=np.full(shape, 0.7), m=m, K=K)
    return keys, fields


def headline(n: int = 256):
    """The benchmark workload: the three-phase injection physics at 256^3 (seed 1001)."""
    return three_phase_injection(n, seed=1001)


def _correlated_lnK(rng, n: int, corr_cells: float, ln_sigma: float) -> np.ndarray:
    """ln K on an n^3 grid: white noise smoothed by a Gaussian filter of
    `corr_cells` cells standard deviation, rescaled to mean ln(K_MEAN) and
    standard deviation `ln_sigma`."""
    from scipy.ndimage import gaussian_filter

    w = gaussian_filter(rng.standard_normal((n, n, n)), sigma=corr_cells, mode="reflect")
    w = (w - w.mean()) / w.std()
    return np.log(K_MEAN) + ln_sigma * w


ENSEMBLE_EDGES = (32, 48, 64, 80, 96)


def ensemble(count: int = 64, seed: int = 1003, dt: float = 1e-6):
    """Ensemble of two-phase Buckley-Leverett realizations: each cube edge
    drawn uniformly from {32, 48, 64, 80, 96}; K lognormal with geometric mean
    6.64e-11 m^2 and ln-sigma 1.0, correlated by a Gaussian filter of 4 cells;
    porosity 0.4 (K / K_mean)^0.1 clipped to [0.25, 0.5]; physics and
    initial state of buckley_leverett_2ph, except dt = 1e-6 s: the K tails
    (up to ~500 K_mean) put the explicit pressure-diffusion limit near 1e-6 s
    (dt = 1e-5 diverges within 100 steps on a 96^3 realization, 1e-6 runs
    500 steps on the realization of largest K).  Returns (keys, [fields, ...]):
    keys of the common physics (grid keys are per realization, in each
    fields dict under "dims" as (Nx, Ny, Nz))."""
    rng = np.random.default_rng(seed)
    edges = rng.choice(np.array(ENSEMBLE_EDGES), size=count)
    keys, _ = buckley_leverett_2ph(int(edges[0]))
    # SOURCE covers the whole top face of the largest realization
    keys = dict(keys, DT=dt, SOURCE=max(ENSEMBLE_EDGES))
    out = []
    for n in edges:
        n = int(n)
        lnK = _correlated_lnK(rng, n, 4.0, 1.0)
        K = np.exp(lnK)
        m = np.clip(0.4 * (K / K_MEAN) ** 0.1, 0.25, 0.5)
        shape = (n, n, n)
        out.append(dict(dims=(n, n, n), P_w=np.full(shape, keys["P_ATM"]), S_n=np.full(shape, 0.85), m=m, K=K))
    return keys, out


def history_matching(n: int = 96, nsteps: int = 400, seed: int = 1004, columns: int = 16, every: int = 50,
                     phases: int = 3):
    """Permeability inversion on an n^3 three-phase grid (physics of
    three_phase_injection; phases=2: of buckley_leverett_2ph, with the same
    dt): a seeded "true" ln K (ln-sigma 1.0 around
    ln 6.64e-11, 6-cell Gaussian correlation), the trial ln K = its mean
    everywhere, porosity 0.4(1 + 0.05 N(0,1)) clipped to [0.3, 0.5] for both,
    and `columns` seeded observation columns (i, k) whose interior cells
    (j = 1 .. n-2) are observed every `every` steps up to nsteps.

    dt = 1e-6 s: the true field's K tails reach ~100 K_mean, and dt = 1e-5
    diverges within 50 steps at 96^3 (dt = 1e-6 runs the 400 steps).

    Returns (keys, state, lnK_true, lnK_trial, obs) with state the initial
    primary fields and porosity, obs = {cells, steps} (the observed values
    come from a forward run with lnK_true)."""
    keys, st = three_phase_injection(n, seed=seed) if phases == 3 else buckley_leverett_2ph(n)
    keys = dict(keys, DT=1e-6)
    rng = np.random.default_rng(seed + 1)
    lnK_true = _correlated_lnK(rng, n, 6.0, 1.0)
    lnK_trial = np.full((n, n, n), np.log(K_MEAN))
    st.pop("K")
    ik = rng.choice((n - 2) * (n - 2), size=columns, replace=False)
    i_c, k_c = 1 + ik % (n - 2), 1 + ik // (n - 2)
    j = np.arange(1, n - 1)
    cells = (i_c[:, None] + j[None, :] * n + k_c[:, None] * n * n).ravel().astype(np.int64)
    steps = np.arange(every, nsteps + 1, every, dtype=np.int64)
    return keys, st, lnK_true, lnK_trial, {"cells": cells, "steps": steps}


# The reference's Three-phase/defines.ini deck (values only).
REF_3PH_DECK = dict(
    MODEL=3, HX=5e-2, HY=5e-2, HZ=5e-2, DT=1e-3, TAU=2e-3, S_W_GR=0.4, S_N_GR=0.3, S_W_INIT=0.4,
    S_N_INIT=0.3, SOURCE=0, L=1e-7, C_N=1000, C_W=1500, C_G=330, BETA_W=4.4e-7, BETA_N=1.0e-6,
    BETA_G=9.2e-6, MU_W=1e-3, MU_N=1e-3, MU_G=1.84e-5, RO_W=1000, RO_N=850, RO_G=1.4, P_ATM=1.0e5,
    G_CONST=9.807, ITERATIONS=7, LAMBDA_0=3.25, LAMBDA_1=3.25, P_D_NW_0=3125.0, P_D_NW_1=3125.0,
    P_D_GN_0=1042.0, P_D_GN_1=1042.0, K_0=6.64e-11, K_1=7.15e-12, M_0=0.4, M_1=0.39, S_WR_0=0.0,
    S_WR_1=0.0, S_NR_0=0.0, S_NR_1=0.0, S_GR_0=0.0, S_GR_1=0.0, Q=0.15e-1,
)


def energy_3ph(nx: int = 256, ny: int = 256, nz: int = 128, seed: int = 1002, closure: str = "reference"):
    """The reference's three-phase model with the energy equation
    (-DTHREE_PHASE -DENERGY; its Three-phase/defines.ini deck, reference
    closures, 3 Newton iterations, dt = 1e-4 s) on an nx x ny x nz grid.
    Initial state: hydrostatic water pressure, S = (0.4, 0.3, 0.3), T = 285 K
    (three-phase.cpp:470-487), porosity 0.4(1 + 0.05 N(0,1)) clipped to
    [0.3, 0.5], uniform K_0.  The reference has no two-phase energy model, so
    this is its own energy model on the grid of the non-isothermal workload.
    closure="corey": the relative permeabilities of the two-phase workload
    instead (Corey-2, S_wr 0.1, S_nr 0.05), the rest unchanged."""
    rng = np.random.default_rng(seed)
    keys = dict(REF_3PH_DECK, ENERGY=1, ITERATIONS=3, DT=1e-4, NX=nx, NY=ny, NZ=nz)
    if closure == "corey":
        keys.update(CLOSURE="corey", COREY_N=2, S_WR_0=0.1, S_NR_0=0.05)
    shape = (nz, ny, nx)
    j = np.arange(ny).reshape(1, ny, 1)
    P = np.broadcast_to(keys["P_ATM"] + 1000.0 * keys["G_CONST"] * keys["HY"] * j, shape).copy()
    fields = dict(P_w=P, P_n=P.copy(), P_g=P.copy(), S_w=np.full(shape, 0.4), S_n=np.full(shape, 0.3),
                  T=np.full(shape, 285.0), m=_porosity_noise(rng, shape), K=np.full(shape, keys["K_0"]))
    return keys, fields
