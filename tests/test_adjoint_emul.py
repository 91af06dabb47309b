"""The discrete adjoint of the fused three-phase step (capaz_adjoint.h),
built for the host, against central finite differences of the oracle's
forward step (the CPU restatement of the reference, which the fused kernels
match bit for bit in the Corey closures).

The objective is the history-matching one: half the sum of squared S_w and
S_n mismatches at observed cells and steps.  A directional derivative g.v
of the adjoint gradient must match (J(theta + eps v) - J(theta - eps v)) /
(2 eps) for random directions v in ln K and in porosity.  At eps = 1e-4 the
central difference is accurate to ~1e-8 relative here, so a wrong sign, a
dropped term or a transposed operand in any pass shows up as an O(1)
relative error.
"""
import numpy as np
import pytest

from capaz_cases import case, random_state
from emul import emul
from oracle import oracle

pytestmark = pytest.mark.skipif(not (emul.available() and oracle.available()), reason="host builds missing")


def problem(nx, ny, nz, seed, nsteps, **over):
    keys = case("3ph", nx, ny, nz, CLOSURE="corey", S_WR_0=0.1, S_NR_0=0.05, SOURCE=1, DT=1e-4, ITERATIONS=4)
    keys.update(over)
    st = random_state(keys, seed)
    rng = np.random.default_rng(seed + 1)
    shape = st["P_w"].shape
    lnK = np.log(keys["K_0"]) + 0.5 * rng.standard_normal(shape)
    st["K"] = np.exp(lnK)
    n = int(np.prod(shape))
    cells = rng.choice(n, min(12, n), replace=False)
    steps = sorted({max(1, nsteps // 2), nsteps})
    obs = {"cells": cells, "steps": steps,
           "S_w": rng.uniform(0.2, 0.5, (len(steps), cells.size)),
           "S_n": rng.uniform(0.1, 0.4, (len(steps), cells.size))}
    return keys, st, lnK, obs, rng


def oracle_J(keys, st, lnK, m, nsteps, obs):
    f = dict(st, K=np.exp(lnK), m=m)
    d = oracle.OracleDomain(keys, f)
    seen = {}
    for t in range(1, nsteps + 1):
        d.run(1, first_step=t)
        if t in obs["steps"]:
            seen[t] = (d.field("S_w").ravel().copy(), d.field("S_n").ravel().copy())
    return emul.objective(seen, obs)


@pytest.mark.parametrize("dims,seed,nsteps,corey_n", [((7, 6, 5), 11, 6, 2.0), ((9, 5, 1), 12, 4, 2.0),
                                                      ((1, 8, 6), 13, 4, 2.0), ((7, 6, 5), 14, 5, 2.6)])
def test_adjoint_gradient_matches_central_differences(dims, seed, nsteps, corey_n):
    keys, st, lnK, obs, rng = problem(*dims, seed, nsteps, COREY_N=corey_n)
    J, g_lnK, g_m = emul.gradient(keys, st, nsteps, obs)
    assert J == oracle_J(keys, st, lnK, st["m"], nsteps, obs)  # same forward, same bits
    eps = 1e-4
    for _ in range(3):
        v = rng.standard_normal(lnK.shape)
        fd = (oracle_J(keys, st, lnK + eps * v, st["m"], nsteps, obs)
              - oracle_J(keys, st, lnK - eps * v, st["m"], nsteps, obs)) / (2 * eps)
        ad = float(np.sum(g_lnK * v))
        assert abs(fd - ad) <= 1e-5 * abs(fd), f"ln K direction: fd {fd:.12e} adjoint {ad:.12e}"
        vm = 0.05 * rng.standard_normal(lnK.shape)
        fd = (oracle_J(keys, st, lnK, st["m"] + eps * vm, nsteps, obs)
              - oracle_J(keys, st, lnK, st["m"] - eps * vm, nsteps, obs)) / (2 * eps)
        ad = float(np.sum(g_m * vm))
        assert abs(fd - ad) <= 1e-4 * abs(fd), f"porosity direction: fd {fd:.12e} adjoint {ad:.12e}"


@pytest.mark.parametrize("dims,seed,nsteps", [((7, 6, 5), 31, 5), ((9, 5, 1), 32, 4)])
def test_reference_closure_gradient_matches_central_differences(dims, seed, nsteps):
    """The reference's van Genuchten k_r (three-phase.cpp:146-186): k_n
    depends on both S_w and S_n, so the chain rule's cross term dXi_n/dS_w is
    exercised; the forward step is the oracle's, glibc pow and all."""
    keys, st, lnK, obs, rng = problem(*dims, seed, nsteps, CLOSURE="reference")
    J, g_lnK, g_m = emul.gradient(keys, st, nsteps, obs)
    assert J == oracle_J(keys, st, lnK, st["m"], nsteps, obs)
    eps = 1e-4
    for _ in range(3):
        v = rng.standard_normal(lnK.shape)
        fd = (oracle_J(keys, st, lnK + eps * v, st["m"], nsteps, obs)
              - oracle_J(keys, st, lnK - eps * v, st["m"], nsteps, obs)) / (2 * eps)
        ad = float(np.sum(g_lnK * v))
        assert abs(fd - ad) <= 1e-5 * abs(fd), f"ln K direction: fd {fd:.12e} adjoint {ad:.12e}"
        vm = 0.05 * rng.standard_normal(lnK.shape)
        fd = (oracle_J(keys, st, lnK, st["m"] + eps * vm, nsteps, obs)
              - oracle_J(keys, st, lnK, st["m"] - eps * vm, nsteps, obs)) / (2 * eps)
        ad = float(np.sum(g_m * vm))
        assert abs(fd - ad) <= 1e-4 * abs(fd), f"porosity direction: fd {fd:.12e} adjoint {ad:.12e}"


def test_reference_closure_derivatives_match_central_differences():
    """vg_d's partial derivatives of the van Genuchten k_r against central
    differences of its own values (the values against the forward's k_r are
    pinned by the gradient test above)."""
    keys = case("3ph", 4, 4, 4, CLOSURE="reference")
    rng = np.random.default_rng(5)
    for _ in range(200):
        x, y = rng.uniform(0.01, 0.6), rng.uniform(0.01, 0.35)
        k, d = emul.vg_d(keys, x, y)
        h = 1e-6
        kxp, _ = emul.vg_d(keys, x + h, y)
        kxm, _ = emul.vg_d(keys, x - h, y)
        kyp, _ = emul.vg_d(keys, x, y + h)
        kym, _ = emul.vg_d(keys, x, y - h)
        fd_w = (kxp[0] - kxm[0]) / (2 * h)
        fd_nx, fd_ny = (kxp[1] - kxm[1]) / (2 * h), (kyp[1] - kym[1]) / (2 * h)
        fd_g = -(kxp[2] - kxm[2]) / (2 * h)  # dk_g/dSge = -dk_g/dSwe at fixed Sne
        for got, want in ((d[0], fd_w), (d[1], fd_nx), (d[2], fd_ny), (d[3], fd_g)):
            assert abs(got - want) <= 1e-6 * max(abs(want), 1e-3), (x, y, got, want)


@pytest.mark.parametrize("closure", ["reference", "corey"])
def test_coefficient_derivatives_match_central_differences(closure):
    """coef_d's dXi_a/dS_w and dXi_a/dS_n (the chain rule through the
    effective saturations, including the van Genuchten k_n's dependence on
    S_w) against central differences of its own Xi."""
    keys = case("3ph", 4, 4, 4, CLOSURE=closure, S_WR_0=0.1, S_NR_0=0.05)
    rng = np.random.default_rng(6)
    h = 1e-7
    for _ in range(100):
        Sw, Sn = rng.uniform(0.15, 0.55), rng.uniform(0.1, 0.3)
        K = keys["K_0"] * np.exp(rng.standard_normal())
        Xi, dSw, dSn = emul.coef_d(keys, 1e5, Sw, Sn, K)
        for ph in range(3):
            fw = (emul.coef_d(keys, 1e5, Sw + h, Sn, K)[0][ph] - emul.coef_d(keys, 1e5, Sw - h, Sn, K)[0][ph]) / (2 * h)
            fn = (emul.coef_d(keys, 1e5, Sw, Sn + h, K)[0][ph] - emul.coef_d(keys, 1e5, Sw, Sn - h, K)[0][ph]) / (2 * h)
            scale = max(abs(fw), abs(fn), 1e-3 * abs(Xi[ph]), 1e-30)
            assert abs(dSw[ph] - fw) <= 1e-5 * scale, (closure, ph, "S_w", dSw[ph], fw)
            assert abs(dSn[ph] - fn) <= 1e-5 * scale, (closure, ph, "S_n", dSn[ph], fn)


def problem2(nx, ny, nz, seed, nsteps, **over):
    # S_nr 0: the clipped k_r's kink (S_e = 1 at S_n = S_nr) stays below the states drawn
    keys = case("2ph", nx, ny, nz, CLOSURE="corey", COREY_N=2, S_NR_0=0.0, SOURCE=1, DT=1e-4, ITERATIONS=4)
    keys.update(over)
    st = random_state(keys, seed)
    rng = np.random.default_rng(seed + 1)
    shape = st["P_w"].shape
    lnK = np.log(keys["K_0"]) + 0.5 * rng.standard_normal(shape)
    st["K"] = np.exp(lnK)
    n = int(np.prod(shape))
    cells = rng.choice(n, min(12, n), replace=False)
    steps = sorted({max(1, nsteps // 2), nsteps})
    obs = {"cells": cells, "steps": steps,
           "S_w": rng.uniform(0.6, 0.9, (len(steps), cells.size)),
           "S_n": rng.uniform(0.1, 0.4, (len(steps), cells.size))}
    return keys, st, lnK, obs, rng


def oracle_J2(keys, st, lnK, m, nsteps, obs):
    # two phases: S_w = 1 - S_n of the step (the oracle's S_w array is set at
    # the start of the next step)
    d = oracle.OracleDomain(keys, dict(st, K=np.exp(lnK), m=m))
    seen = {}
    for t in range(1, nsteps + 1):
        d.run(1, first_step=t)
        if t in obs["steps"]:
            Sn = d.field("S_n").ravel().copy()
            seen[t] = (1. - Sn, Sn)
    return emul.objective(seen, obs)


@pytest.mark.parametrize("dims,seed,nsteps,over", [((7, 6, 5), 41, 5, {}), ((9, 6, 1), 42, 4, {}),
                                                   ((7, 6, 5), 43, 5, dict(P_BOTTOM=1.0e5)),
                                                   ((6, 7, 4), 44, 4, dict(COREY_N=2.6, SOURCE=0))])
def test_two_phase_gradient_matches_central_differences(dims, seed, nsteps, over):
    """The two-phase adjoint (newton2 backwards, the two-phase Border's
    transpose including the hydrostatic bottom, the two-phase transposed
    stencil) against central differences of the oracle's two-phase steps."""
    keys, st, lnK, obs, rng = problem2(*dims, seed, nsteps, **over)
    J, g_lnK, g_m = emul.gradient(keys, st, nsteps, obs)
    assert J == oracle_J2(keys, st, lnK, st["m"], nsteps, obs)
    # J's differences lose digits below this step here (the relative error of
    # the central difference grows from ~1e-6 at 4e-4 to ~1e-4 at 2.5e-5);
    # a dropped or mis-signed term is an O(1) error
    eps = 4e-4
    for _ in range(3):
        v = rng.standard_normal(lnK.shape)
        fd = (oracle_J2(keys, st, lnK + eps * v, st["m"], nsteps, obs)
              - oracle_J2(keys, st, lnK - eps * v, st["m"], nsteps, obs)) / (2 * eps)
        ad = float(np.sum(g_lnK * v))
        assert abs(fd - ad) <= 1e-4 * abs(fd), f"ln K direction: fd {fd:.12e} adjoint {ad:.12e}"
        vm = 0.05 * rng.standard_normal(lnK.shape)
        fd = (oracle_J2(keys, st, lnK, st["m"] + eps * vm, nsteps, obs)
              - oracle_J2(keys, st, lnK, st["m"] - eps * vm, nsteps, obs)) / (2 * eps)
        ad = float(np.sum(g_m * vm))
        assert abs(fd - ad) <= 1e-4 * abs(fd), f"porosity direction: fd {fd:.12e} adjoint {ad:.12e}"


@pytest.mark.parametrize("over", [dict(ITERATIONS=2), dict(ITERATIONS=2, P_BOTTOM=1.0e5), dict(ITERATIONS=1)])
def test_two_phase_step_transpose_matches_central_differences(over):
    """One two-phase step's transposed Jacobian: for random weights w on the
    new state (P_w, S_n of every cell, boundary cells included) and a random
    direction v of the old state, the adjoint step's <lambda_t, v> equals the
    central difference of <w, x_{t+1}(x_t + eps v)> through the host build of
    the step (the oracle's bits).  Few Newton iterations keep the Newton
    matrix's own derivative terms, (dM/dx) u, visible; the hydrostatic bottom
    face puts the Border transpose's density and saturation terms in."""
    keys, st, lnK, obs, rng = problem2(7, 6, 5, 51, 1, DT=1e-3, **over)
    shape = st["P_w"].shape

    def step(state):
        d = emul.EmulDomain(keys, state)
        d.fused_step()
        return d.arrays["P_w"].copy(), d.arrays["S_n"].copy()

    wP, wS = 1e-5 * rng.standard_normal(shape), rng.standard_normal(shape)
    vP, vS = 20. * rng.standard_normal(shape), 2e-3 * rng.standard_normal(shape)
    state = {n: st[n] for n in ("P_w", "S_n", "m", "K")}
    full = {n: np.zeros(shape) for n in emul.ADJ_STATE}
    full.update(state)
    lam1 = {n: np.zeros(shape) for n in emul.ADJ_STATE}
    lam1["P_w"], lam1["S_n"] = wP, wS
    grad = {"lnK": np.zeros(shape), "m": np.zeros(shape)}
    lam0 = emul.adjoint_step(keys, full, lam1, grad)
    ad = float(np.sum(lam0["P_w"] * vP) + np.sum(lam0["S_n"] * vS))
    eps = 1e-3
    Pp, Sp = step(dict(state, P_w=st["P_w"] + eps * vP, S_n=st["S_n"] + eps * vS))
    Pm, Sm = step(dict(state, P_w=st["P_w"] - eps * vP, S_n=st["S_n"] - eps * vS))
    fd = (np.sum(wP * (Pp - Pm)) + np.sum(wS * (Sp - Sm))) / (2 * eps)
    assert abs(fd - ad) <= 1e-6 * abs(fd), (fd, ad)


def test_gradient_vanishes_where_nothing_is_observed_downstream():
    # one step, observed cells only: the gradient lives on their stencils
    keys, st, lnK, obs, rng = problem(8, 8, 8, 21, 1)
    obs = dict(obs, cells=np.array([3 + 3 * 8 + 3 * 64]), steps=[1], S_w=obs["S_w"][:1, :1], S_n=obs["S_n"][:1, :1])
    J, g_lnK, g_m = emul.gradient(keys, st, 1, obs)
    nz = np.argwhere(g_lnK != 0)
    assert len(nz) > 0
    assert np.abs(nz - np.array([3, 3, 3])).max() <= 1  # the 7-point stencil of the observed cell
    assert np.count_nonzero(g_m) == 1 and g_m[3, 3, 3] != 0


def test_adjoint_at_zero_driving_forces_matches_central_differences():
    # The inversion workload starts from a uniform pressure with a uniform
    # trial K, so many x/z faces carry exactly zero driving force, where the
    # upwind flux has a kink: the adjoint takes the mean of the one-sided
    # slopes there, which is what a central difference measures.
    from capaz_amd import workloads

    n, N = 16, 24
    keys, st, lnK_true, lnK0, ob = workloads.history_matching(n=n, nsteps=N, columns=4, every=N // 4)
    steps = [int(s) for s in ob["steps"]]

    def run(lnK, m):
        d = oracle.OracleDomain(keys, dict(st, K=np.exp(lnK), m=m))
        seen = {}
        for t in range(1, N + 1):
            d.run(1, first_step=t)
            if t in steps:
                seen[t] = (d.field("S_w").ravel().copy(), d.field("S_n").ravel().copy())
        return seen

    truth = run(lnK_true, st["m"])
    obs = {"cells": ob["cells"], "steps": steps, "S_w": np.array([truth[s][0][ob["cells"]] for s in steps]),
           "S_n": np.array([truth[s][1][ob["cells"]] for s in steps])}
    J, gK, gm = emul.gradient(keys, dict(st, K=np.exp(lnK0)), N, obs)
    assert J == emul.objective(run(lnK0, st["m"]), obs)
    rng = np.random.default_rng(77)
    eps = 1e-4
    for _ in range(2):
        vK = rng.standard_normal(lnK0.shape)
        vm = 0.02 * rng.standard_normal(lnK0.shape)
        fd = (emul.objective(run(lnK0 + eps * vK, st["m"] + eps * vm), obs)
              - emul.objective(run(lnK0 - eps * vK, st["m"] - eps * vm), obs)) / (2 * eps)
        ad = float(np.sum(gK * vK) + np.sum(gm * vm))
        assert abs(fd - ad) <= 1e-5 * abs(fd), (fd, ad)


@pytest.mark.parametrize("dims,seed,nsteps", [((7, 6, 5), 11, 6), ((9, 5, 1), 12, 4), ((1, 8, 6), 13, 4)])
def test_tangent_linear_and_adjoint_agree_to_rounding(dims, seed, nsteps):
    # the dot-product test: the tangent-linear step's dJ along (v_K, v_m)
    # equals the adjoint gradient's <g, v>; any error in the transpose (a
    # missed term, a transposed index) shows as O(1), rounding as ~1e-15
    keys, st, lnK, obs, rng = problem(*dims, seed, nsteps)
    J, g_lnK, g_m = emul.gradient(keys, st, nsteps, obs)
    for _ in range(3):
        vK = rng.standard_normal(lnK.shape)
        vm = 0.05 * rng.standard_normal(lnK.shape)
        Jt, dJ = emul.tangent(keys, st, nsteps, obs, vK, vm)
        assert Jt == J
        ad = float(np.sum(g_lnK * vK) + np.sum(g_m * vm))
        assert abs(dJ - ad) <= 1e-12 * abs(dJ), (dJ, ad)
