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
