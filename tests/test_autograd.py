"""The torch.autograd wrapper of the history-matching objective
(capaz_amd/autograd.py) on the CPU: the problem object is the host build of
the forward step and of the discrete adjoint, so the wrapper's plumbing --
gradients handed to torch, the chain rule through torch operations around
the simulator -- is checked without a GPU.  The device problem
(HistoryMatchingProblem over capaz_adjoint) is checked in
tests/test_autograd_gpu.py."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from capaz_amd.autograd import history_matching_loss  # noqa: E402
from capaz_cases import case, random_state  # noqa: E402
from emul import emul  # noqa: E402

pytestmark = pytest.mark.skipif(not emul.available(), reason="host build missing")


class HostProblem:
    def __init__(self, keys, st, obs, nsteps):
        self.keys, self.st, self.obs, self.nsteps = keys, st, obs, nsteps
        self.shape = st["P_w"].shape

    def objective_and_gradient(self, lnK, m):
        return emul.gradient(self.keys, dict(self.st, K=np.exp(lnK), m=m), self.nsteps, self.obs)


def make(seed=61):
    keys = case("3ph", 7, 6, 5, CLOSURE="corey", S_WR_0=0.1, S_NR_0=0.05, SOURCE=1, DT=1e-4, ITERATIONS=4)
    st = random_state(keys, seed)
    rng = np.random.default_rng(seed)
    shape = st["P_w"].shape
    lnK = np.log(keys["K_0"]) + 0.5 * rng.standard_normal(shape)
    cells = rng.choice(int(np.prod(shape)), 10, replace=False)
    obs = {"cells": cells, "steps": [2, 4], "S_w": rng.uniform(0.2, 0.5, (2, 10)), "S_n": rng.uniform(0.1, 0.4, (2, 10))}
    return HostProblem(keys, st, obs, 4), lnK, st["m"], rng


def test_backward_hands_torch_the_adjoint_gradient():
    prob, lnK, m, _ = make()
    J, gK, gm = prob.objective_and_gradient(lnK, m)
    tK = torch.tensor(lnK, requires_grad=True)
    tm = torch.tensor(m, requires_grad=True)
    loss = history_matching_loss(tK, tm, prob)
    assert loss.item() == J
    (3.0 * loss).backward()
    assert np.array_equal(tK.grad.numpy(), 3.0 * gK)
    assert np.array_equal(tm.grad.numpy(), 3.0 * gm)


def test_chain_rule_through_a_parametrisation():
    # ln K = mean + s * z and m = 0.4 * sigmoid-free affine map of w: torch owns the outer derivatives
    prob, lnK, m, rng = make(62)
    mean = float(lnK.mean())
    z = torch.tensor((lnK - mean) / 0.5, requires_grad=True)
    w = torch.tensor(m / 0.4, requires_grad=True)
    loss = history_matching_loss(mean + 0.5 * z, 0.4 * w, prob)
    loss.backward()
    _, gK, gm = prob.objective_and_gradient(lnK, m)
    assert np.allclose(z.grad.numpy(), 0.5 * gK, rtol=1e-12, atol=0)
    assert np.allclose(w.grad.numpy(), 0.4 * gm, rtol=1e-12, atol=0)
