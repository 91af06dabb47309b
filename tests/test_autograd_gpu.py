"""The torch.autograd wrapper over the device adjoint (HistoryMatchingProblem
on device 0): the loss and its gradients are capaz_objective's J and
capaz_adjoint's gradient, and a directional derivative of the loss matches
central differences of the loss itself."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_torch_loss_and_gradient_on_the_device_adjoint():
    torch = pytest.importorskip("torch")
    from capaz_amd import capaz
    from capaz_amd.autograd import HistoryMatchingProblem, history_matching_loss
    from capaz_cases import case, random_state

    keys = case("3ph", 16, 14, 10, CLOSURE="corey", S_WR_0=0.1, S_NR_0=0.05, SOURCE=1, DT=1e-4, ITERATIONS=4)
    st = random_state(keys, 71)
    rng = np.random.default_rng(71)
    shape = st["P_w"].shape
    lnK = np.log(keys["K_0"]) + 0.5 * rng.standard_normal(shape)
    cells = rng.choice(int(np.prod(shape)), 16, replace=False)
    obs = {"cells": cells, "steps": [3, 6], "S_w": rng.uniform(0.2, 0.5, (2, 16)), "S_n": rng.uniform(0.1, 0.4, (2, 16))}
    prob = HistoryMatchingProblem(keys, st, obs, 6)
    tK = torch.tensor(lnK, requires_grad=True)
    tm = torch.tensor(st["m"], requires_grad=True)
    loss = history_matching_loss(tK, tm, prob)
    loss.backward()
    ref = capaz.Domain(keys, dict(st, K=np.exp(lnK)))
    J, gK, gm = ref.adjoint(6, obs)
    ref.close()
    assert loss.item() == J
    assert np.array_equal(tK.grad.numpy(), gK) and np.array_equal(tm.grad.numpy(), gm)
    eps = 1e-4
    v = rng.standard_normal(shape)
    with torch.no_grad():
        fd = (history_matching_loss(tK + eps * torch.tensor(v), tm, prob)
              - history_matching_loss(tK - eps * torch.tensor(v), tm, prob)).item() / (2 * eps)
    ad = float((tK.grad * torch.tensor(v)).sum())
    assert abs(fd - ad) <= 1e-5 * abs(fd)
    prob.close()
