"""torch.autograd access to the history-matching objective.

``history_matching_loss(lnK, m, problem)`` returns J(ln K, m) as a 0-d
tensor; its backward pass hands torch the gradient that the discrete adjoint
(capaz_adjoint, include/capaz.h) computed with the forward run, so the
simulator is a differentiable layer of any torch model -- a permeability
and porosity inversion, a learned prior, a loss term.

The simulator's state lives in its own device buffers; the tensors passed
in and out are float64 tensors of the grid's shape (Nz, Ny, Nx), x fastest,
on the host.  The library and torch each bring a HIP runtime: tensors given
to this wrapper stay on the CPU, so the two never share device memory.
"""
from __future__ import annotations

import numpy as np
import torch

from . import capaz

STATE = ("P_w", "S_w", "S_n", "P_n", "P_g")


class HistoryMatchingProblem:
    """One forward problem: the deck `keys`, the initial primary state
    (P_w, S_w, S_n, P_n, P_g; K and m are the inputs), the observations
    {cells, steps, S_w, S_n} and the number of steps."""

    def __init__(self, keys: dict, state: dict, obs: dict, nsteps: int, checkpoint_every: int = 0, device: int = 0):
        self.keys = dict(keys)
        self.state = {f: np.asarray(state[f], dtype=np.float64) for f in STATE}
        self.obs = obs
        self.nsteps = int(nsteps)
        self.checkpoint_every = int(checkpoint_every)
        self.domain = capaz.Domain(keys, None, device)
        self.shape = self.domain.shape

    def _reset(self, lnK: np.ndarray, m: np.ndarray) -> None:
        for f, v in self.state.items():
            self.domain.set_field(f, v)
        self.domain.set_field("K", np.exp(lnK))
        self.domain.set_field("m", m)

    def objective(self, lnK: np.ndarray, m: np.ndarray) -> float:
        self._reset(lnK, m)
        return self.domain.objective(self.nsteps, self.obs)

    def objective_and_gradient(self, lnK: np.ndarray, m: np.ndarray):
        self._reset(lnK, m)
        return self.domain.adjoint(self.nsteps, self.obs, self.checkpoint_every)

    def close(self) -> None:
        self.domain.close()


class _HistoryMatchingLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, lnK: torch.Tensor, m: torch.Tensor, problem) -> torch.Tensor:
        to_np = lambda t: t.detach().to("cpu", torch.float64).numpy().reshape(problem.shape)
        J, g_lnK, g_m = problem.objective_and_gradient(to_np(lnK), to_np(m))
        ctx.save_for_backward(torch.from_numpy(g_lnK).to(lnK), torch.from_numpy(g_m).to(m))
        return torch.tensor(J, dtype=lnK.dtype, device=lnK.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        g_lnK, g_m = ctx.saved_tensors
        return grad_out * g_lnK, grad_out * g_m, None


def history_matching_loss(lnK: torch.Tensor, m: torch.Tensor, problem) -> torch.Tensor:
    """J(ln K, m) of `problem` (a HistoryMatchingProblem, or any object with
    the same objective_and_gradient and shape), differentiable in both."""
    return _HistoryMatchingLoss.apply(lnK, m, problem)
