"""Differentiable design objective: a torch.autograd.Function over the
native forward / reverse-time adjoint.

``DesignProblem`` holds one scenario with a design region and a DFT
objective (see ``scenarios.adjoint_design``); ``objective(eps_r)`` returns J
as a 0-d tensor whose backward() yields dJ/d eps_r from the native adjoint
(``xfd_forward_checkpointed`` / ``xfd_backward``).  The simulation runs on the
GPU through the C-ABI; torch only carries the parameter and its gradient.
"""

from __future__ import annotations

import numpy as np
import torch

from . import native


class DesignProblem:
    def __init__(self, spec: str, design_origin, design_shape, steps: int, checkpoint_interval: int = 100,
                 precision: str = "f32"):
        self.sim = native.Simulation(spec, precision)
        self.origin = tuple(int(v) for v in design_origin)
        self.shape = tuple(int(v) for v in design_shape)
        self.steps = int(steps)
        self.interval = int(checkpoint_interval)
        self.last_value = None

    def run_forward(self, eps_r: np.ndarray) -> float:
        if tuple(eps_r.shape) != self.shape:
            raise ValueError(f"eps_r shape {eps_r.shape} != design shape {self.shape}")
        self.sim.set_design(self.origin, eps_r)
        self.last_value = self.sim.forward(self.steps, self.interval)
        return self.last_value

    def run_backward(self) -> np.ndarray:
        return self.sim.backward(self.shape)

    def objective(self, eps_r: torch.Tensor) -> torch.Tensor:
        return _DesignObjective.apply(eps_r, self)


class _DesignObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps_r: torch.Tensor, problem: DesignProblem):
        e = eps_r.detach().to("cpu", torch.float64).numpy()
        value = problem.run_forward(e)
        ctx.problem = problem
        ctx.dtype = eps_r.dtype
        return torch.tensor(value, dtype=torch.float64, device=eps_r.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        g = ctx.problem.run_backward()
        gt = torch.from_numpy(g).to(device=grad_out.device, dtype=torch.float64)
        return (grad_out * gt).to(ctx.dtype), None
