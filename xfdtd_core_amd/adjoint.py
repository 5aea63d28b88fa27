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
        # counts forward runs; the checkpoints on the device belong to the
        # latest one only
        self.generation = 0

    def run_forward(self, eps_r: np.ndarray) -> float:
        if tuple(eps_r.shape) != self.shape:
            raise ValueError(f"eps_r shape {eps_r.shape} != design shape {self.shape}")
        self.sim.set_design(self.origin, eps_r)
        self.last_value = self.sim.forward(self.steps, self.interval)
        self.generation += 1
        return self.last_value

    def run_backward(self, generation: int | None = None, eps_r: np.ndarray | None = None) -> np.ndarray:
        """Gradient of the forward run `generation` (default: the latest).
        The device keeps the checkpoints of the latest forward only, so an
        older generation is run forward again from its saved eps_r first."""
        if generation is not None and generation != self.generation:
            if eps_r is None:
                raise RuntimeError(f"backward of forward run {generation}, but the checkpoints on the device "
                                   f"belong to run {self.generation} and no eps_r was saved to recompute them")
            self.run_forward(eps_r)
        return self.sim.backward(self.shape)

    def objective(self, eps_r: torch.Tensor) -> torch.Tensor:
        return _DesignObjective.apply(eps_r, self)


class _DesignObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps_r: torch.Tensor, problem: DesignProblem):
        e = eps_r.detach().to("cpu", torch.float64).numpy()
        value = problem.run_forward(e)
        ctx.problem = problem
        ctx.generation = problem.generation
        ctx.eps_r = e
        ctx.dtype = eps_r.dtype
        return torch.tensor(value, dtype=torch.float64, device=eps_r.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        g = ctx.problem.run_backward(ctx.generation, ctx.eps_r)
        gt = torch.from_numpy(g).to(device=grad_out.device, dtype=torch.float64)
        return (grad_out * gt).to(ctx.dtype), None
