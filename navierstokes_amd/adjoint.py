"""Reverse mode behind torch.autograd.Function.

`ObserverObjective.apply(dp, solver, nsteps, every)` runs the HIP time
advance of the 2-D Euler equations from the base state with the initial
pressure perturbation `dp` (a (jmax, imax) float64 device tensor) and returns
the observer objective

    J = dt * sum_{n=1..nsteps} sum_arc (p_n - pb)^2

as a 0-d tensor.  Its backward runs the library's hand-written adjoint
kernels (checkpoints every `every` steps, recomputation inside each segment)
and returns dJ/d(dp).  Both precision modes of the solver work; the
perturbation form returns its FP32 adjoint widened to float64.

The reference has no adjoint; this is the reverse mode of its LDDRK6 step
(rk.f90:12-75) for the inviscid equations.
"""
from __future__ import annotations

import torch

from .solver import Solver


def _base_planes(solver: Solver, device) -> torch.Tensor:
    cache = getattr(solver, "_base_dev", None)
    if cache is None or cache.device != device:
        ub = torch.from_numpy(solver.case.Ub).permute(2, 0, 1).contiguous()
        cache = ub.to(device=device, dtype=torch.float64)
        solver._base_dev = cache
    return cache


def _run_forward(solver: Solver, dp: torch.Tensor, nsteps: int, every: int, iter0: int) -> float:
    U0 = _base_planes(solver, dp.device).clone()
    U0[3] += dp
    torch.cuda.current_stream(dp.device).synchronize()  # the library runs on its own stream
    solver.put_state_dev(U0)
    return solver.adj_forward(iter0, nsteps, every)


class ObserverObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, dp: torch.Tensor, solver: Solver, nsteps: int, every: int, iter0: int = 0):
        c = solver.case
        if tuple(dp.shape) != (c.jmax, c.imax) or dp.dtype != torch.float64 or not dp.is_cuda:
            raise ValueError(f"dp must be a float64 device tensor of shape {(c.jmax, c.imax)}")
        J = _run_forward(solver, dp, nsteps, every, iter0)
        ctx.solver = solver
        ctx.args = (nsteps, every, iter0)
        # the solver holds the checkpoints of its latest forward only: the
        # backward re-runs this forward when another one came after it
        ctx.generation = solver.adj_generation
        ctx.save_for_backward(dp.detach().clone())
        return torch.tensor(J, dtype=torch.float64, device=dp.device)

    @staticmethod
    def backward(ctx, gJ: torch.Tensor):
        solver = ctx.solver
        c = solver.case
        if solver.adj_generation != ctx.generation:
            (dp,) = ctx.saved_tensors
            _run_forward(solver, dp, *ctx.args)
            ctx.generation = solver.adj_generation
        g = torch.empty((c.nk, c.jmax, c.imax), dtype=torch.float64, device=gJ.device)
        torch.cuda.current_stream(gJ.device).synchronize()
        solver.adj_backward_dev(g)
        return gJ * g[3], None, None, None, None


def observer_objective(dp: torch.Tensor, solver: Solver, nsteps: int, every: int, iter0: int = 0) -> torch.Tensor:
    return ObserverObjective.apply(dp, solver, nsteps, every, iter0)
