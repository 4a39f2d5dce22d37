"""PyTorch harness over the HIP discrete adjoint.

``ProbeFields.apply(eps, mu, problem)`` maps per-node permittivity and
permeability (torch float64 tensors of npts) to the six fields at a set of
probe nodes after ``nsteps`` RK steps of the HIP forward; its backward runs
the checkpointed adjoint of libnekcem_hip.so (transposed curl, transposed
flux/lift, reversed low-storage RK; checkpoints every ``every`` steps and a
recompute of each segment) and returns dL/deps, dL/dmu.  Any loss written in
torch on the probe fields can then be differentiated with autograd, e.g.

    out = ProbeFields.apply(eps, mu, problem)        # (6, nprobes)
    loss = ((out[5] - target) ** 2).sum()             # probe Ez
    loss.backward()                                   # eps.grad, mu.grad

The reference has no adjoint; the forward it differentiates is the step of
cem_maxwell_op_rk (src/cem_maxwell.F:327-345) with a usersrc-style source.
PyTorch is only the autograd plumbing: fields move between torch and the
library through host arrays at the two ends of the run.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .maxwell import MaxwellSolver

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None


@dataclass
class ProbeProblem:
    """Everything fixed while eps and mu vary."""
    solver: MaxwellSolver
    H0: np.ndarray            # (3, npts) initial state
    E0: np.ndarray
    dt: float
    nsteps: int
    every: int                # checkpoint interval (steps)
    probes: np.ndarray        # node indices (0-based)
    t0: float = 0.0


def _run_forward(problem: ProbeProblem, eps: np.ndarray, mu: np.ndarray):
    s = problem.solver
    s.set_materials(eps, mu)
    s.time = problem.t0
    s.set_state(problem.H0, problem.E0)
    s.adjoint_forward(problem.dt, problem.nsteps, problem.every)
    H, E = s.get_state()
    return np.concatenate([H[:, problem.probes], E[:, problem.probes]])


if torch is not None:

    class ProbeFields(torch.autograd.Function):
        """(eps, mu) -> fields (Hx, Hy, Hz, Ex, Ey, Ez) at the probes."""

        @staticmethod
        def forward(ctx, eps, mu, problem: ProbeProblem):
            e = eps.detach().cpu().numpy().astype(np.float64)
            m = mu.detach().cpu().numpy().astype(np.float64)
            out = _run_forward(problem, e, m)
            ctx.problem = problem
            return torch.from_numpy(out).to(eps.device)

        @staticmethod
        def backward(ctx, grad_out):
            p = ctx.problem
            n = p.solver.npts
            g = grad_out.detach().cpu().numpy().astype(np.float64)
            lamH, lamE = np.zeros((3, n)), np.zeros((3, n))
            for c in range(3):
                np.add.at(lamH[c], p.probes, g[c])
                np.add.at(lamE[c], p.probes, g[3 + c])
            ge, gm = p.solver.adjoint_backward(lamH, lamE)
            dev = grad_out.device
            return torch.from_numpy(ge).to(dev), torch.from_numpy(gm).to(dev), None
