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

``DrudeProbeFields.apply(gamma, wp, problem)`` does the same with the
per-element Drude damping and plasma frequency as the parameters.

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
    s.adjoint_set_grads(True)
    s.set_materials(eps, mu)
    s.time = problem.t0
    s.set_state(problem.H0, problem.E0)
    s.adjoint_forward(problem.dt, problem.nsteps, problem.every)
    H, E = s.get_state()
    return np.concatenate([H[:, problem.probes], E[:, problem.probes]])


@dataclass
class DrudeProbeProblem:
    """A probe problem whose parameters are the Drude gamma and omega_p of
    each Drude element (set_drude's a = gamma, b = omega_p^2 at every listed
    node of the element); eps and mu stay as the solver holds them."""
    solver: MaxwellSolver
    H0: np.ndarray
    E0: np.ndarray
    dt: float
    nsteps: int
    every: int
    probes: np.ndarray
    dindex: np.ndarray        # listed volume nodes (0-based)
    owner: np.ndarray         # per listed node: its Drude element (0..ne-1)
    t0: float = 0.0


def _run_forward_drude(problem: DrudeProbeProblem, gamma: np.ndarray, wp: np.ndarray):
    s = problem.solver
    s.adjoint_set_grads(False)      # only the Drude gradients are asked for
    s.set_drude(problem.dindex, gamma[problem.owner], (wp * wp)[problem.owner])
    s.time = problem.t0
    s.set_state(problem.H0, problem.E0)
    s.adjoint_forward(problem.dt, problem.nsteps, problem.every)
    H, E = s.get_state()
    return np.concatenate([H[:, problem.probes], E[:, problem.probes]])


def _seed(probes, n, g):
    lamH, lamE = np.zeros((3, n)), np.zeros((3, n))
    for c in range(3):
        np.add.at(lamH[c], probes, g[c])
        np.add.at(lamE[c], probes, g[3 + c])
    return lamH, lamE


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
            g = grad_out.detach().cpu().numpy().astype(np.float64)
            ge, gm = p.solver.adjoint_backward(*_seed(p.probes, p.solver.npts, g))
            dev = grad_out.device
            return torch.from_numpy(ge).to(dev), torch.from_numpy(gm).to(dev), None

    class DrudeProbeFields(torch.autograd.Function):
        """(gamma, omega_p) per Drude element -> fields at the probes; the
        backward runs the same checkpointed adjoint, whose Drude part
        reverses the polarisation-current ODE (dr_J/da = -J, dr_J/db = E)."""

        @staticmethod
        def forward(ctx, gamma, wp, problem: DrudeProbeProblem):
            g = gamma.detach().cpu().numpy().astype(np.float64)
            w = wp.detach().cpu().numpy().astype(np.float64)
            out = _run_forward_drude(problem, g, w)
            ctx.problem = problem
            ctx.wp = w
            return torch.from_numpy(out).to(gamma.device)

        @staticmethod
        def backward(ctx, grad_out):
            p = ctx.problem
            g = grad_out.detach().cpu().numpy().astype(np.float64)
            p.solver.adjoint_backward(*_seed(p.probes, p.solver.npts, g))
            ga, gb = p.solver.adjoint_drude_grad()
            ne = len(ctx.wp)
            g_gamma = np.bincount(p.owner, weights=ga, minlength=ne)
            g_wp = 2.0 * ctx.wp * np.bincount(p.owner, weights=gb, minlength=ne)
            dev = grad_out.device
            return torch.from_numpy(g_gamma).to(dev), torch.from_numpy(g_wp).to(dev), None
