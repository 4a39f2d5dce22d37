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
PyTorch is only the autograd plumbing.  The tensors stay where they are:
with GPU tensors the parameters, the initial state, the probe fields, the
adjoint seeds and the gradients go between torch and the library as device
pointers (device-to-device copies); only the material and Drude tables,
which the library builds on the host, stage their per-node inputs through
host memory.  CPU tensors go through the same calls as host pointers.
torch ships its own HIP runtime; the two share one HSA runtime, and torch
sees the GPU, only if torch's is loaded first: import torch (and call
torch.cuda.init()) before the first MaxwellSolver of the process.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
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
    _dev: dict = field(default_factory=dict, repr=False)   # the above as tensors, per device


def _run_forward(problem: ProbeProblem, eps: np.ndarray, mu: np.ndarray):
    """The probe fields for host arrays eps, mu (numpy; finite differences)."""
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
    _dev: dict = field(default_factory=dict, repr=False)


def _run_forward_drude(problem: DrudeProbeProblem, gamma: np.ndarray, wp: np.ndarray):
    """The probe fields for host arrays gamma, wp (numpy; finite differences)."""
    s = problem.solver
    s.adjoint_set_grads(False)      # only the Drude gradients are asked for
    s.set_drude(problem.dindex, gamma[problem.owner], (wp * wp)[problem.owner])
    s.time = problem.t0
    s.set_state(problem.H0, problem.E0)
    s.adjoint_forward(problem.dt, problem.nsteps, problem.every)
    H, E = s.get_state()
    return np.concatenate([H[:, problem.probes], E[:, problem.probes]])


if torch is not None:

    _dp = C.POINTER(C.c_double)

    def _ptr(t):
        """double* of a contiguous float64 tensor (host or device memory)."""
        return None if t is None else C.cast(C.c_void_p(t.data_ptr()), _dp)

    def _ready(*ts):
        """The library copies synchronously, outside torch's streams: let
        torch's pending work on these tensors finish first."""
        for t in ts:
            if t is not None and t.is_cuda:
                torch.cuda.synchronize(t.device)
                return

    def _problem_tensors(problem, dev):
        """H0, E0 (3, npts), the probe indices (and for a Drude problem the
        owners and the 1-based node list) on dev, made once per device."""
        key = str(dev)
        if key not in problem._dev:
            n = problem.solver.npts
            H0 = torch.as_tensor(np.asarray(problem.H0), dtype=torch.float64, device=dev).reshape(3, n).contiguous()
            E0 = torch.as_tensor(np.asarray(problem.E0), dtype=torch.float64, device=dev).reshape(3, n).contiguous()
            pr = torch.as_tensor(np.asarray(problem.probes), dtype=torch.long, device=dev)
            extra = {}
            if isinstance(problem, DrudeProbeProblem):
                extra["owner"] = torch.as_tensor(np.asarray(problem.owner), dtype=torch.long, device=dev)
                extra["dindex"] = np.ascontiguousarray(np.asarray(problem.dindex) + 1, dtype=np.int32)
            problem._dev[key] = (H0, E0, pr, extra)
        return problem._dev[key]

    def _start(problem, dev):
        """Time and state to the problem's start (device to device)."""
        s = problem.solver
        H0, E0, _, _ = _problem_tensors(problem, dev)
        s.time = problem.t0
        _ready(H0)
        _lib.check(s._L.nkc_set_state(s._h, _ptr(H0), _ptr(E0), None, None, s.npts))

    def _probe_fields(problem, dev):
        """The six fields at the probes after the forward, a (6, nprobes)
        tensor on dev: the state copied device to device, then gathered."""
        s = problem.solver
        n = s.npts
        _, _, pr, _ = _problem_tensors(problem, dev)
        H = torch.empty((3, n), dtype=torch.float64, device=dev)
        E = torch.empty((3, n), dtype=torch.float64, device=dev)
        _ready(H)
        _lib.check(s._L.nkc_get_state(s._h, _ptr(H), _ptr(E), None, None, n))
        return torch.cat([H[:, pr], E[:, pr]])

    def _adjoint(problem, grad_out, want_material):
        """The backward sweep seeded by grad_out at the probes (index_add: a
        probe listed twice gets both seeds); dL/deps, dL/dmu when asked."""
        s = problem.solver
        n = s.npts
        dev = grad_out.device
        _, _, pr, _ = _problem_tensors(problem, dev)
        g = grad_out.detach().to(torch.float64)
        lamH = torch.zeros((3, n), dtype=torch.float64, device=dev).index_add_(1, pr, g[:3])
        lamE = torch.zeros((3, n), dtype=torch.float64, device=dev).index_add_(1, pr, g[3:])
        ge = torch.empty(n, dtype=torch.float64, device=dev) if want_material else None
        gm = torch.empty(n, dtype=torch.float64, device=dev) if want_material else None
        _ready(lamH)
        _lib.check(s._L.nkc_adjoint_backward(s._h, _ptr(lamH), _ptr(lamE), n, _ptr(ge), _ptr(gm)))
        return ge, gm

    class ProbeFields(torch.autograd.Function):
        """(eps, mu) -> fields (Hx, Hy, Hz, Ex, Ey, Ez) at the probes."""

        @staticmethod
        def forward(ctx, eps, mu, problem: ProbeProblem):
            s = problem.solver
            dev = eps.device
            e = eps.detach().to(torch.float64).contiguous()
            m = mu.detach().to(device=dev, dtype=torch.float64).contiguous()
            s.adjoint_set_grads(True)
            _ready(e, m)
            _lib.check(s._L.nkc_set_materials(s._h, _ptr(e), _ptr(m)))
            _start(problem, dev)
            s.adjoint_forward(problem.dt, problem.nsteps, problem.every)
            ctx.problem = problem
            return _probe_fields(problem, dev)

        @staticmethod
        def backward(ctx, grad_out):
            ge, gm = _adjoint(ctx.problem, grad_out, True)
            return ge, gm, None

    class DrudeProbeFields(torch.autograd.Function):
        """(gamma, omega_p) per Drude element -> fields at the probes; the
        backward runs the same checkpointed adjoint, whose Drude part
        reverses the polarisation-current ODE (dr_J/da = -J, dr_J/db = E)."""

        @staticmethod
        def forward(ctx, gamma, wp, problem: DrudeProbeProblem):
            s = problem.solver
            dev = gamma.device
            _, _, _, extra = _problem_tensors(problem, dev)
            owner, di = extra["owner"], extra["dindex"]
            g = gamma.detach().to(torch.float64)
            w = wp.detach().to(device=dev, dtype=torch.float64)
            par = torch.stack([g[owner], (w * w)[owner]], dim=1).contiguous()   # (a, b) per listed node
            s.adjoint_set_grads(False)      # only the Drude gradients are asked for
            _ready(par)
            _lib.check(s._L.nkc_set_drude(s._h, len(di), di.ctypes.data_as(C.POINTER(C.c_int32)), _ptr(par),
                                          None, None))
            s.ndrude = len(di)
            _start(problem, dev)
            s.adjoint_forward(problem.dt, problem.nsteps, problem.every)
            ctx.problem = problem
            ctx.wp = w
            ctx.ne = g.shape[0]
            return _probe_fields(problem, dev)

        @staticmethod
        def backward(ctx, grad_out):
            p = ctx.problem
            s = p.solver
            dev = grad_out.device
            _adjoint(p, grad_out, False)
            _, _, _, extra = _problem_tensors(p, dev)
            owner = extra["owner"]
            nd = len(extra["dindex"])
            ga = torch.empty(nd, dtype=torch.float64, device=dev)
            gb = torch.empty(nd, dtype=torch.float64, device=dev)
            _ready(ga)
            _lib.check(s._L.nkc_adjoint_get_drude(s._h, _ptr(ga), _ptr(gb)))
            g_gamma = torch.zeros(ctx.ne, dtype=torch.float64, device=dev).index_add_(0, owner, ga)
            g_wp = 2.0 * ctx.wp * torch.zeros(ctx.ne, dtype=torch.float64, device=dev).index_add_(0, owner, gb)
            return g_gamma, g_wp, None
