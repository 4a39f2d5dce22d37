"""The torch harness with GPU tensors (nekcem_amd/adjoint.py): parameters,
initial state, probe fields, seeds and gradients stay in device memory and
cross the C-ABI as device pointers.

torch ships its own HIP runtime; the library links the system one.  They
share one HSA runtime (and torch sees the GPU) only when torch's is loaded
first, so the check runs in a child interpreter that initialises torch's GPU
before the library, as a user of the harness does (the test runner's own
process has the library loaded already).  The child compares, bit for bit,
ProbeFields and DrudeProbeFields on cuda tensors against the same calls on
CPU tensors (host pointers) and prints one JSON line.
"""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import json, sys
import torch
torch.cuda.init()                      # torch's HIP runtime first
assert torch.cuda.is_available()
sys.path.insert(0, ROOT)
import numpy as np
from nekcem_amd.adjoint import ProbeFields, ProbeProblem, DrudeProbeFields, DrudeProbeProblem
from nekcem_amd.maxwell import MaxwellSolver
from nekcem_amd.mesh import BoxMesh

rng = np.random.default_rng(3)
mesh = BoxMesh((2, 2, 2), 5, (0, 0, 0), (1, 1, 1), ("PEC",) * 3)
n, n3 = mesh.npts, mesh.n1 ** 3
x, y, z = mesh.coords()
eps = 2 + 0.3 * np.sin(2 * np.pi * x) * np.cos(np.pi * y)
amp = np.exp(-((x - .4) ** 2 + (y - .5) ** 2 + (z - .6) ** 2) / (2 * .15 ** 2)) * mesh.bmn
probes = rng.choice(n, 8, replace=False)
dt = mesh.dt(0.25)
H0, E0 = np.zeros((3, n)), 0.01 * rng.standard_normal((3, n))
w = rng.standard_normal((6, 8))
res = {}

s = MaxwellSolver(mesh, eps=eps)
s.set_source(amp, (0, 0, 1), 0.05, 0.03)
prob = ProbeProblem(solver=s, H0=H0, E0=E0, dt=dt, nsteps=20, every=7, probes=probes)
outs = {}
for dev in ("cuda:0", "cpu"):
    e = torch.tensor(eps, dtype=torch.float64, device=dev, requires_grad=True)
    m = torch.ones(n, dtype=torch.float64, device=dev, requires_grad=True)
    out = ProbeFields.apply(e, m, prob)
    (torch.tensor(w, device=dev) * out ** 2).sum().backward()
    outs[dev] = (out.detach().cpu().numpy(), e.grad.cpu().numpy(), m.grad.cpu().numpy(),
                 out.device.type, e.grad.device.type)
a, b = outs["cuda:0"], outs["cpu"]
res["probe_on_gpu"] = a[3] == "cuda" and a[4] == "cuda"
res["probe_equal"] = all(bool(np.array_equal(u, v)) for u, v in zip(a[:3], b[:3]))
res["probe_grad_nonzero"] = bool(np.max(np.abs(a[1])) > 0)
s.close()

s = MaxwellSolver(mesh, eps=eps)
s.set_source(amp, (0, 0, 1), 0.05, 0.03)
elems = np.array([0, 3, 5])
dindex = (elems[:, None] * n3 + np.arange(n3)[None, :]).ravel()
owner = np.repeat(np.arange(len(elems)), n3)
gam, wp = rng.uniform(0.05, 0.2, 3), rng.uniform(2.0, 3.0, 3)
s.set_drude(dindex, gam[owner], (wp * wp)[owner])
prob = DrudeProbeProblem(solver=s, H0=H0, E0=E0, dt=dt, nsteps=20, every=7, probes=probes,
                         dindex=dindex, owner=owner)
outs = {}
for dev in ("cuda:0", "cpu"):
    tg = torch.tensor(gam, dtype=torch.float64, device=dev, requires_grad=True)
    tw = torch.tensor(wp, dtype=torch.float64, device=dev, requires_grad=True)
    out = DrudeProbeFields.apply(tg, tw, prob)
    (torch.tensor(w, device=dev) * out ** 2).sum().backward()
    outs[dev] = (out.detach().cpu().numpy(), tg.grad.cpu().numpy(), tw.grad.cpu().numpy(), tg.grad.device.type)
a, b = outs["cuda:0"], outs["cpu"]
res["drude_on_gpu"] = a[3] == "cuda"
res["drude_out_equal"] = bool(np.array_equal(a[0], b[0]))
res["drude_grad_close"] = bool(np.allclose(a[1], b[1], rtol=1e-12, atol=0) and np.allclose(a[2], b[2], rtol=1e-12, atol=0))
s.close()
print("RESULT " + json.dumps(res), flush=True)
"""


def test_torch_harness_with_gpu_tensors():
    """ProbeFields / DrudeProbeFields on cuda tensors: outputs and gradients
    on the GPU, equal to the CPU-tensor path (Drude gradients, summed per
    element with atomics on the GPU, to 1e-12)."""
    print("[case] torch harness on GPU tensors (child interpreter)", flush=True)
    code = "ROOT = " + repr(ROOT) + "\n" + CHILD
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    line = [l for l in r.stdout.splitlines() if l.startswith("RESULT ")][-1]
    res = json.loads(line[len("RESULT "):])
    assert all(res.values()), res
