"""Parity at the orders of the dispersive workload (24^3 elements, N = 9,
a Drude block in the central cube, a Gaussian pulse): the stage kernel's
source/Drude/Lorentz instance at N = 7, 8 and 9 against the oracle, in one
residual (resH, resE and resJ) and over 40 RK steps under every one-element
layout (lanes packed or not, free or 128-VGPR register budget), and the
Drude adjoint at the same orders against the oracle's adjoint and central
differences of the GPU forward.

N = 8 and 9 form the z-face fluxes in phase 1 (N = 7 in phase 2); layouts
2 and 3 read the RK register level by level, 0 and 1 up front; layouts 1
and 3 pack the four lane roles (at N = 7 the packed workgroup equals the
unpacked one); layouts 6 and 7 split the element six ways (free and
occupancy-bound register budgets).  Tolerances as in the other parity tests: 1e-11 of the
field's max for one residual, 1e-11 relative L2 (bm1-weighted) per field
after the steps; the adjoint's gradients within 1e-9 of the oracle's.
"""
import numpy as np
import pytest

from nekcem_amd.maxwell import MaxwellSolver
from nekcem_amd.mesh import BoxMesh

pytestmark = pytest.mark.gpu

ORDERS = [7, 8, 9]
LO, HI = (0.0, 0.0, 0.0), (1.0, 1.2, 0.9)
BC = ("PEC", "P", "PEC")


def _norm(u, w):
    return float(np.sqrt(np.sum(w * u * u)))


def _rel_l2(g, o, w):
    big = max(_norm(o[c], w) for c in range(len(o)))
    return max(_norm(g[c] - o[c], w) / max(_norm(o[c], w), 1e-6 * big) for c in range(len(o)))


def _physics(mesh, rng):
    """Source amplitude (bm1-weighted Gaussian, absent from some elements)
    and a Drude region that covers some elements whole and others in part,
    with per-node gamma and omega_p^2."""
    x, y, z = mesh.coords()
    amp = np.exp(-((x - 0.45) ** 2 + (y - 0.6) ** 2 + (z - 0.5) ** 2) / (2 * 0.2 ** 2)) * mesh.bmn
    amp[y > 0.9] = 0.0
    dindex = np.nonzero((x < 0.7) & (z > 0.25))[0]
    a = rng.uniform(0.05, 0.2, len(dindex))
    b = rng.uniform(1.0, 4.0, len(dindex))
    return amp, dindex, a, b


def _pair(oracle, N, seed):
    rng = np.random.default_rng(seed)
    nel = (2, 2, 2)
    mesh = BoxMesh(nel, N, LO, HI, BC)
    eps = np.repeat(rng.uniform(1, 3, mesh.nelt), mesh.n1 ** 3)
    ref = oracle.BoxCase(oracle.BoxSpec(nel=nel, N=N, lo=LO, hi=HI, bc=BC), eps=eps)
    return rng, mesh, eps, ref


@pytest.mark.parametrize("N", ORDERS)
def test_source_drude_residual_high_orders(oracle, N):
    """One residual of a random state and random current with the source on
    and the Drude block: resH, resE, resJ within 1e-11 of the oracle's max."""
    print(f"[case] source+Drude residual N={N}", flush=True)
    rng, mesh, eps, ref = _pair(oracle, N, 40 + N)
    amp, dindex, a, b = _physics(mesh, rng)
    n = mesh.npts
    H, E = rng.standard_normal((3, n)), rng.standard_normal((3, n))
    J = rng.standard_normal((3, len(dindex)))
    ref.set_source(amp, direction=(0.3, -0.2, 1.0), t0=0.4, tw=0.15)
    ref.set_drude(dindex, a, b)
    gpu = MaxwellSolver(mesh, eps=eps)
    gpu.set_source(amp, direction=(0.3, -0.2, 1.0), t0=0.4, tw=0.15)
    gpu.set_drude(dindex, a, b, J=J)
    gpu.set_state(H, E)
    for t in (0.35, 0.5):
        rH, rE = gpu.residual(t)
        oH, oE, oJ = ref.residual(H, E, t=t, J=J.reshape(-1))
        gJ = gpu.get_drude()[2]
        for g_, o_ in ((rH, oH), (rE, oE), (gJ, oJ.reshape(3, -1))):
            for c in range(3):
                assert np.max(np.abs(g_[c] - o_[c])) <= 1e-11 * np.max(np.abs(o_[c])), (t, c)
    gpu.close()


_STEPS_REF = {}


def _steps_reference(oracle, N):
    """The oracle's 40 steps from a random state and current with the pulse
    firing inside the window (cached per order)."""
    if N not in _STEPS_REF:
        rng, mesh, eps, ref = _pair(oracle, N, 50 + N)
        amp, dindex, a, b = _physics(mesh, rng)
        n = mesh.npts
        H, E = 0.1 * rng.standard_normal((3, n)), 0.1 * rng.standard_normal((3, n))
        J = 0.1 * rng.standard_normal((3, len(dindex)))
        dt = mesh.dt(0.25)
        src = dict(direction=(0.0, 0.4, 1.0), t0=10 * dt, tw=6 * dt)
        ref.set_source(amp, **src)
        ref.set_drude(dindex, a, b)
        out = ref.run(H, E, dt, 40, J=J.reshape(-1))
        _STEPS_REF[N] = (mesh, eps, amp, dindex, a, b, H, E, J, dt, src, out)
    return _STEPS_REF[N]


@pytest.mark.parametrize("layout", [0, 1, 2, 3, 6, 7])
@pytest.mark.parametrize("N", ORDERS)
def test_source_drude_steps_every_layout(oracle, N, layout):
    """40 RK steps (graph replay) with the pulse and the Drude block under
    each one-element layout: fields, J and J's RK register within 1e-11
    relative L2 of the oracle; the device time equals the oracle's."""
    print(f"[case] source+Drude 40 steps N={N} layout={layout}", flush=True)
    mesh, eps, amp, dindex, a, b, H, E, J, dt, src, out = _steps_reference(oracle, N)
    gpu = MaxwellSolver(mesh, eps=eps)
    gpu.set_source(amp, **src)
    gpu.set_drude(dindex, a, b, J=J)
    gpu.set_layout(layout)           # after the set-up, which re-tunes
    gpu.set_state(H, E)
    gpu.step(dt, 40)
    gH, gE = gpu.get_state()
    gJ, gkJ, _ = gpu.get_drude()
    assert gpu.layout()["layout"] == layout
    w = mesh.bmn
    assert _rel_l2(np.concatenate([gH, gE]), np.concatenate([out["H"], out["E"]]), w) <= 1e-11
    wJ = w[dindex]
    assert _rel_l2(gJ, out["J"].reshape(3, -1), wJ) <= 1e-11
    assert _rel_l2(gkJ, out["kJ"].reshape(3, -1), wJ) <= 1e-11
    assert gpu.time == out["time"]
    gpu.close()


def _adjoint_problem(oracle, N, K):
    rng = np.random.default_rng(60 + N)
    nel = (2, 2, 2)
    lo, hi = (0.0, 0.0, 0.0), (2.0, 2.0, 2.0)
    mesh = BoxMesh(nel, N, lo, hi, ("PEC",) * 3)
    n, n3 = mesh.npts, mesh.n1 ** 3
    x, y, z = mesh.coords()
    elems = np.array([0, 3, 5, 6])       # Drude elements (whole), per-element gamma, omega_p
    ne = len(elems)
    dindex = (elems[:, None] * n3 + np.arange(n3)[None, :]).ravel()
    gam, wp = rng.uniform(0.05, 0.2, ne), rng.uniform(2.0, 3.0, ne)
    eps = 1.5 + 0.2 * np.sin(np.pi * x) * np.cos(np.pi * y)
    amp = np.exp(-((x - 1) ** 2 + (y - 1) ** 2 + (z - 1) ** 2) / (2 * 0.3 ** 2)) * mesh.bmn
    probes = rng.choice(n, 8, replace=False)
    dt = mesh.dt(0.25)
    src = ((0.0, 0.0, 1.0), 8 * dt, 4 * dt)
    ref = oracle.BoxCase(oracle.BoxSpec(nel=nel, N=N, lo=lo, hi=hi, bc=("PEC",) * 3), eps=eps)
    ref.set_source(amp, *src)
    return rng, mesh, elems, dindex, gam, wp, eps, amp, probes, dt, src, ref


@pytest.mark.parametrize("N", ORDERS)
def test_drude_adjoint_high_orders(oracle, N):
    """Checkpointed Drude adjoint (40 steps, every 9: a ragged last segment)
    with the pulse: gradients with respect to per-node eps and per-node
    gamma, omega_p^2 within 1e-9 of the oracle's discrete adjoint, and the
    per-element gamma / omega_p gradients within 1e-6 of central
    differences (h = 1e-5) of the GPU forward."""
    K, every = 40, 9
    print(f"[case] Drude adjoint N={N} K={K} every={every}", flush=True)
    rng, mesh, elems, dindex, gam, wp, eps, amp, probes, dt, src, ref = _adjoint_problem(oracle, N, K)
    n, n3, ne = mesh.npts, mesh.n1 ** 3, len(elems)
    H0, E0 = np.zeros((3, n)), np.zeros((3, n))

    def ref_params(g, w):
        ref.set_drude(dindex, np.repeat(g, n3), np.repeat(w * w, n3))

    ref_params(gam * 1.3, wp * 0.8)
    target = ref.run(H0, E0, dt, K)["E"][2, probes].copy()
    ref_params(gam, wp)
    L_ref, ge_ref, _, _, ga_ref, gb_ref = ref.loss_and_grad(H0, E0, dt, K, probes, target, drude=True)

    s = MaxwellSolver(mesh, eps=eps)
    s.set_materials(eps)
    s.set_source(amp, *src)
    s.set_drude(dindex, np.repeat(gam, n3), np.repeat(wp * wp, n3))
    s.time = 0.0
    s.set_state(H0, E0)
    s.adjoint_forward(dt, K, every)
    E = s.get_state()[1]
    diff = E[2, probes] - target
    L = float(np.sum(diff * diff))
    assert L > 0 and abs(L - L_ref) <= 1e-9 * L_ref
    lamE = np.zeros((3, n))
    np.add.at(lamE[2], probes, 2 * diff)
    ge, _ = s.adjoint_backward(np.zeros((3, n)), lamE)
    ga, gb = s.adjoint_drude_grad()
    for g_, o_ in ((ge, ge_ref), (ga, ga_ref), (gb, gb_ref)):
        assert np.max(np.abs(o_)) > 0
        assert np.max(np.abs(g_ - o_)) <= 1e-9 * np.max(np.abs(o_))
    g_gam = ga.reshape(ne, n3).sum(axis=1)
    g_wp = 2 * wp * gb.reshape(ne, n3).sum(axis=1)

    def loss(g, w):
        s.set_drude(dindex, np.repeat(g, n3), np.repeat(w * w, n3))
        s.time = 0.0
        s.set_state(H0, E0)
        s.step(dt, K)
        d = s.get_state()[1][2, probes] - target
        return float(np.sum(d * d))

    h = 1e-5
    for which in ("gamma", "wp"):
        d = rng.standard_normal(ne)
        if which == "gamma":
            fd = (loss(gam + h * d, wp) - loss(gam - h * d, wp)) / (2 * h)
            an = float(np.dot(g_gam, d))
        else:
            fd = (loss(gam, wp + h * d) - loss(gam, wp - h * d)) / (2 * h)
            an = float(np.dot(g_wp, d))
        assert abs(fd - an) <= 1e-6 * abs(fd), (which, fd, an)
    s.close()
