"""The HIP discrete adjoint: transposed operator and material gradients,
against the oracle's vector-Jacobian product and central differences."""
import numpy as np
import pytest

from nekcem_amd.maxwell import MaxwellSolver
from nekcem_amd.mesh import BoxMesh

pytestmark = pytest.mark.gpu

CASES = [
    dict(nel=(2, 3, 2), N=5, bc=("P", "P", "P"), upwind=True),
    dict(nel=(2, 2, 2), N=7, bc=("PEC", "PEC", "PEC"), upwind=True),
    dict(nel=(3, 2, 2), N=4, bc=("P", "PEC", "P"), upwind=False),
    dict(nel=(2, 2, 2), N=8, bc=("PEC", "P", "P"), upwind=True),
    dict(nel=(2, 1, 2), N=9, bc=("P", "PEC", "P"), upwind=True),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"N{c['N']}-{'-'.join(c['bc'])}-{'up' if c['upwind'] else 'cen'}")
@pytest.mark.parametrize("material", ["element", "node"])
def test_adjoint_apply_is_transpose(oracle, case, material):
    """<k, r(u)> = <B^T P k, u> on the GPU, and B^T P k equals the oracle's
    vjp of the same residual."""
    rng = np.random.default_rng(3)
    nel, N, bc, up = case["nel"], case["N"], case["bc"], case["upwind"]
    print(f"[case] adjoint transpose nel={nel} N={N} bc={bc} upwind={up} {material}", flush=True)
    mesh = BoxMesh(nel, N, (0, 0, 0), (1.0, 1.2, 0.9), bc)
    n = mesh.npts
    if material == "element":
        eps = np.repeat(rng.uniform(1, 4, mesh.nelt), mesh.n1 ** 3)
        mu = np.repeat(rng.uniform(1, 2, mesh.nelt), mesh.n1 ** 3)
    else:
        eps, mu = rng.uniform(1, 4, n), rng.uniform(1, 2, n)
    g = MaxwellSolver(mesh, eps=eps, mu=mu, ifupwind=up)
    u = (rng.standard_normal((3, n)), rng.standard_normal((3, n)))
    k = (rng.standard_normal((3, n)), rng.standard_normal((3, n)))
    g.set_state(*u)
    rH, rE = g.residual()
    aH, aE = g.adjoint_apply(*k)
    lhs = np.sum(k[0] * rH) + np.sum(k[1] * rE)
    rhs = np.sum(aH * u[0]) + np.sum(aE * u[1])
    assert abs(lhs - rhs) <= 1e-12 * (np.sum(np.abs(k[0] * rH)) + np.sum(np.abs(k[1] * rE)))
    ref = oracle.BoxCase(oracle.BoxSpec(nel=nel, N=N, lo=(0, 0, 0), hi=(1.0, 1.2, 0.9), bc=bc, upwind=up),
                         eps=eps, mu=mu)
    oH, oE, _, _ = ref.vjp(*u, *k)
    for a_, b_ in ((aH, oH), (aE, oE)):
        for c in range(3):
            assert np.max(np.abs(a_[c] - b_[c])) <= 1e-11 * np.max(np.abs(b_[c]))
    g.close()


def _config4_small(oracle, nel=(2, 2, 2), N=5, K=20, every=7, seed=1):
    rng = np.random.default_rng(seed)
    lo, hi = (0, 0, 0), (1, 1, 1)
    mesh = BoxMesh(nel, N, lo, hi, ("PEC",) * 3)
    n = mesh.npts
    x, y, z = mesh.coords()
    eps = 2 + 0.3 * np.sin(2 * np.pi * x) * np.cos(np.pi * y) * np.sin(np.pi * z + 0.3)
    amp = np.exp(-((x - .4) ** 2 + (y - .5) ** 2 + (z - .6) ** 2) / (2 * .15 ** 2)) * mesh.bmn
    probes = rng.choice(n, 8, replace=False)
    ref = oracle.BoxCase(oracle.BoxSpec(nel=nel, N=N, lo=lo, hi=hi, bc=("PEC",) * 3), eps=eps)
    ref.set_source(amp, (0, 0, 1), 0.05, 0.03)
    dt = mesh.dt(0.25)
    H0, E0 = np.zeros((3, n)), np.zeros((3, n))
    ref.set_materials(eps * 1.05, None)
    target = ref.run(H0, E0, dt, K)["E"][2, probes].copy()
    ref.set_materials(eps, None)
    return mesh, eps, amp, probes, target, ref, dt, K, every, H0, E0, rng


def _gpu_loss(mesh, eps, amp, probes, target, dt, K, H0, E0, solver=None):
    s = solver or MaxwellSolver(mesh, eps=eps)
    if solver is not None:
        s.set_materials(eps)
    s.set_source(amp, (0, 0, 1), 0.05, 0.03)
    s.time = 0.0
    s.set_state(H0, E0)
    s.step(dt, K)
    E = s.get_state()[1]
    d = E[2, probes] - target
    return float(np.sum(d * d)), s


def test_adjoint_gradient_matches_oracle_and_fd(oracle):
    """Checkpointed GPU adjoint (20 steps, checkpoint every 7) vs the
    oracle's adjoint, and vs central FD (h = 1e-6) of the GPU forward."""
    mesh, eps, amp, probes, target, ref, dt, K, every, H0, E0, rng = _config4_small(oracle)
    n = mesh.npts
    L_ref, ge_ref, gm_ref, _ = ref.loss_and_grad(H0, E0, dt, K, probes, target)
    s = MaxwellSolver(mesh, eps=eps)
    s.set_materials(eps)
    s.set_source(amp, (0, 0, 1), 0.05, 0.03)
    s.set_state(H0, E0)
    s.adjoint_forward(dt, K, every)
    E = s.get_state()[1]
    diff = E[2, probes] - target
    L = float(np.sum(diff * diff))
    assert abs(L - L_ref) <= 1e-9 * L_ref
    lamE = np.zeros((3, n))
    np.add.at(lamE[2], probes, 2 * diff)
    ge, gm = s.adjoint_backward(np.zeros((3, n)), lamE)
    assert np.max(np.abs(ge - ge_ref)) <= 1e-9 * np.max(np.abs(ge_ref))
    assert np.max(np.abs(gm - gm_ref)) <= 1e-9 * np.max(np.abs(gm_ref))
    # the final forward state survives the backward sweep
    assert np.array_equal(s.get_state()[1], E)
    # a second sweep over the same forward, and windows of 3 steps inside
    # the 7-step segments, give the same gradient bit for bit
    ge2, gm2 = s.adjoint_backward(np.zeros((3, n)), lamE)
    assert np.array_equal(ge2, ge) and np.array_equal(gm2, gm)
    s.adjoint_set_store(3)
    s.time = 0.0
    s.set_state(H0, E0)
    s.adjoint_forward(dt, K, every)
    info = s.adjoint_info()
    assert info["window"] == 3 and info["segments"] == 3
    ge3, gm3 = s.adjoint_backward(np.zeros((3, n)), lamE)
    assert np.array_equal(ge3, ge) and np.array_equal(gm3, gm)
    s.adjoint_set_store(0)
    for _ in range(3):
        d = rng.standard_normal(n)
        lp, _ = _gpu_loss(mesh, eps + 1e-6 * d, amp, probes, target, dt, K, H0, E0, s)
        lm, _ = _gpu_loss(mesh, eps - 1e-6 * d, amp, probes, target, dt, K, H0, E0, s)
        fd = (lp - lm) / 2e-6
        assert abs(fd - np.dot(ge, d)) <= 1e-6 * abs(fd)
    s.close()


def test_torch_probe_fields_autograd(oracle):
    """ProbeFields under torch autograd: an Ez probe loss written in torch
    gives the oracle's eps and mu gradients, and a loss mixing H and E
    probes agrees with central differences (h = 1e-6) of the GPU forward
    in eps and in mu."""
    torch = pytest.importorskip("torch")
    from nekcem_amd.adjoint import ProbeFields, ProbeProblem, _run_forward
    mesh, eps, amp, probes, target, ref, dt, K, every, H0, E0, rng = _config4_small(oracle)
    n = mesh.npts
    print("[case] torch ProbeFields 2x2x2 N=5 PEC, 20 steps, checkpoint every 7", flush=True)
    L_ref, ge_ref, gm_ref, _ = ref.loss_and_grad(H0, E0, dt, K, probes, target)
    s = MaxwellSolver(mesh, eps=eps)
    s.set_source(amp, (0, 0, 1), 0.05, 0.03)
    prob = ProbeProblem(solver=s, H0=H0, E0=E0, dt=dt, nsteps=K, every=every, probes=probes)
    dev = "cuda:0" if torch.cuda.is_available() else "cpu"
    e = torch.tensor(eps, dtype=torch.float64, device=dev, requires_grad=True)
    m = torch.ones(n, dtype=torch.float64, device=dev, requires_grad=True)
    tgt = torch.tensor(target, dtype=torch.float64, device=dev)
    out = ProbeFields.apply(e, m, prob)
    assert out.shape == (6, len(probes))
    loss = ((out[5] - tgt) ** 2).sum()
    assert abs(loss.item() - L_ref) <= 1e-9 * L_ref
    loss.backward()
    assert np.max(np.abs(e.grad.cpu().numpy() - ge_ref)) <= 1e-9 * np.max(np.abs(ge_ref))
    assert np.max(np.abs(m.grad.cpu().numpy() - gm_ref)) <= 1e-9 * np.max(np.abs(gm_ref))

    # a loss over H and E probes, checked against FD in eps and in mu
    e2 = torch.tensor(eps, dtype=torch.float64, device=dev, requires_grad=True)
    m2 = torch.ones(n, dtype=torch.float64, device=dev, requires_grad=True)
    w = torch.tensor(rng.standard_normal((6, len(probes))), dtype=torch.float64, device=dev)
    out2 = ProbeFields.apply(e2, m2, prob)
    (w * out2 ** 2).sum().backward()
    wn = w.cpu().numpy()

    def lossf(ev, mv):
        return float(np.sum(wn * _run_forward(prob, ev, mv) ** 2))

    one = np.ones(n)
    for grad, which in ((e2.grad, "eps"), (m2.grad, "mu")):
        d = rng.standard_normal(n)
        if which == "eps":
            fd = (lossf(eps + 1e-6 * d, one) - lossf(eps - 1e-6 * d, one)) / 2e-6
        else:
            fd = (lossf(eps, one + 1e-6 * d) - lossf(eps, one - 1e-6 * d)) / 2e-6
        assert abs(fd - float(np.dot(grad.cpu().numpy(), d))) <= 1e-6 * abs(fd), which
    s.close()


def test_drude_adjoint_matches_oracle_and_fd(oracle):
    """Drude region (central elements, per-element gamma and omega_p) with a
    Gaussian source, 40 steps checkpointed every 7: GPU gradients with
    respect to per-element gamma and omega_p and per-node eps equal the
    oracle's adjoint (1e-9) and central FD (h = 1e-6) of the GPU forward
    (1e-6)."""
    rng = np.random.default_rng(6)
    nel, N, K, every = (3, 3, 3), 4, 40, 7
    print(f"[case] Drude adjoint nel={nel} N={N} K={K} every={every}", flush=True)
    lo, hi = (0, 0, 0), (1, 1, 1)
    mesh = BoxMesh(nel, N, lo, hi, ("PEC",) * 3)
    n, n3 = mesh.npts, mesh.n1 ** 3
    x, y, z = mesh.coords()
    cen = [v.reshape(-1, n3).mean(axis=1) for v in (x, y, z)]
    inside = np.ones(mesh.nelt, bool)
    for v in cen:
        inside &= np.abs(v - 0.5) < 0.25
    elems = np.nonzero(inside)[0]
    ne = len(elems)
    assert ne >= 1
    dindex = (elems[:, None] * n3 + np.arange(n3)[None, :]).ravel()
    # omega_p ~ 2-3 and 40 steps: the loss depends on gamma strongly enough
    # for central differences at h = 1e-6 to resolve 1e-6 relative
    gam, wp = rng.uniform(0.05, 0.2, ne), rng.uniform(2.0, 3.0, ne)
    eps = 1.5 + 0.2 * np.sin(np.pi * x) * np.cos(np.pi * y)
    amp = np.exp(-((x - .5) ** 2 + (y - .5) ** 2 + (z - .5) ** 2) / (2 * .15 ** 2)) * mesh.bmn
    probes = rng.choice(n, 8, replace=False)
    dt = mesh.dt(0.25)
    H0, E0 = np.zeros((3, n)), np.zeros((3, n))

    ref = oracle.BoxCase(oracle.BoxSpec(nel=nel, N=N, lo=lo, hi=hi, bc=("PEC",) * 3), eps=eps)
    ref.set_source(amp, (0, 0, 1), 0.05, 0.03)

    def ref_params(g, w):
        ref.set_drude(dindex, np.repeat(g, n3), np.repeat(w * w, n3))

    ref_params(gam * 1.3, wp * 0.8)
    target = ref.run(H0, E0, dt, K)["E"][2, probes].copy()
    ref_params(gam, wp)
    L_ref, ge_ref, _, _, ga_ref, gb_ref = ref.loss_and_grad(H0, E0, dt, K, probes, target, drude=True)

    s = MaxwellSolver(mesh, eps=eps)
    s.set_materials(eps)
    s.set_source(amp, (0, 0, 1), 0.05, 0.03)
    s.set_drude(dindex, np.repeat(gam, n3), np.repeat(wp * wp, n3))
    s.time = 0.0
    s.set_state(H0, E0)
    s.adjoint_forward(dt, K, every)
    E = s.get_state()[1]
    diff = E[2, probes] - target
    L = float(np.sum(diff * diff))
    assert abs(L - L_ref) <= 1e-9 * L_ref
    lamE = np.zeros((3, n))
    np.add.at(lamE[2], probes, 2 * diff)
    ge, _ = s.adjoint_backward(np.zeros((3, n)), lamE)
    ga, gb = s.adjoint_drude_grad()
    assert np.array_equal(s.get_state()[1], E)
    assert np.max(np.abs(ge - ge_ref)) <= 1e-9 * np.max(np.abs(ge_ref))
    assert np.max(np.abs(ga - ga_ref)) <= 1e-9 * np.max(np.abs(ga_ref))
    assert np.max(np.abs(gb - gb_ref)) <= 1e-9 * np.max(np.abs(gb_ref))
    g_gam = ga.reshape(ne, n3).sum(axis=1)
    g_wp = 2 * wp * gb.reshape(ne, n3).sum(axis=1)
    # windows of 2 steps inside the 7-step segments: the same bits (from the
    # same initial state: set_drude zeroes the current again)
    s.adjoint_set_store(2)
    s.set_drude(dindex, np.repeat(gam, n3), np.repeat(wp * wp, n3))
    s.time = 0.0
    s.set_state(H0, E0)
    s.adjoint_forward(dt, K, every)
    assert s.adjoint_info()["window"] == 2
    ge2, _ = s.adjoint_backward(np.zeros((3, n)), lamE)
    ga2, gb2 = s.adjoint_drude_grad()
    assert np.array_equal(ge2, ge) and np.array_equal(ga2, ga) and np.array_equal(gb2, gb)
    # without the eps/mu gradients (no residual store): the same Drude bits
    s.adjoint_set_store(0)
    s.adjoint_set_grads(False)
    s.set_drude(dindex, np.repeat(gam, n3), np.repeat(wp * wp, n3))
    s.time = 0.0
    s.set_state(H0, E0)
    s.adjoint_forward(dt, K, every)
    ge3, _ = s.adjoint_backward(np.zeros((3, n)), lamE)
    ga3, gb3 = s.adjoint_drude_grad()
    assert np.array_equal(ga3, ga) and np.array_equal(gb3, gb) and not np.any(ge3)
    s.adjoint_set_grads(True)

    def loss(g, w):
        s.set_drude(dindex, np.repeat(g, n3), np.repeat(w * w, n3))
        s.time = 0.0
        s.set_state(H0, E0)
        s.step(dt, K)
        d = s.get_state()[1][2, probes] - target
        return float(np.sum(d * d))

    h = 1e-6
    for which in ("gamma", "wp"):
        d = rng.standard_normal(ne)
        if which == "gamma":
            fd = (loss(gam + h * d, wp) - loss(gam - h * d, wp)) / (2 * h)
            an = np.dot(g_gam, d)
        else:
            fd = (loss(gam, wp + h * d) - loss(gam, wp - h * d)) / (2 * h)
            an = np.dot(g_wp, d)
        assert abs(fd - an) <= 1e-6 * abs(fd), which

    # the same gradient through the torch harness
    torch = pytest.importorskip("torch")
    from nekcem_amd.adjoint import DrudeProbeFields, DrudeProbeProblem
    s.adjoint_set_store(0)
    owner = np.repeat(np.arange(ne), n3)
    prob = DrudeProbeProblem(solver=s, H0=H0, E0=E0, dt=dt, nsteps=K, every=every, probes=probes,
                             dindex=dindex, owner=owner)
    tg = torch.tensor(gam, dtype=torch.float64, requires_grad=True)
    tw = torch.tensor(wp, dtype=torch.float64, requires_grad=True)
    out = DrudeProbeFields.apply(tg, tw, prob)
    ((out[5] - torch.tensor(target)) ** 2).sum().backward()
    assert np.allclose(tg.grad.numpy(), g_gam, rtol=1e-12, atol=1e-12 * np.max(np.abs(g_gam)))
    assert np.allclose(tw.grad.numpy(), g_wp, rtol=1e-12, atol=1e-12 * np.max(np.abs(g_wp)))
    s.close()
