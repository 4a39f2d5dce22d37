"""Parity of the HIP path (libnekcem_hip.so via its C-ABI) with the CPU
oracle on the same mesh, materials and initial state.

Tolerances: the contract is 1e-11 relative L2 per field (bm1-weighted, the
cem_error norm) after the stated step count; single residual evaluations
are held to 1e-11 of the field's max (random input; the two sides build
their D matrices independently, which differ by ~1e-15 relative).  The HIP kernel uses per-element
constant metrics and the strong-form arrangement of the same arithmetic, so
agreement is to round-off, not bitwise.
"""
import numpy as np
import pytest

from nekcem_amd.maxwell import MaxwellSolver, cem_error
from nekcem_amd.mesh import BoxMesh
from nekcem_amd import usr as cases

pytestmark = pytest.mark.gpu


def _norm(u, bm1):
    return float(np.sqrt(np.sum(bm1 * u * u)))


def parity(gH, gE, oH, oE, bm1, floor=1e-6):
    """Largest per-field relative L2 (bm1-weighted) difference.  A field whose
    own norm is below `floor` of the largest field's norm (e.g. a component
    that is zero in the exact solution and holds only truncation error) is
    measured against the largest field's norm instead."""
    g = np.concatenate([gH, gE])
    o = np.concatenate([oH, oE])
    norms = [_norm(o[c], bm1) for c in range(6)]
    big = max(norms)
    return max(_norm(g[c] - o[c], bm1) / (norms[c] if norms[c] >= floor * big else big)
               for c in range(6))


def _pair(oracle, nel, N, bc=("P", "P", "P"), lo=(0, 0, 0), hi=(2 * np.pi,) * 3,
          eps=None, mu=None, upwind=True, graph=True):
    print(f"[case] nel={nel} N={N} bc={bc} upwind={upwind} graph={graph}", flush=True)
    mesh = BoxMesh(nel, N, lo, hi, bc)
    spec = oracle.BoxSpec(nel=nel, N=N, lo=lo, hi=hi, bc=bc, upwind=upwind)
    ref = oracle.BoxCase(spec, eps=eps, mu=mu)
    gpu = MaxwellSolver(mesh, eps=eps, mu=mu, ifupwind=upwind, graph=graph)
    return mesh, ref, gpu


RESIDUAL_CASES = [
    dict(nel=(3, 2, 4), N=5, bc=("P", "P", "P")),
    dict(nel=(2, 2, 2), N=7, bc=("PEC", "PEC", "PEC"), lo=(-1, -1, -1), hi=(1, 1, 1)),
    dict(nel=(2, 3, 2), N=3, bc=("P", "PEC", "P")),
    dict(nel=(2, 2, 3), N=6, bc=("PEC", "P", "P")),
    dict(nel=(2, 2, 2), N=8, bc=("P", "P", "PEC")),
    dict(nel=(1, 2, 2), N=9, bc=("P", "P", "P")),
    dict(nel=(2, 1, 2), N=4, bc=("P", "P", "P"), upwind=False),
    # one element: every periodic face is paired with the element's own
    # opposite face
    dict(nel=(1, 1, 1), N=6, bc=("P", "P", "P")),
    dict(nel=(1, 1, 1), N=5, bc=("PEC", "P", "PEC")),
]


@pytest.mark.parametrize("case", RESIDUAL_CASES, ids=lambda c: f"N{c['N']}-{'-'.join(c['bc'])}")
@pytest.mark.parametrize("material", ["vacuum", "element", "node"])
def test_residual_parity(oracle, case, material):
    rng = np.random.default_rng(7)
    n1 = case["N"] + 1
    nelt = int(np.prod(case["nel"]))
    npts = nelt * n1 ** 3
    eps = mu = None
    if material == "element":
        eps = np.repeat(rng.uniform(1, 4, nelt), n1 ** 3)
        mu = np.repeat(rng.uniform(1, 2, nelt), n1 ** 3)
    elif material == "node":
        eps = rng.uniform(1, 4, npts)
        mu = rng.uniform(1, 2, npts)
    mesh, ref, gpu = _pair(oracle, eps=eps, mu=mu, **case)
    assert gpu.info()["mat_mode"] == (1 if material == "node" else 0)
    H = rng.standard_normal((3, npts))
    E = rng.standard_normal((3, npts))
    gpu.set_state(H, E)
    rH, rE = gpu.residual()
    oH, oE = ref.residual(H, E)
    for a, b in ((rH, oH), (rE, oE)):
        for c in range(3):
            scale = np.max(np.abs(b[c]))
            assert np.max(np.abs(a[c] - b[c])) <= 1e-11 * scale, (c, np.max(np.abs(a[c] - b[c])), scale)
    gpu.close()


def test_config1_100_steps_parity(oracle):
    """4^3 periodic box on [0,2pi]^3, N=7, vacuum, upwind, (1,1,1) mode,
    CFL 0.25, 100 steps: GPU vs oracle within 1e-11 relative L2 per field,
    and both at the analytic solution to the same error."""
    mesh, ref, gpu = _pair(oracle, (4, 4, 4), 7)
    x, y, z = mesh.coords()
    H, E = cases.boxper_solution(0.0, x, y, z)
    dt = mesh.dt(0.25)
    assert abs(dt - ref.dt_from_cfl(0.25)) < 1e-15
    gpu.set_state(H, E)
    gpu.step(dt, 100)
    gH, gE = gpu.get_state()
    out = ref.run(H, E, dt, 100)
    Hs, Es = cases.boxper_solution(out["time"], x, y, z)
    assert parity(gH, gE, out["H"], out["E"], mesh.bmn) <= 1e-11
    for g, o, s in ((gH, out["H"], Hs), (gE, out["E"], Es)):
        for c in range(3):
            eg, eo = cem_error(g[c], s[c], mesh.bmn)[0], ref.cem_error(o[c], s[c])[0]
            assert abs(eg - eo) <= 1e-3 * eo + 1e-15
    assert abs(gpu.time - out["time"]) < 1e-13
    gpu.close()


def test_3dboxper_analytic_on_gpu(oracle):
    """tests/3dboxper on the HIP path: 4x4x8, N=8, dt=2e-4, 50 steps, checked
    like its userchk (istep <= 10 and every iocomm=50) at L2 <= 5e-10,
    Linf <= 5e-9."""
    mesh = BoxMesh((4, 4, 8), 8)
    gpu = MaxwellSolver(mesh)
    gpu.cem_maxwell_init_fields(cases.boxper_solution)
    gpu.set_dt(-0.0002)
    x, y, z = mesh.coords()
    seen = []

    def userchk(s):
        H, E = s.get_state()
        Hs, Es = cases.boxper_solution(s.time, x, y, z)
        for u, ex in ((H, Hs), (E, Es)):
            for c in range(3):
                l2, li = cem_error(u[c], ex[c], mesh.bmn)
                if l2 > 5e-10 or li > 5e-9:
                    return False
        seen.append(s.istep)
        return True

    gpu.time_advancing_pde(50, userchk, chk_every=50)
    assert seen == list(range(1, 11)) + [50]
    gpu.close()


def test_3dboxpec_analytic_on_gpu(oracle):
    """tests/3dboxpec on the HIP path: 3x3x3 PEC box, N=8, dt=0.005, 1000
    steps, L2 <= 5e-8, Linf <= 5e-7 (checked every iocomm=10 steps)."""
    mesh = BoxMesh((3, 3, 3), 8, (-1, -1, -1), (1, 1, 1), ("PEC",) * 3)
    gpu = MaxwellSolver(mesh)
    gpu.cem_maxwell_init_fields(cases.boxpec_solution)
    gpu.set_dt(-0.005)
    x, y, z = mesh.coords()
    worst = [0.0]

    def userchk(s):
        H, E = s.get_state()
        Hs, Es = cases.boxpec_solution(s.time, x, y, z)
        for u, ex in ((H, Hs), (E, Es)):
            for c in range(3):
                l2, li = cem_error(u[c], ex[c], mesh.bmn)
                worst[0] = max(worst[0], l2)
                if l2 > 5e-8 or li > 5e-7:
                    return False
        return True

    gpu.time_advancing_pde(1000, userchk, chk_every=100)
    assert 0 < worst[0] <= 5e-8
    gpu.close()


def test_graph_replay_matches_direct_launches(oracle):
    mesh = BoxMesh((3, 3, 3), 7)
    rng = np.random.default_rng(11)
    eps = np.repeat(rng.uniform(1, 4, mesh.nelt), mesh.n1 ** 3)
    H = rng.standard_normal((3, mesh.npts))
    E = rng.standard_normal((3, mesh.npts))
    outs = []
    for graph in (True, False):
        g = MaxwellSolver(mesh, eps=eps, graph=graph)
        g.set_state(H, E)
        for _ in range(3):          # replays with host work in between
            g.step(1e-3, 2)
            g.get_state()
        outs.append(g.get_state(with_k=True))
        g.close()
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)


def test_blowup_detected():
    mesh = BoxMesh((2, 2, 2), 5)
    g = MaxwellSolver(mesh)
    H = np.zeros((3, mesh.npts))
    E = np.ones((3, mesh.npts))
    g.set_state(H, E)
    assert abs(g.check_finite() - 1.0) < 1e-15
    E[1, 17] = np.nan
    g.set_state(H, E)
    from nekcem_amd._lib import NkcError, NKC_ERR_BLOWUP
    with pytest.raises(NkcError) as ei:
        g.check_finite()
    assert ei.value.code == NKC_ERR_BLOWUP
    g.close()


def test_config2_shape_reduced_parity(oracle):
    """The 32^3 workload's physics on an 8^3 box: per-element eps ~ U[1,4]
    (seed 1), 8 seeded plane waves, upwind, CFL 0.25, 20 steps."""
    nel = (8, 8, 8)
    eps = np.repeat(cases.element_eps(512, seed=1), 512)
    mesh, ref, gpu = _pair(oracle, nel, 7, eps=eps)
    x, y, z = mesh.coords()
    H, E = cases.plane_waves(x, y, z, seed=2)
    dt = mesh.dt(0.25)
    gpu.set_state(H, E)
    gpu.step(dt, 20)
    gH, gE = gpu.get_state()
    out = ref.run(H, E, dt, 20)
    assert parity(gH, gE, out["H"], out["E"], mesh.bmn) <= 1e-11
    gpu.close()


def test_source_and_drude_residual_parity(oracle):
    """Gaussian source + Drude region in one residual evaluation (random
    state and random J)."""
    rng = np.random.default_rng(21)
    nel, N = (3, 2, 2), 5
    mesh, ref, gpu = _pair(oracle, nel, N, bc=("PEC", "P", "PEC"),
                           eps=np.repeat(rng.uniform(1, 3, 12), 216))
    npts = mesh.npts
    x, y, z = mesh.coords()
    amp = np.exp(-((x - 3) ** 2 + (y - 2) ** 2 + (z - 3) ** 2) / 0.5) * mesh.bmn
    amp[np.abs(x - 3) > 2] = 0.0          # some elements carry no source
    dindex = np.nonzero((x > 1.5) & (x < 5) & (z < 4))[0]
    a = rng.uniform(0.05, 0.2, len(dindex))
    b = rng.uniform(0.5, 2.0, len(dindex))
    J = rng.standard_normal((3, len(dindex)))
    H = rng.standard_normal((3, npts))
    E = rng.standard_normal((3, npts))
    ref.set_source(amp, direction=(0.3, -0.2, 1.0), t0=0.4, tw=0.15)
    ref.set_drude(dindex, a, b)
    gpu.set_source(amp, direction=(0.3, -0.2, 1.0), t0=0.4, tw=0.15)
    gpu.set_drude(dindex, a, b, J=J)
    gpu.set_state(H, E)
    for t in (0.35, 0.5):
        rH, rE = gpu.residual(t)
        oH, oE, oJ = ref.residual(H, E, t=t, J=J.reshape(-1))
        gJ = gpu.get_drude()[2]
        for a_, b_ in ((rH, oH), (rE, oE), (gJ, oJ.reshape(3, -1))):
            for c in range(3):
                assert np.max(np.abs(a_[c] - b_[c])) <= 1e-11 * np.max(np.abs(b_[c]))
    gpu.close()


def test_source_and_drude_steps_parity(oracle):
    """40 RK steps with the source switching on and a Drude block: fields and
    J within 1e-11 relative L2 of the oracle; device time matches."""
    rng = np.random.default_rng(5)
    nel, N = (3, 3, 3), 6
    lo, hi = (0, 0, 0), (1, 1, 1)
    mesh, ref, gpu = _pair(oracle, nel, N, bc=("PEC",) * 3, lo=lo, hi=hi)
    x, y, z = mesh.coords()
    amp = np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2) / (2 * 0.1 ** 2)) * mesh.bmn
    inner = (np.abs(x - 0.5) < 0.17) & (np.abs(y - 0.5) < 0.17) & (np.abs(z - 0.5) < 0.17)
    dindex = np.nonzero(inner)[0]
    a = np.full(len(dindex), 0.1)
    b = np.full(len(dindex), 4.0)
    for side in (ref, gpu):
        side.set_source(amp, direction=(0, 0, 1), t0=0.05, tw=0.02)
        side.set_drude(dindex, a, b)
    H = np.zeros((3, mesh.npts))
    E = np.zeros((3, mesh.npts))
    dt = mesh.dt(0.25)
    gpu.set_state(H, E)
    gpu.step(dt, 40)
    gH, gE = gpu.get_state()
    gJ = gpu.get_drude()[0]
    out = ref.run(H, E, dt, 40)
    assert np.max(np.abs(out["E"])) > 1e-3       # the pulse did fire
    assert parity(gH, gE, out["H"], out["E"], mesh.bmn) <= 1e-11
    oJ = out["J"].reshape(3, -1)
    assert np.sqrt(np.sum((gJ - oJ) ** 2)) <= 1e-11 * np.sqrt(np.sum(oJ ** 2))
    assert gpu.time == out["time"]
    gpu.close()


def test_lorentz_residual_and_steps_parity(oracle):
    """A Lorentz block (cem_maxwell_lorentz) beside a Drude block, with the
    source: one residual (resH, resE and resJ/resP of a random state) within
    1e-11 of the oracle, and 40 RK steps of fields, J and P within 1e-11
    relative L2."""
    print("[case] Lorentz + Drude + source, 3x3x3 N=6 PEC", flush=True)
    rng = np.random.default_rng(7)
    nel, N = (3, 3, 3), 6
    mesh, ref, gpu = _pair(oracle, nel, N, bc=("PEC",) * 3, lo=(0, 0, 0), hi=(1, 1, 1))
    x, y, z = mesh.coords()
    n = mesh.npts
    amp = np.exp(-((x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2) / (2 * 0.1 ** 2)) * mesh.bmn
    lblock = np.nonzero((x > 0.34) & (y < 0.66) & (z > 0.34))[0]
    dblock = np.nonzero((x < 0.32) & (y > 0.68))[0]
    la, lb, lc = rng.uniform(0.05, 0.3, len(lblock)), rng.uniform(1, 3, len(lblock)), rng.uniform(0.5, 2, len(lblock))
    for side in (ref, gpu):
        side.set_source(amp, direction=(0, 1, 0), t0=0.05, tw=0.02)
        side.set_drude(dblock, 0.1, 2.0)
        side.set_lorentz(lblock, la, lb, lc)
    # residual of a random state with a random Lorentz state
    H, E = rng.standard_normal((3, n)), rng.standard_normal((3, n))
    L0 = rng.standard_normal((6, len(lblock)))
    gpu.set_lorentz(lblock, la, lb, lc, L=L0)
    gpu.set_state(H, E)
    rH, rE = gpu.residual(0.04)
    rL = gpu.get_lorentz()[2]
    oH, oE, oJ, oL = ref.residual(H, E, t=0.04, J=np.zeros(3 * len(dblock)), L=L0.reshape(-1))
    for a_, b_ in ((rH, oH), (rE, oE), (rL, oL.reshape(6, -1))):
        assert np.max(np.abs(a_ - b_)) <= 1e-11 * np.max(np.abs(b_))
    # 40 steps from rest
    gpu.set_lorentz(lblock, la, lb, lc)
    H0, E0 = np.zeros((3, n)), np.zeros((3, n))
    dt = mesh.dt(0.25)
    gpu.time = 0.0
    gpu.set_state(H0, E0)
    gpu.step(dt, 40)
    gH, gE = gpu.get_state()
    gL = gpu.get_lorentz()[0]
    out = ref.run(H0, E0, dt, 40)
    assert np.max(np.abs(out["E"])) > 1e-3
    assert parity(gH, gE, out["H"], out["E"], mesh.bmn) <= 1e-11
    oLs = out["L"].reshape(6, -1)
    assert np.max(np.abs(oLs)) > 0
    assert np.sqrt(np.sum((gL - oLs) ** 2)) <= 1e-11 * np.sqrt(np.sum(oLs ** 2))
    with pytest.raises(Exception):
        gpu.adjoint_forward(dt, 4, 2)          # the Lorentz adjoint is refused
    gpu.close()


@pytest.mark.parametrize("N", [4, 5, 6, 7, 8, 9])
def test_lane_layouts_bit_identical_and_tuner(N):
    """All stage-kernel layouts built for the order give the same bits
    (graph replay, 3 steps, per-element eps, periodic x, PEC y/z): the four
    one-element layouts, at N <= 5 the two two-element ones on an odd
    element count (the last workgroup's second block is empty), at N >= 7
    the two six-role ones; the tuner times them and picks one; a layout not
    built for the order is refused."""
    print(f"[case] lane layouts N={N}", flush=True)
    from nekcem_amd._lib import NkcError
    from nekcem_amd.maxwell import MaxwellSolver, NLAYOUT
    from nekcem_amd.mesh import BoxMesh
    rng = np.random.default_rng(N)
    nel = (3, 3, 1) if N <= 5 else (3, 2, 2)
    mesh = BoxMesh(nel, N, (0, 0, 0), (1.0, 0.8, 1.2), ("P", "PEC", "PEC"))
    n = mesh.npts
    eps = np.repeat(rng.uniform(1, 4, mesh.nelt), mesh.n1 ** 3)
    H, E = rng.standard_normal((3, n)), rng.standard_normal((3, n))
    dt = mesh.dt(0.25)
    layouts = list(range(6)) if N <= 5 else [0, 1, 2, 3] + ([6, 7] if N >= 7 else [])
    out = []
    for layout in layouts:
        s = MaxwellSolver(mesh, eps=eps)
        s.set_layout(layout)
        s.set_state(H, E)
        s.step(dt, 3)
        out.append(s.get_state())
        if layout == layouts[-1]:
            t = s.tune()
            assert np.isfinite(t["us"][0]) and np.isfinite(t["us"][2]) and t["layout"] in layouts
            assert all(np.isfinite(t["us"][p]) == (N <= 5) for p in (4, 5))
            assert all(np.isfinite(t["us"][p]) == (N >= 7) for p in (6, 7))
            assert s.get_state()[1].tobytes() == out[-1][1].tobytes()   # tuning leaves the state
            with pytest.raises(NkcError, match="not built"):
                s.set_layout(4 if N > 5 else 6)
        s.close()
    for layout, o in zip(layouts[1:], out[1:]):
        for a, b in zip(out[0], o):
            assert np.array_equal(a, b), (layout, float(np.max(np.abs(a - b))))


@pytest.mark.parametrize("N", [4, 7, 9])
def test_curl_parity(oracle, N):
    """nkc_curl (the weighted volume curl alone) against the oracle's
    cem_maxwell volume term on a random state, PEC/periodic box with
    per-element materials (which the curl does not see): <= 1e-11 of the
    field's max."""
    rng = np.random.default_rng(N)
    nel = (2, 3, 2)
    n1 = N + 1
    nelt = int(np.prod(nel))
    eps = np.repeat(rng.uniform(1, 4, nelt), n1 ** 3)
    mesh, ref, gpu = _pair(oracle, nel, N, bc=("P", "PEC", "P"), lo=(0, -1, 0), hi=(1, 1, 2), eps=eps)
    H = rng.standard_normal((3, mesh.npts))
    E = rng.standard_normal((3, mesh.npts))
    gpu.set_state(H, E)
    rH, rE = gpu.curl()
    oH, oE = ref.curl(H, E)
    for a, b in ((rH, oH), (rE, oE)):
        assert np.max(np.abs(a - b)) <= 1e-11 * np.max(np.abs(b))


def test_one_element_steps_every_layout(oracle):
    """A one-element periodic box (N = 5) stepped under every built layout:
    with the two-element layouts the single workgroup carries one valid
    element and one spare block.  5 steps against the oracle, 1e-11."""
    from nekcem_amd.maxwell import NLAYOUT
    rng = np.random.default_rng(3)
    done = 0
    for layout in range(NLAYOUT):
        mesh, ref, gpu = _pair(oracle, (1, 1, 1), 5)
        try:
            gpu.set_layout(layout)
        except Exception as err:     # a layout not built for this order
            assert "not built" in str(err)
            gpu.close()
            continue
        x, y, z = mesh.coords()
        H, E = cases.boxper_solution(0.0, x, y, z)
        H = H + 0.1 * rng.standard_normal(H.shape)
        gpu.set_state(H, E)
        dt = mesh.dt(0.25)
        gpu.step(dt, 5)
        gH, gE = gpu.get_state()
        out = ref.run(H, E, dt, 5)
        assert parity(gH, gE, out["H"], out["E"], mesh.bmn) <= 1e-11, layout
        gpu.close()
        done += 1
    assert done == 6            # N = 5 builds layouts 0..5
