"""The reference's own case files on the HIP path: 3dboxpec from its .rea
and 3dboxper from its .rea + .re2, built by nkc_create_hex and checked like
their userchk with the on-device cem_error."""
import os

import numpy as np
import pytest

from nekcem_amd.maxwell import MaxwellSolver
from nekcem_amd.mesh import BoxMesh
from nekcem_amd import usr

pytestmark = pytest.mark.gpu
CASES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cases")


def _rescale(lo, hi):
    """usrdat2 of the two cases: map each axis' extent onto [lo, hi]."""
    def f(xc, yc, zc):
        out = []
        for c in (xc, yc, zc):
            a, b = c.min(), c.max()
            out.append(lo + (hi - lo) * (c - a) / (b - a))
        return out
    return f


def test_3dboxpec_from_rea():
    """3dboxpec.rea (SIZE lxi = 8): 1000 steps of dt = 0.005 from the
    analytic mode, L2 <= 5e-8 and Linf <= 5e-7 every 100 steps."""
    print("[case] 3dboxpec from .rea, N=8, 1000 steps", flush=True)
    s = MaxwellSolver.from_rea(os.path.join(CASES, "3dboxpec.rea"), 8, usrdat2=_rescale(-1.0, 1.0))
    s.init_solution("3dboxpec", 0.0)
    dt = s.mesh.dt(s.param[12])
    assert dt == 0.005
    worst = 0.0
    for _ in range(int(s.param[11]) // 100):
        s.step(dt, 100)
        l2, li = s.cem_error_device("3dboxpec")
        worst = max(worst, float(np.max(l2)))
        assert np.all(l2 <= 5e-8) and np.all(li <= 5e-7), (s.istep, l2, li)
    assert worst > 0
    s.close()


def test_3dboxper_from_re2():
    """3dboxper.rea + .re2 (lxi = 8): 50 steps of dt = 2e-4, checked at steps
    1..10 and 50 like its userchk, L2 <= 5e-10, Linf <= 5e-9; and the same
    state as the box built from extents to 1e-13."""
    print("[case] 3dboxper from .rea + .re2, N=8, 50 steps", flush=True)
    s = MaxwellSolver.from_rea(os.path.join(CASES, "3dboxper.rea"), 8, usrdat2=_rescale(0.0, 2 * np.pi))
    s.init_solution("3dboxper", 0.0)
    dt = s.mesh.dt(s.param[12])
    for step in range(1, int(s.param[11]) + 1):
        s.step(dt, 1)
        if step <= 10 or step == 50:
            l2, li = s.cem_error_device("3dboxper")
            assert np.all(l2 <= 5e-10) and np.all(li <= 5e-9), (step, l2, li)
    H, E = s.get_state()
    ref = MaxwellSolver(BoxMesh((4, 4, 8), 8))
    ref.init_solution("3dboxper", 0.0)
    ref.step(dt, 50)
    Hr, Er = ref.get_state()
    scale = max(np.max(np.abs(Hr)), np.max(np.abs(Er)))
    assert np.max(np.abs(H - Hr)) <= 1e-13 * scale and np.max(np.abs(E - Er)) <= 1e-13 * scale
    s.close()
    ref.close()


def _graded_hex(verts, bc):
    """Corners, codes and parameters (the reader arrays) of a graded box
    lattice: per-axis vertex coordinates verts, 'P' or 'PEC' per axis."""
    nel = [len(v) - 1 for v in verts]
    nelt = nel[0] * nel[1] * nel[2]
    xc, yc, zc = np.zeros((nelt, 8)), np.zeros((nelt, 8)), np.zeros((nelt, 8))
    cbc = np.full((nelt, 6), "E  ", dtype="<U3")
    par = np.zeros((nelt, 6, 5))
    # Nek face f (1-based) -> (axis, side) and its opposite face
    fax = {1: (1, 0), 2: (0, 1), 3: (1, 1), 4: (0, 0), 5: (2, 0), 6: (2, 1)}
    opp = {1: 3, 3: 1, 2: 4, 4: 2, 5: 6, 6: 5}
    for e in range(nelt):
        c = [e % nel[0], (e // nel[0]) % nel[1], e // (nel[0] * nel[1])]
        x0, x1 = verts[0][c[0]], verts[0][c[0] + 1]
        y0, y1 = verts[1][c[1]], verts[1][c[1] + 1]
        z0, z1 = verts[2][c[2]], verts[2][c[2] + 1]
        xc[e] = [x0, x1, x1, x0, x0, x1, x1, x0]
        yc[e] = [y0, y0, y1, y1, y0, y0, y1, y1]
        zc[e] = [z0] * 4 + [z1] * 4
        for f in range(1, 7):
            ax, sd = fax[f]
            d = list(c)
            d[ax] += 1 if sd else -1
            inside = 0 <= d[ax] < nel[ax]
            if not inside and bc[ax] == "PEC":
                cbc[e, f - 1] = "PEC"
                continue
            d[ax] %= nel[ax]
            cbc[e, f - 1] = "E  " if inside else "P  "
            par[e, f - 1, 0] = d[0] + nel[0] * (d[1] + nel[1] * d[2]) + 1
            par[e, f - 1, 1] = opp[f]
    return xc, yc, zc, cbc, par


@pytest.mark.parametrize("upwind", [True, False])
def test_graded_hex_mesh_parity(oracle, upwind):
    """nkc_create_hex on a graded lattice (element sizes differing per axis
    and per element), periodic in x, PEC in y and z, per-element eps: one
    residual and 20 RK steps within 1e-11 of the oracle built on the same
    vertices."""
    print(f"[case] graded hex mesh upwind={upwind}", flush=True)
    from nekcem_amd import usr
    from oracle.oracle import BoxSpec
    verts = (np.array([0.0, 0.3, 0.5, 1.1]), np.array([0.0, 0.5, 0.7]), np.array([0.0, 0.2, 0.9, 1.0]))
    bc = ("P", "PEC", "PEC")
    N = 5
    xc, yc, zc, cbc, par = _graded_hex(verts, bc)
    nelt = xc.shape[0]
    rng = np.random.default_rng(9)
    eps_e = rng.uniform(1, 4, nelt)
    s = MaxwellSolver.from_corners(N, xc, yc, zc, cbc, par, eps_elem=eps_e, ifupwind=upwind)
    n1 = N + 1
    ref = oracle.BoxCase(BoxSpec(nel=(3, 2, 3), N=N, bc=bc, upwind=upwind, verts=verts),
                         eps=np.repeat(eps_e, n1 ** 3))
    n = s.npts
    for a, b in zip(s.mesh.coords(), (ref.xm1, ref.ym1, ref.zm1)):
        assert np.max(np.abs(a - b)) <= 1e-14
    H, E = rng.standard_normal((3, n)), rng.standard_normal((3, n))
    s.set_state(H, E)
    rH, rE = s.residual()
    oH, oE = ref.residual(H, E)
    for a_, b_ in ((rH, oH), (rE, oE)):
        assert np.max(np.abs(a_ - b_)) <= 1e-11 * np.max(np.abs(b_))
    dt = ref.dt_from_cfl(0.25)
    assert abs(dt - s.mesh.dt(0.25)) <= 1e-15
    s.step(dt, 20)
    gH, gE = s.get_state()
    out = ref.run(H, E, dt, 20)
    bm = ref.bm1
    for a_, b_ in ((gH, out["H"]), (gE, out["E"])):
        for c in range(3):
            den = np.sqrt(np.sum(bm * b_[c] ** 2))
            assert np.sqrt(np.sum(bm * (a_[c] - b_[c]) ** 2)) <= 1e-11 * den
    s.close()
