"""The graphene sheet current on the HIP path: the reference's 3dgraphene
case, built from its .rea + .box, against the oracle (parity) and against
its exact solution at its userchk tolerances; sheets on an interior z plane
(both sides listed), an interior x plane and a periodic wrap face (one side
listed) and a PEC wall at N = 7 and 4, with random per-point sheet
parameters and two materials, against the oracle; every layout
bit-identical; graph replay equal to direct launches; set/get round trip and
argument checks.

Tolerances: parity 1e-11 relative L2 per field (bm1-weighted) and for the
sheet state (relative L2 over the listed points); residuals 1e-11 of the
field's max."""
import os

import numpy as np
import pytest

from nekcem_amd.maxwell import MaxwellSolver
from nekcem_amd import usr
from tests.test_gpu_parity import parity

pytestmark = pytest.mark.gpu
CASES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cases")


def _sheet_index(faces, n2):
    """0-based face-point numbers of the listed (element, Ed face) pairs."""
    return np.concatenate([np.arange((e * 6 + f - 1) * n2, (e * 6 + f) * n2) for e, f in faces])


def _graphene(oracle, with_oracle=True):
    """3dgraphene on both sides: the HIP solver from its .rea/.box (N = 8,
    usrdat2 to 5 x 10 x 5, PML from param 77-79, the incident faces, the
    sheet on both sides of y = 0) and the oracle's box; initial state =
    usersol(0), pmlbn = mu H, pmldn = eps E, sheet state from userini."""
    print("[case] 3dgraphene", flush=True)
    s = MaxwellSolver.from_rea(os.path.join(CASES, "3dgraphene.rea"), 8,
                               usrdat2=usr.rescale_usrdat2((5.0, 10.0, 5.0)))
    n1, n2 = 9, 81
    assert s.mesh.nelt == 4 * 12 * 4
    lo_y, hi_y = s.mesh.lo[:, 1], s.mesh.hi[:, 1]
    upper = usr.probe_upper(lo_y, hi_y, n1)
    thick, order, referr = int(s.param[77]), s.param[78], s.param[79]
    ext = s.setup_pml(thick, order, referr)
    faces, k, om, ha, ea = usr.graphene_incident(upper, lo_y, hi_y)
    assert len(faces) == 16
    s.set_incident(faces, k, om, ha, ea)
    sheet = usr.sheet_faces(lo_y, hi_y)
    assert len(sheet) == 32
    gidx = _sheet_index(sheet, n2)
    G0 = np.repeat(usr.graphene_initial_state().reshape(15, 1), len(gidx), axis=1)
    s.set_graphene(gidx, usr.GRAPHENE_SHEET, G0)
    x, y, z = s.mesh.coords()
    tol = 1e-9
    pml_elem = (hi_y <= ext["inner"][2] + tol) | (lo_y >= ext["inner"][3] - tol)
    upn, pmn = np.repeat(upper, n1 ** 3), np.repeat(pml_elem, n1 ** 3)
    one = np.ones(s.npts)

    def sol(t):
        return usr.graphene_solution(t, y, one, one, upn, pmn, ext["inner"], ext["outer"], order, referr)

    H, E = sol(0.0)
    s.time = 0.0
    s.set_state(H, E)
    s.set_pml_fields(H, E)
    ref = None
    if with_oracle:
        spec = oracle.BoxSpec(nel=(4, 12, 4), N=8, lo=(-2.5, -5.0, -2.5), hi=(2.5, 5.0, 2.5), bc=("P", "PML", "P"))
        ref = oracle.BoxCase(spec)
        ref.set_pml(thick, order, referr)
        on = np.zeros(ref.nxzfl, dtype=np.int8)
        for e, f in faces:
            on[(e * 6 + f - 1) * n2:(e * 6 + f) * n2] = 1
        ref.set_incident(on, k, om, ha, ea)
        ref.set_graphene(gidx, usr.GRAPHENE_SHEET)
        assert np.array_equal(np.nonzero(ref.pmltag)[0], np.nonzero(pml_elem)[0])
    return s, ref, sol, gidx, G0


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_graphene_steps_parity(oracle):
    """40 RK steps of 3dgraphene (PML, incident wave, sheet): fields, B, D
    and the sheet state and register within 1e-11 relative L2 of the
    oracle."""
    s, ref, sol, gidx, G0 = _graphene(oracle)
    H0, E0 = sol(0.0)
    dt = 0.005
    s.step(dt, 40)
    gH, gE = s.get_state()
    gP = s.get_pml("state")
    out = ref.run(H0, E0, dt, 40, P=np.concatenate([H0, E0]), G=G0)
    assert parity(gH, gE, out["H"], out["E"], ref.bm1) <= 1e-11
    inpml = np.repeat(ref.pmltag != 0, 9 ** 3)
    assert parity(gP[:3, inpml], gP[3:, inpml], out["P"][:3, inpml], out["P"][3:, inpml], ref.bm1[inpml]) <= 1e-11
    G, kG = s.get_graphene("state"), s.get_graphene("register")
    print(f"[3dgraphene] sheet state rel L2 {_rel(G, out['G']):.2e}, register {_rel(kG, out['kG']):.2e}", flush=True)
    assert _rel(G, out["G"]) <= 1e-11 and _rel(kG, out["kG"]) <= 1e-11
    assert _rel(G, G0) > 1e-3     # the state moved


def test_graphene_residual(oracle):
    """One residual of a random state (fields, B, D, sheet state) at
    t = 0.3 matches the oracle's: the sheet current of the present state
    enters the flux, and the state itself is not advanced."""
    s, ref, _, gidx, _ = _graphene(oracle)
    n = s.npts
    rng = np.random.default_rng(5)
    H, E = rng.standard_normal((3, n)), rng.standard_normal((3, n))
    B, D = rng.standard_normal((3, n)), rng.standard_normal((3, n))
    G = rng.standard_normal((15, len(gidx)))
    s.set_state(H, E)
    s.set_pml_fields(B, D)
    s.set_graphene(gidx, usr.GRAPHENE_SHEET, G)
    rH, rE = s.residual(0.3)
    oH, oE, _, _ = ref.residual(H, E, t=0.3, P=np.concatenate([B, D]), G=G)
    for a, b in ((rH, oH), (rE, oE)):
        assert np.max(np.abs(a - b)) <= 1e-11 * np.max(np.abs(b))
    assert np.array_equal(s.get_graphene("state"), G)


def test_graphene_userchk(oracle):
    """The reference's 3dgraphene userchk on the HIP path: 1000 steps of
    dt = 0.005 (param 11, 12), cem_error against usersol at steps 1..10 and
    every 50 (iocomm), L2 and Linf within its per-field tolerances."""
    s, _, sol, _, _ = _graphene(oracle, with_oracle=False)
    dt = s.mesh.dt(s.param[12])
    assert dt == 0.005 and int(s.param[11]) == 1000
    l2tol, litol = np.array(usr.GRAPHENE_L2TOL), np.array(usr.GRAPHENE_LINFTOL)
    worst = np.zeros(6)
    for step in range(1, 1001):
        s.step(dt, 1)
        if step <= 10 or step % 50 == 0:
            H, E = sol(s.time)
            l2, li = s.cem_error_fields(H, E)
            worst = np.maximum(worst, l2 / l2tol)
            assert np.all(l2 <= l2tol) and np.all(li <= litol), (step, l2, li)
    print(f"[3dgraphene] worst L2 / tolerance per field: {worst}", flush=True)
    assert worst[0] > 1e-2


def _planes(oracle, N=7, nel=(4, 3, 4), seed=3):
    """A PEC-walled (y) box, periodic in x and z, two materials (eps 1 / 2.5
    by element parity), random smooth fields; sheets with random per-point
    parameters and state on: the interior z plane between layers 1 and 2
    (both sides), the interior x plane between columns 1 and 2 (the x+ side
    only), the y- PEC wall of the bottom layer and, in the first z layer,
    the periodic x wrap (the x- side only)."""
    n1 = N + 1
    n2, n3 = n1 * n1, n1 ** 3
    print(f"[case] sheet planes N={N} nel={nel}", flush=True)
    spec = oracle.BoxSpec(nel=nel, N=N, lo=(0.0, 0.0, 0.0), hi=(1.0, 0.75, 1.0), bc=("P", "PEC", "P"))
    ref = oracle.BoxCase(spec)
    ex, ey, ez = ref.exyz
    eps = np.repeat(np.where((ex + ey + ez) % 2 == 0, 1.0, 2.5), n3)
    mu = np.repeat(np.where(ez >= 2, 1.5, 1.0), n3)
    ref = oracle.BoxCase(spec, eps=eps, mu=mu)
    faces = []
    for e in range(ref.nelt):
        if ez[e] == 1:
            faces.append((e, 6))
        if ez[e] == 2:
            faces.append((e, 5))
        if ex[e] == 1:
            faces.append((e, 2))
        if ey[e] == 0:
            faces.append((e, 1))
        if ex[e] == 0 and ez[e] == 0:
            faces.append((e, 4))     # x- face on the periodic wrap, one side listed
    gidx = _sheet_index(faces, n2)
    rng = np.random.default_rng(seed)
    n = len(gidx)
    lo = np.array([0.0, 0.5, -0.2, 10.0, 1.0, -1.0, -5.0, -0.2, 10.0, 1.0, -1.0, -5.0])
    hi = np.array([1.0, 2.0, 0.0, 60.0, 5.0, 1.0, 5.0, 0.0, 60.0, 5.0, 1.0, 5.0])
    par = lo + (hi - lo) * rng.random((n, 12))
    G0 = 0.1 * rng.standard_normal((15, n))
    ref.set_graphene(gidx, par)
    s = MaxwellSolver.from_box(nel, N, spec.lo, spec.hi, spec.bc)
    s.set_materials(eps, mu)
    s.set_graphene(gidx, par, G0)
    x, y, z = s.mesh.coords()
    H = np.stack([usr.smooth_field(x, y, z, seed + c) for c in range(3)])
    E = np.stack([usr.smooth_field(x, y, z, seed + 3 + c) for c in range(3)])
    s.set_state(H, E)
    return s, ref, H, E, G0, (gidx, par)


@pytest.mark.parametrize("N", [7, 4], ids=["N7", "N4"])
def test_graphene_planes_parity(oracle, N):
    """Sheets on z and x planes and a PEC wall (N = 7 takes the phase-2
    z-face path, N = 4 the phase-1 one and the two-element layouts), 30
    steps: fields, sheet state and register within 1e-11 of the oracle."""
    s, ref, H, E, G0, _ = _planes(oracle, N=N)
    dt = 0.002
    s.step(dt, 30)
    gH, gE = s.get_state()
    out = ref.run(H, E, dt, 30, G=G0)
    assert parity(gH, gE, out["H"], out["E"], ref.bm1) <= 1e-11
    G, kG = s.get_graphene("state"), s.get_graphene("register")
    assert _rel(G, out["G"]) <= 1e-11 and _rel(kG, out["kG"]) <= 1e-11
    # the sheet changed the fields: the same run without it differs
    ref.gindex = None
    bare = ref.run(H, E, dt, 30)
    assert parity(gH, gE, bare["H"], bare["E"], ref.bm1) > 1e-6


@pytest.mark.parametrize("N", [7, 4], ids=["N7", "N4"])
def test_graphene_layouts_bit_identical(oracle, N):
    """Every built stage-kernel layout gives the same bits with sheets."""
    outs = []
    from nekcem_amd.maxwell import NLAYOUT
    for layout in range(NLAYOUT):
        s, _, _, _, _, _ = _planes(oracle, N=N)
        try:
            s.set_layout(layout)
        except Exception as err:     # a layout not built for this order, or without the sheet
            assert "not built" in str(err) or "carries no" in str(err)
            s.close()
            continue
        s.step(0.002, 4)
        outs.append((layout, s.get_state(), s.get_graphene("state")))
        s.close()
    assert len(outs) >= 4
    for layout, (H, E), G in outs[1:]:
        assert np.array_equal(H, outs[0][1][0]) and np.array_equal(E, outs[0][1][1]), layout
        assert np.array_equal(G, outs[0][2]), layout


def test_graphene_materials_follow(oracle):
    """nkc_set_materials after nkc_set_graphene refreshes the sheet's Yfac
    and yconduc: the run matches the oracle built with the new materials."""
    s, ref, H, E, G0, (gidx, par) = _planes(oracle, N=4)
    n = s.npts
    eps2 = np.full(n, 3.0)
    mu2 = np.full(n, 1.2)
    s.set_materials(eps2, mu2)
    s.set_state(H, E)
    ref2 = type(ref)(ref.spec, eps=eps2, mu=mu2)
    ref2.set_graphene(gidx, par)
    s.step(0.002, 10)
    gH, gE = s.get_state()
    out = ref2.run(H, E, 0.002, 10, G=G0)
    assert parity(gH, gE, out["H"], out["E"], ref2.bm1) <= 1e-11


def test_graphene_set_get_and_checks(oracle):
    """Round trip of the state, removal, and the argument checks: an index
    out of range, one listed twice, a short leading dimension, a state that
    is not finite; the adjoint refuses a sheet."""
    from nekcem_amd._lib import NkcError
    s, ref, H, E, G0, (gidx, par) = _planes(oracle, N=4)
    assert np.array_equal(s.get_graphene("state"), G0)
    assert np.array_equal(s.get_graphene("register"), np.zeros_like(G0))
    with pytest.raises(Exception, match="graphene"):
        s.adjoint_forward(0.002, 4, 2)
    nxzfl = s.mesh.nelt * 6 * 25
    import ctypes as C
    from nekcem_amd._lib import check
    dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int32)
    p12, f18 = np.zeros(12 * nxzfl), np.zeros(18 * nxzfl)

    def raw(index, ld):
        gi = np.asarray(index, dtype=np.int32)
        check(s._L.nkc_set_graphene(s._h, len(gi), gi.ctypes.data_as(ip), p12.ctypes.data_as(dp),
                                    f18.ctypes.data_as(dp), ld))
    with pytest.raises(NkcError, match="out of range"):
        raw([nxzfl + 1], nxzfl)
    with pytest.raises(NkcError, match="out of range"):
        raw([0], nxzfl)
    with pytest.raises(NkcError, match="listed twice"):
        raw([3, 3], nxzfl)
    with pytest.raises(NkcError, match="ld"):
        raw([1], nxzfl - 1)
    f18[3 * nxzfl] = np.nan          # fjn(1, 1, 2): point 1's Drude slot
    with pytest.raises(NkcError, match="not finite"):
        raw([1], nxzfl)
    f18[3 * nxzfl] = 0.0
    # a refused call leaves no sheet: the handle steps as the bare oracle
    s.set_state(H, E)
    s.step(0.002, 5)
    gH, gE = s.get_state()
    ref.gindex = None
    out = ref.run(H, E, 0.002, 5)
    assert parity(gH, gE, out["H"], out["E"], ref.bm1) <= 1e-11
    # and so does one whose sheet was removed
    s.set_graphene(gidx, par, G0)
    s.set_graphene([], None)
    s.set_state(H, E)
    s.step(0.002, 5)
    gH2, gE2 = s.get_state()
    assert np.array_equal(gH, gH2) and np.array_equal(gE, gE2)


def test_graphene_graph_equals_direct(oracle):
    """The captured step (two graphs, one per starting copy) and direct
    launches give the same bits for fields and sheet state over an odd
    number of steps split across calls."""
    from nekcem_amd._lib import check
    outs = []
    for graph in (True, False):
        s, _, _, _, _, _ = _planes(oracle, N=4)
        if not graph:
            check(s._L.nkc_set_graph(s._h, 0))
        s.step(0.002, 3)
        s.step(0.002, 4)
        outs.append((s.get_state(), s.get_graphene("state"), s.get_graphene("register")))
        s.close()
    (H0, E0), G0, K0 = outs[0]
    (H1, E1), G1, K1 = outs[1]
    assert np.array_equal(H0, H1) and np.array_equal(E0, E1)
    assert np.array_equal(G0, G1) and np.array_equal(K0, K1)
