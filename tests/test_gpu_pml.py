"""PML, incident field and the sinusoidal source on the HIP path: the
reference's 3ddielectric and 3dboxpml cases, built from their .rea + .box,
against the oracle (parity) and against 3ddielectric's exact solution at its
userchk tolerances.

Tolerances: parity 1e-11 relative L2 per field (bm1-weighted) and for the
PML's auxiliary fields; residuals 1e-11 of the field's max; sigma 1e-13
relative (the two sides evaluate the same closed form from coordinates that
differ by round-off)."""
import os

import numpy as np
import pytest

from nekcem_amd.maxwell import MaxwellSolver
from nekcem_amd import usr
from tests.test_gpu_parity import parity

pytestmark = pytest.mark.gpu
CASES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cases")


def _dielectric(oracle, twomat, with_oracle=True):
    """3ddielectric on both sides: the HIP solver from its .rea/.box (N = 8,
    usrdat2 to 5 x 10 x 5, uservp materials, PML from param 77-79, the
    incident faces) and the oracle's box; initial state = usersol(0) with
    pmlbn = mu H, pmldn = eps E (userini)."""
    print(f"[case] 3ddielectric twomat={twomat}", flush=True)
    s = MaxwellSolver.from_rea(os.path.join(CASES, "3ddielectric.rea"), 8,
                               usrdat2=usr.rescale_usrdat2((5.0, 10.0, 5.0)))
    n1 = 9
    lo_y, hi_y = s.mesh.lo[:, 1], s.mesh.hi[:, 1]
    upper = usr.probe_upper(lo_y, hi_y, n1)
    eps, mu = usr.dielectric_materials(upper, n1, twomat)
    s.set_materials(eps, mu)
    thick, order, referr = int(s.param[77]), s.param[78], s.param[79]
    ext = s.setup_pml(thick, order, referr)
    faces, k, om, ha, ea = usr.dielectric_incident(upper, lo_y, hi_y, twomat)
    s.set_incident(faces, k, om, ha, ea)
    x, y, z = s.mesh.coords()
    tol = 1e-9
    pml_elem = (hi_y <= ext["inner"][2] + tol) | (lo_y >= ext["inner"][3] - tol)
    upn, pmn = np.repeat(upper, n1 ** 3), np.repeat(pml_elem, n1 ** 3)

    def sol(t):
        return usr.dielectric_solution(t, y, eps, mu, upn, pmn, ext["inner"], ext["outer"], order, referr, twomat)

    H, E = sol(0.0)
    s.time = 0.0
    s.set_state(H, E)
    s.set_pml_fields(mu * H, eps * E)
    ref = None
    if with_oracle:
        spec = oracle.BoxSpec(nel=(4, 8, 4), N=8, lo=(-2.5, -5.0, -2.5), hi=(2.5, 5.0, 2.5), bc=("P", "PML", "P"))
        ref = oracle.BoxCase(spec, eps=eps, mu=mu)
        ref.set_pml(thick, order, referr)
        n2 = n1 * n1
        on = np.zeros(ref.nxzfl, dtype=np.int8)
        for e, f in faces:
            on[(e * 6 + f - 1) * n2:(e * 6 + f) * n2] = 1
        ref.set_incident(on, k, om, ha, ea)
        assert np.array_equal(np.nonzero(ref.pmltag)[0], np.nonzero(pml_elem)[0])
        assert np.allclose(ref.pmlinner[2:4], ext["inner"][2:4], rtol=0, atol=1e-12)
        assert np.allclose(ref.pmlouter[2:4], ext["outer"][2:4], rtol=0, atol=1e-12)
    return s, ref, sol, (eps, mu)


def test_dielectric_pml_setup_and_residual(oracle):
    """Sigma from nkc_setup_pml matches pml_calc_sigma of the oracle, and
    one residual of a random state (fields, B, D) with the incident field at
    t = 0.3 matches the oracle's: resH, resE, respmlbn, respmldn."""
    s, ref, _, (eps, mu) = _dielectric(oracle, True)
    n = s.npts
    sig = s.get_pml("sigma")
    osig = ref.pmlsigma.reshape(3, n)
    assert np.max(np.abs(sig - osig)) <= 1e-13 * np.max(np.abs(osig))
    assert np.max(osig[1]) > 1.0 and np.max(np.abs(osig[0])) == 0.0
    rng = np.random.default_rng(11)
    H, E = rng.standard_normal((3, n)), rng.standard_normal((3, n))
    B, D = rng.standard_normal((3, n)), rng.standard_normal((3, n))
    s.set_state(H, E)
    s.set_pml_fields(B, D)
    rH, rE = s.residual(0.3)
    rP = s.get_pml("residual")
    oH, oE, oP = ref.residual(H, E, t=0.3, P=np.concatenate([B, D]))
    inpml = np.repeat(ref.pmltag != 0, 9 ** 3)
    for a, b in ((rH, oH), (rE, oE), (rP[:, inpml], oP[:, inpml])):
        assert np.max(np.abs(a - b)) <= 1e-11 * np.max(np.abs(b))


def test_dielectric_steps_parity(oracle):
    """40 RK steps of 3ddielectric (two materials, PML, incident wave):
    fields, B and D within 1e-11 relative L2 of the oracle."""
    s, ref, sol, (eps, mu) = _dielectric(oracle, True)
    H0, E0 = sol(0.0)
    P0 = np.concatenate([mu * H0, eps * E0])
    dt = 0.005
    s.step(dt, 40)
    gH, gE = s.get_state()
    gP = s.get_pml("state")
    out = ref.run(H0, E0, dt, 40, P=P0)
    bm1 = ref.bm1
    assert parity(gH, gE, out["H"], out["E"], bm1) <= 1e-11
    inpml = np.repeat(ref.pmltag != 0, 9 ** 3)
    # B and D against the oracle as the fields are: a component that is zero
    # in the exact solution (By, Dy: round-off only) is measured against the
    # largest component's norm
    assert parity(gP[:3, inpml], gP[3:, inpml], out["P"][:3, inpml], out["P"][3:, inpml], bm1[inpml]) <= 1e-11


@pytest.mark.parametrize("twomat", [False, True], ids=["1mat", "2mat"])
def test_dielectric_userchk(oracle, twomat):
    """The reference's 3ddielectric userchk on the HIP path: 1000 steps of
    dt = 0.005 (param 11, 12), cem_error against usersol at steps 1..10 and
    every 50 (iocomm), L2 and Linf within its per-field tolerances."""
    s, _, sol, _ = _dielectric(oracle, twomat, with_oracle=False)
    dt = s.mesh.dt(s.param[12])
    assert dt == 0.005 and int(s.param[11]) == 1000
    l2tol = np.array(usr.DIELECTRIC_L2TOL)
    litol = np.array(usr.DIELECTRIC_LINFTOL[twomat])
    worst = np.zeros(6)
    for step in range(1, 1001):
        s.step(dt, 1)
        if step <= 10 or step % 50 == 0:
            H, E = sol(s.time)
            l2, li = s.cem_error_fields(H, E)
            worst = np.maximum(worst, l2 / l2tol)
            assert np.all(l2 <= l2tol) and np.all(li <= litol), (step, l2, li)
    print(f"[3ddielectric twomat={twomat}] worst L2 / tolerance per field: {worst}", flush=True)
    assert worst[0] > 1e-3   # the check sees a nonzero error


def _boxpml(oracle):
    print("[case] 3dboxpml", flush=True)
    s = MaxwellSolver.from_rea(os.path.join(CASES, "3dboxpml.rea"), 8,
                               usrdat2=usr.rescale_usrdat2((2.0, 2.0, 2.0)))
    x, y, z = s.mesh.coords()
    amp = usr.dipole_amplitude(x, y, z, s.mesh.bmn)
    s.set_source(amp, direction=(0.0, 0.0, 1.0), omega=-2.0)
    s.setup_pml(int(s.param[77]), s.param[78], s.param[79])
    s.time = 0.0
    n = s.npts
    s.set_state(np.zeros((3, n)), np.zeros((3, n)))
    ref = None
    if oracle is not None:
        spec = oracle.BoxSpec(nel=(6, 6, 6), N=8, lo=(-1.0,) * 3, hi=(1.0,) * 3, bc=("PML",) * 3)
        ref = oracle.BoxCase(spec)
        ref.set_pml(int(s.param[77]), s.param[78], s.param[79])
        ref.set_source(usr.dipole_amplitude(ref.xm1, ref.ym1, ref.zm1, ref.bm1), direction=(0.0, 0.0, 1.0),
                       omega=-2.0)
    return s, ref


def test_boxpml_parity(oracle):
    """3dboxpml (PML on all six walls: edge and corner elements carry two or
    three sigmas; the dipole's sin(-omega t) source): sigma, then 30 steps
    from rest within 1e-11 of the oracle, fields and B, D."""
    s, ref = _boxpml(oracle)
    n = s.npts
    sig = s.get_pml("sigma")
    osig = ref.pmlsigma.reshape(3, n)
    assert np.max(np.abs(sig - osig)) <= 1e-13 * np.max(np.abs(osig))
    assert np.sum(np.count_nonzero(osig.reshape(3, -1, 729).max(axis=2), axis=0) == 3) == 8
    dt = s.mesh.dt(s.param[12])
    s.step(dt, 30)
    gH, gE = s.get_state()
    gP = s.get_pml("state")
    out = ref.run(np.zeros((3, n)), np.zeros((3, n)), dt, 30)
    assert np.max(np.abs(out["E"][2])) > 1e-4
    # Hz is zero by the z-dipole's symmetry: round-off only, 1.3e-6 of Ez's
    # norm here (a 1e-15 relative change of the source moves it by 1.6e-11
    # of itself on the oracle alone), so it is measured against Ez
    assert parity(gH, gE, out["H"], out["E"], ref.bm1, floor=1e-5) <= 1e-11
    inpml = np.repeat(ref.pmltag != 0, 729)
    d = np.max(np.abs(gP[:, inpml] - out["P"][:, inpml]))
    assert d <= 1e-11 * np.max(np.abs(out["P"][:, inpml]))


def test_boxpml_userchk():
    """The reference's 3dboxpml userchk: 2000 steps of dt = 0.1 dxmin from
    rest with the dipole on, every field's L2 and Linf <= 1 at every 10th
    step (usersol is zero there, so cem_error measures the fields), and the
    radiated field does reach the PML."""
    s, _ = _boxpml(None)
    dt = s.mesh.dt(s.param[12])
    n = s.npts
    Z = np.zeros((3, n))
    peak = 0.0
    for step in range(10, int(s.param[11]) + 1, 10):
        s.step(dt, 10)
        l2, li = s.cem_error_fields(Z, Z)
        peak = max(peak, float(np.max(li)))
        assert np.all(l2 <= 1.0) and np.all(li <= 1.0), (step, l2, li)
    P = s.get_pml("state")
    assert np.max(np.abs(P)) > 1e-6 and peak > 1e-3


def test_pml_layouts_bit_identical(oracle):
    """Every stage-kernel layout gives the same bits with the PML, the
    incident field and two materials (the source-bearing kernel instance)."""
    outs = []
    for layout in range(4):
        s, _, _, _ = _dielectric(oracle, True, with_oracle=False)
        s.set_layout(layout)
        s.step(0.005, 5)
        outs.append((s.get_state(), s.get_pml("state")))
        s.close()
    for (H, E), P in outs[1:]:
        assert np.array_equal(H, outs[0][0][0]) and np.array_equal(E, outs[0][0][1])
        assert np.array_equal(P, outs[0][1])
    # the six-role layouts (built at this order) carry no PML: refused
    from nekcem_amd._lib import NkcError, NKC_ERR_UNSUPPORTED
    s, _, _, _ = _dielectric(oracle, True, with_oracle=False)
    for layout in (6, 7):
        with pytest.raises(NkcError, match="carries no") as ei:
            s.set_layout(layout)
        assert ei.value.code == NKC_ERR_UNSUPPORTED
    s.close()


def test_adjoint_refuses_pml(oracle):
    s, _, _, _ = _dielectric(oracle, False, with_oracle=False)
    with pytest.raises(Exception, match="PML"):
        s.adjoint_forward(0.005, 4, 2)


def test_pml_and_incident_argument_checks(oracle):
    """The set-up entry points refuse what the kernels cannot honour, with a
    message and no state change: a PML face that is not a wall, a PML
    thicker than the mesh, colliding layers, parameters outside
    pml_errchk's ranges, incident faces out of range; and a PML handle
    keeps stepping after a refused call."""
    from nekcem_amd._lib import NkcError
    print("[case] PML / incident argument checks", flush=True)
    s = MaxwellSolver.from_box((2, 4, 2), 4, (0, -1, 0), (1, 1, 1), ("P", "PML", "P"))
    pf = s._pml_faces()
    assert pf.sum() == 2 * 2 * 2     # y- faces of the bottom layer, y+ of the top
    bad = pf.copy()
    bad[0, 1] = 1                    # x+ face of element 0: interior
    with pytest.raises(NkcError, match="not a wall"):
        s.setup_pml(1, 3.0, 1e-8, pml_faces=bad)
    with pytest.raises(NkcError, match="thicker|collide"):
        s.setup_pml(3, 3.0, 1e-8)
    with pytest.raises(NkcError, match="thickness"):
        s.setup_pml(11, 3.0, 1e-8)
    with pytest.raises(NkcError, match="order"):
        s.setup_pml(1, 0.5, 1e-8)
    with pytest.raises(NkcError, match="reflection"):
        s.setup_pml(1, 3.0, 0.0)
    ext = s.setup_pml(2, 3.0, 1e-8)
    assert np.allclose(ext["inner"][2:4], [0.0, 0.0]) and np.allclose(ext["outer"][2:4], [-1.0, 1.0])
    with pytest.raises(NkcError, match="out of range"):
        s.set_incident([(16, 1)], (0, 1, 0), 1.0, (1, 0, 0), (0, 0, 1))
    with pytest.raises(NkcError, match="out of range"):
        s.set_incident([(0, 7)], (0, 1, 0), 1.0, (1, 0, 0), (0, 0, 1))
    n = s.npts
    rng = np.random.default_rng(3)
    s.set_state(rng.standard_normal((3, n)), rng.standard_normal((3, n)))
    s.step(1e-3, 3)
    H, E = s.get_state()
    assert np.all(np.isfinite(H)) and np.all(np.isfinite(E))
    s.close()
