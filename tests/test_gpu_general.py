"""General hexahedra on the HIP path (the per-node-geometry kernel,
geo_kernel.hpp): parity with the oracle's general case on box elements in
rotated local frames (faces paired in other orientations), on a smoothly
curved box and on the reference's cylwave mesh (80 circular-arc edges,
N = 11); the rotated box against the box kernels on the same physical
nodes; the on-device cem_error with per-node mass; and the entry points
the general kernel does not carry refusing.

Tolerances as in the box parity tests: one residual or curl within 1e-11
of the field's max, 20 RK steps within 1e-11 relative L2 (bm1-weighted).
"""
import numpy as np
import pytest

from nekcem_amd import _lib
from nekcem_amd.maxwell import MaxwellSolver, cem_error
from nekcem_amd.mesh import BoxMesh
from tests.meshes_general import curved_box, cylwave_mesh, rotated_box

pytestmark = pytest.mark.gpu


def _rel_l2(g, o, w):
    big = max(float(np.sqrt(np.sum(w * o[c] ** 2))) for c in range(len(o)))
    return max(float(np.sqrt(np.sum(w * (g[c] - o[c]) ** 2))) / max(float(np.sqrt(np.sum(w * o[c] ** 2))),
                                                                         1e-6 * big) for c in range(len(o)))


def _case(oracle, name):
    """(HexMesh, oracle GeneralCase, eps) of a named general mesh."""
    rng = np.random.default_rng(len(name))
    if name.startswith("rotated"):
        N = int(name[-1])
        mesh, _, periods = rotated_box((2, 3, 2), N, (0, 0, 0), (1.0, 1.2, 0.8), ("P", "P", "PEC"),
                                       which=[0, 2, 5, 7, 10], seed=N)
    elif name.startswith("curved"):
        N = int(name[-1])
        mesh, periods = curved_box((2, 2, 2), N), []
    else:
        N = int(name.split("N")[-1])
        mesh, _, zp, _ = cylwave_mesh(N)
        periods = [(0, 0, zp)]
    eps = None
    if "eps" in name:
        eps = rng.uniform(1.0, 3.0, mesh.npts)
    x, y, z = mesh.coords()
    ref = oracle.GeneralCase(mesh.N, x, y, z, mesh.pec_face, periods=periods, eps=eps)
    assert np.array_equal(ref.partner + 1, mesh.face_partner)
    return mesh, ref, eps


CASES = ["rotated-N5", "rotated-N8", "curved-N4", "curved-eps-N7", "cylwave-N5", "cylwave-N11"]


@pytest.mark.parametrize("name", CASES)
def test_general_residual_and_curl_parity(oracle, name):
    print(f"[case] general residual/curl {name}", flush=True)
    mesh, ref, eps = _case(oracle, name)
    rng = np.random.default_rng(5)
    n = mesh.npts
    H, E = rng.standard_normal((3, n)), rng.standard_normal((3, n))
    g = MaxwellSolver(mesh, eps=eps)
    assert g.info()["mat_mode"] == 1 and g.info()["threads_per_wg"] == 256
    g.set_state(H, E)
    for got, want in ((g.residual(), ref.residual(H, E)), (g.curl(), ref.curl(H, E))):
        for a, b in zip(got, want):
            for c in range(3):
                assert np.max(np.abs(a[c] - b[c])) <= 1e-11 * np.max(np.abs(b[c])), (c, np.max(np.abs(a[c] - b[c])))
    g.close()


@pytest.mark.parametrize("name", ["rotated-N5", "curved-eps-N7", "cylwave-N11"])
def test_general_steps_parity(oracle, name):
    """20 RK steps (graph replay) against the oracle, 1e-11 relative L2."""
    print(f"[case] general 20 steps {name}", flush=True)
    mesh, ref, eps = _case(oracle, name)
    rng = np.random.default_rng(6)
    n = mesh.npts
    H, E = 0.1 * rng.standard_normal((3, n)), 0.1 * rng.standard_normal((3, n))
    dt = mesh.dt(0.25)
    assert abs(dt - 0.25 * ref.dxmin) < 1e-15
    g = MaxwellSolver(mesh, eps=eps)
    g.set_state(H, E)
    g.step(dt, 20)
    gH, gE = g.get_state()
    out = ref.run(H, E, dt, 20)
    assert _rel_l2(np.concatenate([gH, gE]), np.concatenate([out["H"], out["E"]]), mesh.bmn) <= 1e-11
    g.close()


def test_rotated_frames_match_the_box_kernels():
    """A box whose elements sit in rotated local frames (general kernel,
    per-node tables, faces paired in other orientations) and the same box
    on the tuned box kernels agree node for node after 10 steps (1e-12
    relative), and so do the on-device cem_error norms."""
    print("[case] rotated frames vs box kernels", flush=True)
    nel, N, lo, hi, bc = (2, 3, 2), 6, (0, 0, 0), (1.0, 1.2, 0.8), ("P", "PEC", "P")
    mesh, idx, _ = rotated_box(nel, N, lo, hi, bc, which=[1, 2, 4, 8, 9, 11], seed=4)
    box = BoxMesh(nel, N, lo, hi, bc)
    rng = np.random.default_rng(7)
    H, E = rng.standard_normal((3, box.npts)), rng.standard_normal((3, box.npts))
    dt = box.dt(0.25)
    gb = MaxwellSolver(box)
    gb.set_state(H, E)
    gb.step(dt, 10)
    bH, bE = gb.get_state()
    gr = MaxwellSolver(mesh)
    gr.set_state(H[:, idx], E[:, idx])
    gr.step(dt, 10)
    rH, rE = gr.get_state()
    for a, b in ((rH, bH[:, idx]), (rE, bE[:, idx])):
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))
    zH, zE = np.zeros_like(H), np.zeros_like(E)
    lb = np.concatenate(gb.cem_error_fields(zH, zE))
    lr = np.concatenate(gr.cem_error_fields(zH[:, idx], zE[:, idx]))
    assert np.allclose(lr, lb, rtol=1e-12, atol=0)
    gb.close()
    gr.close()


def test_general_cem_error_per_node_mass(oracle):
    """cem_error on the device (the nodes' own bm1 on a curved mesh) equals
    the host's mass-weighted norms."""
    print("[case] general cem_error", flush=True)
    mesh, _, _ = _case(oracle, "curved-N4")
    rng = np.random.default_rng(8)
    n = mesh.npts
    H, E, Hx, Ex = (rng.standard_normal((3, n)) for _ in range(4))
    g = MaxwellSolver(mesh)
    g.set_state(H, E)
    l2, li = g.cem_error_fields(Hx, Ex)
    for c, (u, x) in enumerate(zip(list(H) + list(E), list(Hx) + list(Ex))):
        hl2, hli = cem_error(u, x, mesh.bmn)
        assert abs(l2[c] - hl2) <= 1e-13 * hl2 and li[c] == hli
    g.close()


def test_general_refuses_box_only_physics(oracle):
    """Sources, dispersive regions, PML, incident field, graphene, a
    material swap, the adjoint and other layouts are built for box meshes
    only: the general handle refuses them with NKC_ERR_UNSUPPORTED."""
    mesh, _, _ = _case(oracle, "curved-N4")
    g = MaxwellSolver(mesh)
    n = mesh.npts
    calls = [
        lambda: g.set_source(np.ones(n)),
        lambda: g.set_drude([0, 1], 0.1, 1.0),
        lambda: g.set_lorentz([0, 1], 0.1, 1.0, 1.0),
        lambda: g.set_materials(np.ones(n)),
        lambda: g.adjoint_forward(1e-3, 2, 1),
        lambda: g.adjoint_apply(np.zeros((3, n)), np.zeros((3, n))),
        lambda: g.set_layout(2),
    ]
    for call in calls:
        with pytest.raises(_lib.NkcError) as ei:
            call()
        assert ei.value.code == _lib.NKC_ERR_UNSUPPORTED
    g.set_layout(0)
    g.close()


def test_cylwave_from_its_rea():
    """MaxwellSolver.from_rea on the reference's cylwave.rea (its usrdat on
    the corners, usrdat2 on the nodes) takes the general path and steps
    exactly like the mesh built directly."""
    import os
    from nekcem_amd import usr
    print("[case] cylwave from_rea", flush=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cases", "cylwave.rea")
    s = MaxwellSolver.from_rea(path, 11, usrdat2=usr.cylwave_usrdat, usrdat2_nodes=usr.cylwave_usrdat2)
    mesh, R, _, param = cylwave_mesh(11)
    assert s.npts == mesh.npts and s.param[11] == param[11]
    x, y, z = mesh.coords()
    H, E = usr.cylwave_solution(0.0, x, y, z, R)
    g = MaxwellSolver(mesh)
    dt = mesh.dt(param[12])
    for h in (s, g):
        h.set_state(H, E)
        h.step(dt, 5)
    a, b = s.get_state(), g.get_state()
    assert all(np.array_equal(u, v) for u, v in zip(a, b))
    s.close()
    g.close()
