"""The reference's own test registry on the HIP path: every 3-D Maxwell entry
of tests/tests.json (kept as tests/golden/maxwell3d_tests.json) is run from
its case files with the entry's .rea parameter overrides applied as the
reference's runner applies them (bin/runtests rewrites the .rea parameters,
then the case's userchk decides pass or fail), and must pass the case's
userchk tolerances.  cylwave (curved elements, N = 11) runs on the
general-hexahedra kernel."""
import json
import os

import numpy as np
import pytest

from nekcem_amd import usr
from nekcem_amd.maxwell import MaxwellSolver

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
CASES = os.path.join(HERE, "golden", "cases")
with open(os.path.join(HERE, "golden", "maxwell3d_tests.json")) as f:
    MATRIX = json.load(f)


def _override(s, params):
    """ReaState's rewrite of the .rea parameters (bin/runtests:42-63)."""
    for k, v in params.items():
        s.param[int(k)] = float(v)


def _analytic(case, params, rescale):
    from tests.test_gpu_rea import _rescale
    s = MaxwellSolver.from_rea(os.path.join(CASES, f"{case}.rea"), 8, usrdat2=_rescale(*rescale))
    _override(s, params)
    s.init_solution(case, 0.0)
    dt = s.mesh.dt(s.param[12])
    tol = {"3dboxper": (5e-10, 5e-9), "3dboxpec": (5e-8, 5e-7)}[case]
    every = int(s.param[13])              # iocomm: userchk runs at istep <= 10 and every iocomm
    for step in range(1, int(s.param[11]) + 1):
        s.step(dt, 1)
        if step <= 10 or step % every == 0:
            l2, li = s.cem_error_device(case)
            assert np.all(l2 <= tol[0]) and np.all(li <= tol[1]), (case, step, l2, li)
    s.close()


def _dielectric(params):
    from tests.test_gpu_pml import _dielectric as build
    from nekcem_amd import rea
    base = rea.read_case(os.path.join(CASES, "3ddielectric.rea"))["param"]
    twomat = float(params.get("70", base[70])) != 0.0
    s, _, sol, _ = build(None, twomat, with_oracle=False)
    _override(s, params)
    dt = s.mesh.dt(s.param[12])
    l2tol, litol = np.array(usr.DIELECTRIC_L2TOL), np.array(usr.DIELECTRIC_LINFTOL[twomat])
    for step in range(1, int(s.param[11]) + 1):
        s.step(dt, 1)
        if step <= 10 or step % int(s.param[13]) == 0:
            l2, li = s.cem_error_fields(*sol(s.time))
            assert np.all(l2 <= l2tol) and np.all(li <= litol), (step, l2, li)
    s.close()


def _graphene(params):
    from tests.test_gpu_graphene import _graphene as build
    s, _, sol, _, _ = build(None, with_oracle=False)
    _override(s, params)
    dt = s.mesh.dt(s.param[12])
    l2tol, litol = np.array(usr.GRAPHENE_L2TOL), np.array(usr.GRAPHENE_LINFTOL)
    for step in range(1, int(s.param[11]) + 1):
        s.step(dt, 1)
        if step <= 10 or step % int(s.param[13]) == 0:
            l2, li = s.cem_error_fields(*sol(s.time))
            assert np.all(l2 <= l2tol) and np.all(li <= litol), (step, l2, li)
    s.close()


def _boxpml(params):
    from tests.test_gpu_pml import _boxpml as build
    s, _ = build(None)
    _override(s, params)
    dt = s.mesh.dt(s.param[12])
    Z = np.zeros((3, s.npts))             # usersol is zero: cem_error measures the fields
    for step in range(1, int(s.param[11]) + 1):
        s.step(dt, 1)
        if step <= 10 or step % int(s.param[13]) == 0:
            l2, li = s.cem_error_fields(Z, Z)
            assert np.all(l2 <= 1.0) and np.all(li <= 1.0), (step, l2, li)
    s.close()


def _cylwave(params):
    """tests/cylwave: the mesh from its .rea with the case's usrdat and
    usrdat2 (curved elements, N = 11), the TM_01 mode as userini, CFL
    param(12), param(11) steps, userchk at istep <= 10 and every iocomm
    against usersol at L2 <= 5e-9, Linf <= 5e-8 per field."""
    from nekcem_amd.maxwell import MaxwellSolver
    from tests.meshes_general import cylwave_mesh
    mesh, R, _, param = cylwave_mesh(11)
    p = list(param)
    for k, v in params.items():
        p[int(k)] = float(v)
    s = MaxwellSolver(mesh)
    x, y, z = mesh.coords()
    s.set_state(*usr.cylwave_solution(0.0, x, y, z, R))
    dt = mesh.dt(p[12])
    l2tol, litol = np.array(usr.CYLWAVE_L2TOL), np.array(usr.CYLWAVE_LINFTOL)
    for step in range(1, int(p[11]) + 1):
        s.step(dt, 1)
        if step <= 10 or step % int(p[13]) == 0:
            l2, li = s.cem_error_fields(*usr.cylwave_solution(s.time, x, y, z, R))
            assert np.all(l2 <= l2tol) and np.all(li <= litol), (step, l2, li)
    s.close()


RUNNERS = {
    "cylwave": _cylwave,
    "3dboxper": lambda p: _analytic("3dboxper", p, (0.0, 2 * np.pi)),
    "3dboxpec": lambda p: _analytic("3dboxpec", p, (-1.0, 1.0)),
    "3ddielectric": _dielectric,
    "3dgraphene": _graphene,
    "3dboxpml": _boxpml,
}


@pytest.mark.parametrize("name", sorted(MATRIX))
def test_reference_registry_entry(name):
    entry = MATRIX[name]
    assert entry["app"] == "maxwell" and "acc" in entry["parallelization"]
    print(f"[case] {name} params={entry['params']}", flush=True)
    RUNNERS[entry["dir"]](entry["params"])
