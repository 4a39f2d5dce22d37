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
