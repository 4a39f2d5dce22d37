"""A mesh whose field blocks sit beyond 2^31 doubles: 96^3 elements at N = 7
(452,984,832 points, 65 GB of state; the Ez block starts 2.26e9 doubles
into each state buffer), built as a box handle, the 3dboxper mode set on the
device, three RK steps, cem_error of every field on the device against the
analytic mode.  The oracle cannot run this size; the analytic solution is
the size-independent check (at 4^3 x 8 the same mode meets the reference's
tolerance 5e-10, tests/3dboxper/3dboxper.usr:169-216)."""
import numpy as np
import pytest

from nekcem_amd.maxwell import MaxwellSolver

pytestmark = pytest.mark.gpu


def test_field_offsets_beyond_2_31():
    nel, N = 96, 7
    print(f"[case] 3dboxper mode on a {nel}^3 N={N} box", flush=True)
    s = MaxwellSolver.from_box((nel,) * 3, N, (0, 0, 0), (2 * np.pi,) * 3, ("P",) * 3)
    assert 5 * s.info()["fs"] > 2 ** 31
    s.set_layout(2)                   # fixed: no tuner scratch copy of the state
    s.init_solution("3dboxper", 0.0)
    l2, li = s.cem_error_device("3dboxper")
    assert np.max(l2) < 1e-14
    s.step(0.25 * 2 * np.pi / nel / 40, 3)
    l2, li = s.cem_error_device("3dboxper")
    print(f"[scale] L2 {l2} Linf {li}", flush=True)
    assert np.all(np.asarray(l2) < 5e-10) and np.all(np.asarray(li) < 5e-9)
    # a dropped or misplaced field block would leave an O(1) error
    assert np.max(l2) < 1e-13
    s.close()
