"""nkc_cem_error_noalloc, the cem_error entry point that allocates nothing
(its partial norms in the library's fixed device array), against
nkc_cem_error (the same kernels with handle-allocated partials): equal bits
on two calls in a row, the second on a smaller mesh (fewer partials than
the first left in the array), and again on the first mesh after it."""
import ctypes as C

import numpy as np
import pytest

from nekcem_amd import _lib
from nekcem_amd.maxwell import MaxwellSolver

pytestmark = pytest.mark.gpu


def _noalloc(s, sol, t):
    out = np.empty(12)
    _lib.check(s._L.nkc_cem_error_noalloc(s._h, s.SOLUTIONS[sol], float(t), out.ctypes.data_as(C.POINTER(C.c_double))))
    return out


def test_cem_error_noalloc_matches_cem_error():
    print("[case] nkc_cem_error_noalloc: 6^3 N=7 periodic, then 2x1x2 N=4 PEC", flush=True)
    big = MaxwellSolver.from_box((6, 6, 6), 7)                    # 110592 points: 432 partial blocks
    small = MaxwellSolver.from_box((2, 1, 2), 4, (0, 0, 0), (1, 1, 1), ("PEC",) * 3)   # 500 points: 2 blocks
    cases = ((big, "3dboxper"), (small, "3dboxpec"))
    # a box handle has no node coordinates on the device before its first
    # diagnostic: refused, nothing allocated
    with pytest.raises(_lib.NkcError, match="coordinates"):
        _noalloc(big, "3dboxper", 0.0)
    ref = []
    for s, sol in cases:
        s.init_solution(sol, 0.0)
        s.step(s.mesh.dt(0.25), 3)
        l2, linf = s.cem_error_device(sol)
        assert np.all(np.isfinite(l2)) and np.max(l2) > 0
        ref.append(np.concatenate([l2, linf]))
    got = [_noalloc(s, sol, s.time) for s, sol in cases]     # big, then small
    got.append(_noalloc(big, "3dboxper", big.time))           # big again after small
    for g, r in zip(got, ref + ref[:1]):
        assert g.tobytes() == r.tobytes(), (g, r)
    for s, _ in cases:
        s.close()
