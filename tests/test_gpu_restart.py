"""Restart files on the HIP path: nkc_restart_write (cem_restart_out,
src/io.F:411-490 with src/io_co.c) and nkc_restart_read (restart_swap,
src/io.F:637-788).  The file's bytes equal the oracle's restatement of the
layout for the same state; a double-precision file gives the state back
bit for bit (the reference's tests/restart userchk holds it to 1e-15); a
float file to float32 rounding; stepping on from a read-back state equals
stepping on without the restart; files of another mesh are refused.
Parity with the reference's own writer is unpinned (it needs MPI-IO)."""
import os

import numpy as np
import pytest

from nekcem_amd import _lib
from nekcem_amd.maxwell import MaxwellSolver, cem_error

pytestmark = pytest.mark.gpu


def _solver(N=5, nel=(2, 3, 2)):
    return MaxwellSolver.from_box(nel, N, (0, 0, 0), (1.0, 1.2, 0.8), ("PEC", "P", "P"))


def test_restart_bytes_round_trip_and_continuation(oracle, tmp_path):
    print("[case] restart write/read, box 2x3x2 N=5", flush=True)
    rng = np.random.default_rng(1)
    s = _solver()
    n = s.npts
    H, E = rng.standard_normal((3, n)), rng.standard_normal((3, n))
    s.set_state(H, E)
    s.time, s.istep, s.dt = 0.731234567, 120, 2.5e-3
    path = tmp_path / "restart-mpi-binary-N1-t00003.vtk"
    s.restart_out(path, dumpno=3)
    xyz = s.mesh.coords()
    want = oracle.restart_bytes(120, 3, s.mesh.n1, 0.731234567, 2.5e-3, xyz, E, H)
    assert path.read_bytes() == want
    # read back into a solver holding another state
    s.set_state(np.zeros((3, n)), np.ones((3, n)))
    s.time = 0.0
    hdr = s.restart_swap(path)
    assert hdr == dict(istep=120, dumpno=3, time=0.731235, dt=0.0025)   # %f: six decimals
    assert s.time == 0.731235
    gH, gE = s.get_state()
    assert np.array_equal(gH, H) and np.array_equal(gE, E)
    # stepping on from the read-back state = stepping on without the restart
    t = s.mesh.dt(0.25)
    s.step(t, 4)
    a = s.get_state()
    r = _solver()
    r.set_state(H, E)
    r.time = 0.731235
    r.step(t, 4)
    b = r.get_state()
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    # single precision
    s.set_state(H, E)
    s.restart_out(path, dumpno=3, single=True)
    assert path.read_bytes() == oracle.restart_bytes(s.istep, 3, s.mesh.n1, s.time, s.dt, xyz, E, H, double=False)
    s.set_state(np.zeros((3, n)), np.zeros((3, n)))
    s.restart_swap(path)
    gH, gE = s.get_state()
    assert np.array_equal(gH, H.astype(np.float32).astype(np.float64))
    assert np.array_equal(gE, E.astype(np.float32).astype(np.float64))
    # a file of another order or size is refused
    other = _solver(N=6)
    with pytest.raises(_lib.NkcError, match="nx1"):
        other.restart_swap(path)
    other.close()
    small = _solver(nel=(1, 1, 1))
    with pytest.raises(_lib.NkcError, match="POINTS"):
        small.restart_swap(path)
    small.close()
    with pytest.raises(_lib.NkcError):
        s.restart_swap(tmp_path / "missing.vtk")
    s.close()
    r.close()


def test_reference_restart_case(tmp_path):
    """tests/restart/restart.usr: fields filled with 1.0 and written
    (cem_out at uservp), then filled with 2.0; after restart_swap userchk
    finds 1.0 in every field to L2 and Linf <= 1e-15."""
    print("[case] tests/restart semantics", flush=True)
    s = MaxwellSolver.from_box((2, 2, 2), 7, (0, 0, 0), (1, 1, 1), ("P", "P", "P"))
    n = s.npts
    s.set_state(np.ones((3, n)), np.ones((3, n)))
    path = tmp_path / "restart-mpi-binary-N1-t00000.vtk"
    s.dt = 1e-3
    s.restart_out(path, dumpno=0)
    s.set_state(2 * np.ones((3, n)), 2 * np.ones((3, n)))
    s.restart_swap(path)
    gH, gE = s.get_state()
    bm = np.ones(n)
    for u in list(gH) + list(gE):
        l2, li = cem_error(u, np.ones(n), bm)
        assert l2 <= 1e-15 and li <= 1e-15
    s.close()
