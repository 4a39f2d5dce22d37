"""Box-built handles and batched cavities on the HIP path."""
import numpy as np
import pytest

from nekcem_amd import usr
from nekcem_amd.maxwell import BatchSolver, MaxwellSolver
from nekcem_amd.mesh import BoxMesh, box_coords, box_dxmin

pytestmark = pytest.mark.gpu


def test_box_handle_matches_reference_array_handle():
    """nkc_create_box and nkc_create on the same mesh give the same steps."""
    nel, N, bc = (3, 2, 4), 6, ("PEC", "P", "P")
    lo, hi = (0, 0, 0), (1.0, 0.7, 1.3)
    rng = np.random.default_rng(4)
    epse = rng.uniform(1, 4, 24)
    print(f"[case] box vs arrays nel={nel} N={N}", flush=True)
    mesh = BoxMesh(nel, N, lo, hi, bc)
    a = MaxwellSolver(mesh, eps=np.repeat(epse, (N + 1) ** 3))
    b = MaxwellSolver.from_box(nel, N, lo, hi, bc, eps_elem=epse)
    H = rng.standard_normal((3, mesh.npts))
    E = rng.standard_normal((3, mesh.npts))
    dt = mesh.dt(0.2)
    for s in (a, b):
        s.set_state(H, E)
        s.step(dt, 10)
    for x, y in zip(a.get_state(), b.get_state()):
        assert np.max(np.abs(x - y)) <= 1e-12 * np.max(np.abs(x))
    a.close()
    b.close()


def _cavities():
    rng = np.random.default_rng(9)
    out = []
    for n, N in ((2, 5), (3, 7), (2, 9), (4, 6), (3, 8), (2, 7)):
        out.append(dict(nel=(n, n, n), N=N, lo=(0, 0, 0), hi=(1, 1, 1), bc=("PEC",) * 3,
                        eps_elem=rng.uniform(1, 4, n ** 3)))
    return out


def test_batch_matches_single_cavity_handles():
    """Every cavity of a batch (mixed sizes and orders, own dt) evolves
    bit for bit like the same cavity on its own handle."""
    cavs = _cavities()
    print(f"[case] batch of {len(cavs)} cavities", flush=True)
    batch = BatchSolver(cavs)
    assert batch.info()["norders"] == 5
    dts, init = [], []
    for c, cv in enumerate(cavs):
        x, y, z = box_coords(cv["nel"], cv["N"], cv["lo"], cv["hi"])
        H, E = usr.gaussian_ez(x, y, z)
        init.append((H, E))
        batch.set_state(c, H, E)
        dts.append(0.25 * box_dxmin(cv["nel"], cv["N"], cv["lo"], cv["hi"]))
    batch.step(np.array(dts), 15)
    for c, cv in enumerate(cavs):
        s = MaxwellSolver.from_box(cv["nel"], cv["N"], cv["lo"], cv["hi"], cv["bc"], eps_elem=cv["eps_elem"])
        s.set_state(*init[c])
        s.step(dts[c], 15)
        for x, y in zip(s.get_state(), batch.get_state(c)):
            assert np.array_equal(x, y), c
        assert s.time == batch.times[c]
        s.close()
    batch.close()


def test_batch_vs_oracle(oracle):
    cavs = _cavities()[:2]
    batch = BatchSolver(cavs)
    refs, dts = [], []
    for c, cv in enumerate(cavs):
        spec = oracle.BoxSpec(nel=cv["nel"], N=cv["N"], lo=cv["lo"], hi=cv["hi"], bc=cv["bc"])
        n3 = (cv["N"] + 1) ** 3
        ref = oracle.BoxCase(spec, eps=np.repeat(cv["eps_elem"], n3))
        H, E = usr.gaussian_ez(ref.xm1, ref.ym1, ref.zm1)
        batch.set_state(c, H, E)
        dts.append(ref.dt_from_cfl(0.25))
        refs.append((ref, H, E))
    batch.step(np.array(dts), 30)
    for c, (ref, H, E) in enumerate(refs):
        out = ref.run(H, E, dts[c], 30)
        gH, gE = batch.get_state(c)
        for g, o in ((gH, out["H"]), (gE, out["E"])):
            for k in range(3):
                den = np.sqrt(np.sum(ref.bm1 * o[k] ** 2))
                if den > 1e-6:
                    assert np.sqrt(np.sum(ref.bm1 * (g[k] - o[k]) ** 2)) <= 1e-11 * den
    batch.close()
