"""General hexahedral test meshes: box elements in rotated local frames, a
smoothly deformed (curved) box, and the reference's cylwave case."""
import os

import numpy as np

from nekcem_amd import rea, usr
from nekcem_amd.hexmesh import HexMesh
from nekcem_amd.mesh import box_coords

CASES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cases")


def rotate_elements(N, arrays, which, seed=0):
    """Renumber the nodes of the listed elements under a random proper
    rotation of the index cube (an axis permutation with flips, determinant
    +1): each element stays where it is in space, but its r, s, t axes, its
    faces and their in-face orientations change.  arrays: per-node arrays
    (flat, reference order) to renumber alike."""
    n1 = N + 1
    rng = np.random.default_rng(seed)
    out = [np.array(a).reshape(-1, n1, n1, n1).copy() for a in arrays]      # [e, k, j, i]
    for e in which:
        p = rng.permutation(3)
        flips = rng.integers(0, 2, 3)
        det = np.linalg.det(np.eye(3)[p]) * np.prod(np.where(flips == 1, -1.0, 1.0))
        if det < 0:
            flips[0] ^= 1
        for c in out:
            A = c[e].transpose(2, 1, 0).transpose(p)     # [i, j, k], new axis a is old axis p[a]
            for a in range(3):
                if flips[a]:
                    A = np.flip(A, axis=a)
            c[e] = A.transpose(2, 1, 0)
    return [c.reshape(-1) for c in out]


def rotated_box(nel, N, lo, hi, bc, which, seed=0):
    """(HexMesh, index map, periods): a box whose listed elements have
    rotated local frames; index[n] is the node of the plain box mesh that
    node n of the rotated mesh sits on."""
    x, y, z = box_coords(nel, N, lo, hi)
    idx = np.arange(len(x), dtype=np.float64)
    x, y, z, idx = rotate_elements(N, (x, y, z, idx), which, seed)
    periods = [np.eye(3)[a] * (hi[a] - lo[a]) for a in range(3) if bc[a] == "P"]
    mesh = HexMesh.from_coords(N, x, y, z, periods=periods)
    return mesh, idx.astype(np.int64), periods


def curved_box(nel, N, amp=0.08, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
    """A PEC box deformed by a smooth map that fixes its walls' planes
    (x + amp sin(pi y) sin(pi z) ... on the unit cube): curved, non-affine
    elements with conforming faces."""
    x, y, z = box_coords(nel, N, lo, hi)
    u = [(v - l) / (h - l) for v, l, h in zip((x, y, z), lo, hi)]
    s = [np.sin(np.pi * w) for w in u]
    x2 = x + amp * (hi[0] - lo[0]) * s[1] * s[2] * np.sin(2 * np.pi * u[0])
    y2 = y + amp * (hi[1] - lo[1]) * s[0] * s[2] * np.sin(2 * np.pi * u[1])
    z2 = z + amp * (hi[2] - lo[2]) * s[0] * s[1] * np.sin(2 * np.pi * u[2])
    return HexMesh.from_coords(N, x2, y2, z2)


def cylwave_mesh(N=11):
    """The reference's tests/cylwave mesh at order N: its .rea with the
    case's usrdat (corners onto the circle) and usrdat2 (z scaled); returns
    (HexMesh, radius, z period)."""
    r = rea.read_case(os.path.join(CASES, "cylwave.rea"))
    xc, yc, zc = usr.cylwave_usrdat(r["xc"], r["yc"], r["zc"])
    mesh = HexMesh.from_corners(N, xc, yc, zc, r["cbc"], r["bc"], r["curves"], usrdat2=usr.cylwave_usrdat2)
    z = mesh.coords()[2]
    return mesh, usr.xy_radius(xc, yc), float(np.max(z) - np.min(z)), r["param"]
