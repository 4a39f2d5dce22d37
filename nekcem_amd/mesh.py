"""Box-mesh setup in the reference's layout (host side, numpy).

Produces, for a box of ``nelx x nely x nelz`` hexahedra with periodic or PEC
walls, every array NekCEM's setup leaves in COMMON for the Maxwell RK path
and that ``nkc_create`` uploads:

* GLL nodes/weights and derivative matrix (ZWGLL / DGLL,
  src/nek5_speclib.F:107-121, 807-838): interior nodes are the roots of
  P_N' (numpy Legendre), w = 2/(N(N+1)P_N^2), D_ij = P_N(z_i)/(P_N(z_j)
  (z_i - z_j)), D_00 = -N(N+1)/4, D_NN = N(N+1)/4;
* element order e = ex + nelx*(ey + nely*ez) and straight-sided coordinates
  (genbox2 / gen_gtp_vertex, src/nek5_genbox.F:2-145);
* cofactor metrics, Jacobian and mass from D applied to the coordinates
  (XYZRST / GLMAPM1 / GEODAT1, src/nek5_coef.F:555-635, 755-758, 877-930);
* face normals and area weights (AREA3 + UNITVEC, src/nek5_coef.F:1150-1237);
* cemface (cem_set_fc_ptr, src/cem_common.F:214-283);
* the face partner map that replaces the gs handle gsh_face: face node
  (a,b) of a face pairs with node (a,b) of the facing face of the neighbour
  (periodic directions wrap), the pairing setup_dgds2 produces for a box
  (src/nek5_connect11.F:2179-2347);
* the PEC face-node list (cem_maxwell_pec_init, src/cem_maxwell.F:1338-1366)
  and the impedance arrays Z_0, Z_1, Y_0, Y_1 (cem_maxwell_materials,
  src/cem_maxwell.F:262-325);
* dxmin and dt (get_dxmin, src/nek5_courant.F:2-79; set_dt,
  src/cem_drive.F:351-395).

Index arrays are 1-based int32 as in the Fortran COMMON blocks.
"""
from __future__ import annotations

import numpy as np
from numpy.polynomial import legendre as npleg

# Ed's face order -> (axis, side): 1 s=-1 (y-), 2 r=+1 (x+), 3 s=+1 (y+),
# 4 r=-1 (x-), 5 t=-1 (z-), 6 t=+1 (z+)  (facind, src/nek5_connect11.F:1921-1935)
ED_FACES = ((1, 0), (0, 1), (1, 1), (0, 0), (2, 0), (2, 1))


def gll(n1):
    """GLL nodes and weights on [-1, 1]."""
    N = n1 - 1
    c = np.zeros(N + 1)
    c[N] = 1.0
    inner = np.sort(npleg.legroots(npleg.legder(c))) if N > 1 else np.zeros(0)
    z = np.concatenate(([-1.0], inner, [1.0]))
    z = 0.5 * (z - z[::-1])            # exact symmetry
    if n1 % 2:
        z[N // 2] = 0.0
    PN = npleg.legval(z, c)
    w = 2.0 / (N * (N + 1.0) * PN * PN)
    return z, w


def dgll(z):
    """Derivative matrix D(i,j) = dl_j/dr(z_i) (DGLL)."""
    n1 = len(z)
    N = n1 - 1
    c = np.zeros(N + 1)
    c[N] = 1.0
    P = npleg.legval(z, c)
    D = np.zeros((n1, n1))
    for i in range(n1):
        for j in range(n1):
            if i != j:
                D[i, j] = P[i] / (P[j] * (z[i] - z[j]))
    d0 = N * (N + 1.0) / 4.0
    D[0, 0] = -d0
    D[N, N] = d0
    return D


def skp_face_nodes(n1):
    """In-element volume index of in-face node q = a + n1*b for each of Ed's
    faces 1..6 (rows), in skpdat order (src/nek5_connect11.F:1483-1528)."""
    a = np.tile(np.arange(n1), n1)
    b = np.repeat(np.arange(n1), n1)
    m, n2 = n1 - 1, n1 * n1
    out = np.empty((6, n2), dtype=np.int64)
    out[0] = a + n2 * b                 # y-: (a, 0, b)
    out[1] = m + n1 * a + n2 * b        # x+: (N, a, b)
    out[2] = a + n1 * m + n2 * b        # y+: (a, N, b)
    out[3] = n1 * a + n2 * b            # x-: (0, a, b)
    out[4] = a + n1 * b                 # z-: (a, b, 0)
    out[5] = a + n1 * b + n2 * m        # z+: (a, b, N)
    return out


def box_lines(nel, N, lo, hi):
    """Per-axis GLL coordinates of a box: three (nel_ax, n1) arrays."""
    z, _ = gll(N + 1)
    out = []
    for ax in range(3):
        ea = np.arange(nel[ax])
        v0 = lo[ax] + (hi[ax] - lo[ax]) * ea / nel[ax]
        v1 = lo[ax] + (hi[ax] - lo[ax]) * (ea + 1) / nel[ax]
        out.append(v0[:, None] * (0.5 * (1 - z))[None, :] + v1[:, None] * (0.5 * (1 + z))[None, :])
    return out


def box_coords(nel, N, lo=(0.0, 0.0, 0.0), hi=(2 * np.pi,) * 3):
    """xm1, ym1, zm1 (flat, reference volume order) of a box without
    building the other reference arrays."""
    n1 = N + 1
    nelx, nely, nelz = nel
    lx, ly, lz = box_lines(nel, N, lo, hi)
    # (ez, ey, ex, k, j, i) -> element-major, i fastest
    X = np.broadcast_to(lx[None, None, :, None, None, :], (nelz, nely, nelx, n1, n1, n1))
    Y = np.broadcast_to(ly[None, :, None, None, :, None], (nelz, nely, nelx, n1, n1, n1))
    Z = np.broadcast_to(lz[:, None, None, :, None, None], (nelz, nely, nelx, n1, n1, n1))
    return [np.ascontiguousarray(a).reshape(-1) for a in (X, Y, Z)]


def box_dxmin(nel, N, lo=(0.0, 0.0, 0.0), hi=(2 * np.pi,) * 3):
    """get_dxmin of a uniform box (interior-node half spacings)."""
    lines = box_lines(nel, N, lo, hi)
    d2 = min(float(np.min((l[:, 2:] - l[:, :-2]) ** 2)) for l in lines)
    return float(np.sqrt(d2) / 2.0)


def face_geometry(n1, nelt, wxm1, deriv):
    """cemface (1-based) and the unit outward normals and area weights of
    every face node (AREA3 + UNITVEC, src/nek5_coef.F:1150-1237) in the
    reference's face order, from the nine coordinate derivatives xr, xs, xt,
    yr, ys, yt, zr, zs, zt (flat, reference volume order)."""
    n2, n3 = n1 * n1, n1 ** 3
    fn = skp_face_nodes(n1)                                     # (6, n2)
    base = (np.arange(nelt) * n3)[:, None, None]
    vol = base + fn[None, :, :]                                 # (nelt, 6, n2)
    cemface = (vol.reshape(-1) + 1).astype(np.int32)
    xr, xs, xt, yr, ys, yt, zr, zs, zt = (a[vol] for a in deriv)
    unx, uny, unz, area = (np.zeros((nelt, 6, n2)) for _ in range(4))
    wgt = np.tile(wxm1, n1) * np.repeat(wxm1, n1)
    for f, (ax, side) in enumerate(ED_FACES):
        if ax == 0:   u, v = (xs, ys, zs), (xt, yt, zt)      # "R": x_s x x_t
        elif ax == 1: u, v = (xr, yr, zr), (xt, yt, zt)      # "S": x_r x x_t
        else:         u, v = (xr, yr, zr), (xs, ys, zs)      # "T": x_r x x_s
        A = u[1][:, f] * v[2][:, f] - u[2][:, f] * v[1][:, f]
        B = u[2][:, f] * v[0][:, f] - u[0][:, f] * v[2][:, f]
        Cc = u[0][:, f] * v[1][:, f] - u[1][:, f] * v[0][:, f]
        # outward sign per AREA3: R and T faces +A on the + side, S faces
        # +A on the - side
        sgn = (1.0 if side == 1 else -1.0) * (-1.0 if ax == 1 else 1.0)
        area[:, f] = np.sqrt(A * A + B * B + Cc * Cc) * wgt
        unx[:, f], uny[:, f], unz[:, f] = sgn * A, sgn * B, sgn * Cc
    s = np.sqrt(unx * unx + uny * uny + unz * unz)
    return cemface, (unx / s).reshape(-1), (uny / s).reshape(-1), (unz / s).reshape(-1), area.reshape(-1)


def face_materials(mesh, eps=None, mu=None):
    """cem_maxwell_materials (src/cem_maxwell.F:262-325) on any mesh with
    cemface, face_partner and cempec: per-node eps, mu -> (eps, mu, Z_0,
    Z_1, Y_0, Y_1); the face sums are the gs sums over partners, doubled
    on PEC faces."""
    npts = mesh.npts
    eps = np.ones(npts) if eps is None else np.ascontiguousarray(eps, dtype=np.float64)
    mu = np.ones(npts) if mu is None else np.ascontiguousarray(mu, dtype=np.float64)
    vol = mesh.cemface.astype(np.int64) - 1
    zimp = np.sqrt(mu / eps)[vol]
    ycon = np.sqrt(eps / mu)[vol]
    p = mesh.face_partner.astype(np.int64) - 1
    has = p >= 0
    Z0 = zimp.copy()
    Y0 = ycon.copy()
    Z0[has] += zimp[p[has]]          # gs_op_fields(gsh_face, Z_0, add)
    Y0[has] += ycon[p[has]]
    pecmask = np.zeros(mesh.nxzfl, dtype=bool)
    pecmask[mesh.cempec - 1] = True
    Z0[pecmask] *= 2.0
    Y0[pecmask] *= 2.0
    Z1 = Z0 - zimp
    Y1 = Y0 - ycon
    return eps, mu, 0.5 * Z0, Z1, 0.5 * Y0, Y1


class BoxMesh:
    """Reference-layout arrays for a box mesh.

    bc: per direction 'P' (periodic), 'PEC', or 'PML' (a PEC wall behind an
    absorbing layer).
    """

    def __init__(self, nel, N, lo=(0.0, 0.0, 0.0), hi=(2 * np.pi,) * 3, bc=("P", "P", "P")):
        self.nel = tuple(int(x) for x in nel)
        self.N = int(N)
        self.n1 = n1 = self.N + 1
        self.lo, self.hi = tuple(map(float, lo)), tuple(map(float, hi))
        self.bc = tuple(bc)
        nelx, nely, nelz = self.nel
        self.nelt = nelt = nelx * nely * nelz
        n2, n3 = n1 * n1, n1 ** 3
        self.npts = npts = n3 * nelt
        self.nxzfl = 6 * n2 * nelt
        self.z, self.wxm1 = gll(n1)
        self.D = dgll(self.z)
        self.dxm1 = np.ascontiguousarray(self.D.reshape(-1, order="F"))
        w = self.wxm1
        self.w3mn = (w[None, None, :] * w[None, :, None] * w[:, None, None]).reshape(-1)

        e = np.arange(nelt)
        self.exyz = (e % nelx, (e // nelx) % nely, e // (nelx * nely))
        lines = []
        for ax in range(3):
            ea, nea = self.exyz[ax], self.nel[ax]
            v0 = self.lo[ax] + (self.hi[ax] - self.lo[ax]) * ea / nea
            v1 = self.lo[ax] + (self.hi[ax] - self.lo[ax]) * (ea + 1) / nea
            lines.append(v0[:, None] * (0.5 * (1 - self.z))[None, :]
                         + v1[:, None] * (0.5 * (1 + self.z))[None, :])
        # (nelt, k, j, i) views of the coordinates
        self._lines = lines
        self._geometry()
        self._faces()

    # -- coordinates ----------------------------------------------------
    def coords(self):
        """xm1, ym1, zm1 as flat npts arrays (built on demand)."""
        n1, nelt = self.n1, self.nelt
        out = []
        for ax, line in enumerate(self._lines):
            shape = [nelt, 1, 1, 1]
            shape[3 - ax] = n1
            out.append(np.ascontiguousarray(
                np.broadcast_to(line.reshape(shape), (nelt, n1, n1, n1))).reshape(-1))
        return out

    # -- metrics (XYZRST + GLMAPM1 + mass) --------------------------------
    def _geometry(self):
        n1, nelt, D = self.n1, self.nelt, self.D
        # x depends only on i, y only on j, z only on k for box elements, so
        # the nine coordinate derivatives reduce to per-element lines; they
        # are evaluated with D exactly as local_grad3 would.
        dx = self._lines[0] @ D.T      # x_r(i)   (nelt, n1)
        dy = self._lines[1] @ D.T      # y_s(j)
        dz = self._lines[2] @ D.T      # z_t(k)
        # the off-diagonal derivatives (e.g. x_s) are x * rowsum(D): keep them
        rs = D.sum(axis=1)             # (n1,)
        X, Y, Z = self._lines

        def outer(fi, fj, fk):
            # f(i,j,k) = fi[e,i] * fj[e,j] * fk[e,k] with broadcasting
            return (fi[:, None, None, :] * fj[:, None, :, None] * fk[:, :, None, None]).reshape(-1)

        one = np.ones((nelt, n1))
        rsv = np.broadcast_to(rs, (nelt, n1))
        xr = outer(dx, one, one)
        xs = outer(X, rsv, one)
        xt = outer(X, one, rsv)
        yr = outer(rsv, Y, one)
        ys = outer(one, dy, one)
        yt = outer(one, Y, rsv)
        zr = outer(rsv, one, Z)
        zs = outer(one, rsv, Z)
        zt = outer(one, one, dz)
        jac = (xr * ys * zt + xt * yr * zs + xs * yt * zr
               - xr * yt * zs - xs * yr * zt - xt * ys * zr)
        self.jacm = jac
        self.rxmn = ys * zt - yt * zs
        self.rymn = xt * zs - xs * zt
        self.rzmn = xs * yt - xt * ys
        self.sxmn = yt * zr - yr * zt
        self.symn = xr * zt - xt * zr
        self.szmn = xt * yr - xr * yt
        self.txmn = yr * zs - ys * zr
        self.tymn = xs * zr - xr * zs
        self.tzmn = xr * ys - xs * yr
        self.bmn = jac * np.tile(self.w3mn, nelt)
        self._deriv = (xr, xs, xt, yr, ys, yt, zr, zs, zt)

    # -- faces (AREA3, cem_set_fc_ptr, partner map, PEC list) -------------
    def _faces(self):
        n1, nelt = self.n1, self.nelt
        n2, n3 = n1 * n1, n1 ** 3
        self.cemface, self.unxm, self.unym, self.unzm, self.aream = face_geometry(n1, nelt, self.wxm1,
                                                                                  self._deriv)
        del self._deriv

        # partner map (1-based, 0 = none) and PEC list
        partner = np.zeros((nelt, 6, n2), dtype=np.int64)
        pec = np.zeros((nelt, 6), dtype=bool)
        q = np.arange(n2)
        for f, (ax, side) in enumerate(ED_FACES):
            ea, nea = self.exyz[ax], self.nel[ax]
            step = 1 if side == 1 else -1
            nb_idx = [self.exyz[0].copy(), self.exyz[1].copy(), self.exyz[2].copy()]
            nb_idx[ax] = ea + step
            if self.bc[ax] == "P":
                nb_idx[ax] %= nea
                valid = np.ones(nelt, dtype=bool)
            else:
                valid = (nb_idx[ax] >= 0) & (nb_idx[ax] < nea)
                pec[~valid, f] = self.bc[ax] in ("PEC", "PML")   # cem_maxwell_pec_init lists both
            ne = nb_idx[0] + self.nel[0] * (nb_idx[1] + self.nel[1] * nb_idx[2])
            opp = ED_FACES.index((ax, 1 - side))
            slot = (ne[:, None] * 6 + opp) * n2 + q[None, :] + 1
            partner[:, f, :] = np.where(valid[:, None], slot, 0)
        self.face_partner = partner.reshape(-1).astype(np.int32)
        self.pec_face = pec
        pecslots = (np.arange(self.nxzfl).reshape(nelt, 6, n2))[pec]
        self.cempec = (pecslots.reshape(-1) + 1).astype(np.int32)
        self.ncempec = len(self.cempec)

    # -- materials (cem_maxwell_materials) --------------------------------
    def materials(self, eps=None, mu=None):
        """Per-node permittivity/permeability -> (eps, mu, Z_0, Z_1, Y_0, Y_1)."""
        return face_materials(self, eps, mu)

    # -- time step --------------------------------------------------------
    def dxmin(self):
        """get_dxmin: half the smallest distance between the two neighbours
        of an interior node along an index direction."""
        if self.n1 < 3:
            raise ValueError("dxmin needs interior nodes")
        d2 = np.inf
        for ax in range(3):
            diff = self._lines[ax][:, 2:] - self._lines[ax][:, :-2]
            d2 = min(d2, float(np.min(diff * diff)))
        return float(np.sqrt(d2) / 2.0)

    def dt(self, p12):
        """set_dt: dt = |p12| if p12 < 0 else p12*dxmin (Maxwell branch)."""
        if p12 == 0:
            raise ValueError("param(12) must be nonzero")
        return abs(p12) if p12 < 0 else p12 * self.dxmin()

    def element_field(self, per_element):
        """Expand a per-element array to per-node (npts)."""
        return np.repeat(np.asarray(per_element, dtype=np.float64), self.n1 ** 3)
