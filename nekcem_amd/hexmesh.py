"""General hexahedral meshes in the reference's layout (host side, numpy).

For meshes the box setup does not cover -- curved sides, rotated or sheared
elements, faces paired in any orientation -- this builds every array
NekCEM's setup leaves in COMMON for the Maxwell RK path, which
``MaxwellSolver`` hands to nkc_create:

* node coordinates from the element corners and curved-side data NekCEM's
  readers fill: the trilinear map of the 8 corners (GENXYZ,
  src/nek5_genxyz.F:562-680), then for every circular-arc edge ('C') the
  arc's displacement from the straight edge, blended linearly across the
  element (ARCSRF, src/nek5_genxyz.F:2-101);
* cofactor metrics, Jacobian and mass from D applied to the coordinates
  (GLMAPM1 / GEODAT1, src/nek5_coef.F:555-635, 755-758), face normals and
  area weights (AREA3 + UNITVEC, :1150-1237), cemface (cem_set_fc_ptr);
* the face partner map that replaces gsh_face: two faces face each other
  when the boundary data links them (an 'E' or 'P' face names its
  neighbour element and face) or, for an 'E' face without a link, when
  their centroids coincide; their nodes pair at equal coordinates, periodic
  faces after the translation between the two centroids -- the pairing gs
  sets up from the global vertex numbering (setup_dgds2,
  src/nek5_connect11.F:2179-2347);
* the PEC list ('PEC' and 'PML' faces, cem_maxwell_pec_init), the impedance
  arrays (cem_maxwell_materials) and dxmin (get_dxmin, src/nek5_courant.F:
  2-79).
"""
from __future__ import annotations

import numpy as np

from .mesh import ED_FACES, dgll, face_geometry, face_materials, gll, skp_face_nodes

# Preprocessor corner q (0-based) -> (r, s, t) side (0: -1, 1: +1): corners
# 1-4 on t = -1 counter-clockwise from (r, s) = (-1, -1), 5-8 above them.
CORNER_RST = ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))
# EFACE1 (src/nek5_connect11.F:1076-1081): preprocessor face -> symmetric face
EFACE1 = (3, 2, 4, 1, 5, 6)


def _arc(X, Y, xc, yc, isid, radius, z, h):
    """ARCSRF for edge isid (1..8) of one element: add to X, Y (n1, n1, n1)
    [k, j, i] the displacement of the circular arc of the given radius
    through the edge's two corners from the straight edge, blended across
    the element."""
    n1 = len(z)
    p1 = (xc[isid - 1], yc[isid - 1])
    q = 0 if isid == 4 else (4 if isid == 8 else isid)
    p2 = (xc[q], yc[q])
    gap = np.hypot(p1[0] - p2[0], p1[1] - p2[1])
    if abs(2.0 * radius) <= gap * 1.00001:
        raise ValueError(f"radius {radius} too small for an edge of length {gap}")
    xs, ys = p2[1] - p1[1], p1[0] - p2[0]
    xys = np.hypot(xs, ys)
    dth = abs(np.arcsin(0.5 * gap / radius))
    mx, my = 0.5 * (p1[0] + p2[0]), 0.5 * (p1[1] + p2[1])
    xcen = mx - xs / xys * radius * np.cos(dth)
    ycen = my - ys / xys * radius * np.cos(dth)
    th0 = np.arctan2(my - ycen, mx - xcen)
    side = (isid - 1) % 4 + 1
    xcr, ycr = np.zeros(n1), np.zeros(n1)
    for ix in range(n1):
        ixt = n1 - 1 - ix if side > 2 else ix
        r = -z[ix] if radius < 0 else z[ix]
        xcr[ixt] = xcen + abs(radius) * np.cos(th0 + r * dth) - (h[0][ix] * p1[0] + h[1][ix] * p2[0])
        ycr[ixt] = ycen + abs(radius) * np.sin(th0 + r * dth) - (h[0][ix] * p1[1] + h[1][ix] * p2[1])
    sym = EFACE1[side - 1]
    ht = h[(isid - 1) // 4]
    if sym <= 2:
        hr = h[sym - 1]
        X += ht[:, None, None] * xcr[None, :, None] * hr[None, None, :]
        Y += ht[:, None, None] * ycr[None, :, None] * hr[None, None, :]
    else:
        hs = h[sym - 3]
        X += ht[:, None, None] * hs[None, :, None] * xcr[None, None, :]
        Y += ht[:, None, None] * hs[None, :, None] * ycr[None, None, :]


def genxyz(xc, yc, zc, n1, curves=()):
    """Node coordinates (nelt, n1, n1, n1) [e, k, j, i] of the elements with
    corners xc, yc, zc (nelt, 8) and curved edges curves = [(side 1..8,
    element 0-based, parameters, type), ...]; only circular arcs ('C') are
    displaced, other types are refused."""
    z, _ = gll(n1)
    h = (0.5 * (1.0 - z), 0.5 * (1.0 + z))
    xc, yc, zc = (np.asarray(a, dtype=np.float64).reshape(-1, 8) for a in (xc, yc, zc))
    nelt = len(xc)
    out = []
    for c in (xc, yc, zc):
        X = np.zeros((nelt, n1, n1, n1))
        for q, (a, b, t) in enumerate(CORNER_RST):
            X += c[:, q, None, None, None] * (h[t][:, None, None] * h[b][None, :, None] * h[a][None, None, :])[None]
        out.append(X)
    for side, e, par, kind in curves:
        if kind in ("", " "):
            continue
        if kind != "C":
            raise ValueError(f"curved side type {kind!r} (element {e + 1}, side {side}): only circular arcs "
                             "('C') are built")
        if not 1 <= side <= 8:
            raise ValueError(f"arc on side {side} of element {e + 1}")
        _arc(out[0][e], out[1][e], xc[e], yc[e], side, float(par[0]), z, h)
    return out


class HexMesh:
    """Reference-layout arrays of a general hexahedral mesh of order N.

    coords: xm1, ym1, zm1, each (nelt, (N+1)^3) or flat, in the reference's
    volume order.  links (nelt, 6, 2): the neighbour element and face
    (0-based, reference face order) across each face, -1 for none, or
    (-1, -2) for "find the face whose centroid coincides" (an 'E' face
    without a link).  pec (nelt, 6) bool: PEC walls.
    """

    def __init__(self, N, xm1, ym1, zm1, links, pec):
        self.N = int(N)
        self.n1 = n1 = self.N + 1
        n2, n3 = n1 * n1, n1 ** 3
        X = np.asarray(xm1, dtype=np.float64).reshape(-1, n1, n1, n1)
        Y = np.asarray(ym1, dtype=np.float64).reshape(-1, n1, n1, n1)
        Z = np.asarray(zm1, dtype=np.float64).reshape(-1, n1, n1, n1)
        self.nelt = nelt = len(X)
        self.npts = nelt * n3
        self.nxzfl = 6 * n2 * nelt
        self.z, self.wxm1 = gll(n1)
        self.D = D = dgll(self.z)
        self.dxm1 = np.ascontiguousarray(D.reshape(-1, order="F"))
        w = self.wxm1
        self.w3mn = (w[None, None, :] * w[None, :, None] * w[:, None, None]).reshape(-1)
        self._xyz = (X.reshape(-1).copy(), Y.reshape(-1).copy(), Z.reshape(-1).copy())

        # local_grad3 of the coordinates: d/dr along i, d/ds along j, d/dt along k
        def grad(U):
            return (np.einsum("il,ekjl->ekji", D, U).reshape(-1), np.einsum("jl,ekli->ekji", D, U).reshape(-1),
                    np.einsum("kl,elji->ekji", D, U).reshape(-1))

        xr, xs, xt = grad(X)
        yr, ys, yt = grad(Y)
        zr, zs, zt = grad(Z)
        jac = (xr * ys * zt + xt * yr * zs + xs * yt * zr - xr * yt * zs - xs * yr * zt - xt * ys * zr)
        if not np.all(jac > 0):
            raise ValueError("element map with non-positive Jacobian (left-handed or folded element)")
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
        self.cemface, self.unxm, self.unym, self.unzm, self.aream = face_geometry(
            n1, nelt, w, (xr, xs, xt, yr, ys, yt, zr, zs, zt))
        self._pair(np.asarray(links, dtype=np.int64).reshape(nelt, 6, 2), np.asarray(pec, dtype=bool).reshape(nelt, 6))

    # -- face pairing ---------------------------------------------------------
    def _face_xyz(self, e, f):
        fn = skp_face_nodes(self.n1)[f] + e * self.n1 ** 3
        return np.stack([c[fn] for c in self._xyz], axis=1)          # (n2, 3)

    def _pair(self, links, pec):
        n1, nelt = self.n1, self.nelt
        n2 = n1 * n1
        cen = np.array([[self._face_xyz(e, f).mean(axis=0) for f in range(6)] for e in range(nelt)])
        size = float(np.max(np.ptp(np.stack(self._xyz), axis=1)))
        tol = 1e-9 * max(size, 1.0)
        # 'E' faces without a link: the face whose centroid coincides
        key = {}
        for e in range(nelt):
            for f in range(6):
                key.setdefault(tuple(np.round(cen[e, f] / tol * 1e-3).astype(np.int64)), []).append((e, f))
        partner = np.zeros((nelt, 6, n2), dtype=np.int64)
        for e in range(nelt):
            for f in range(6):
                ne, nf = links[e, f]
                if ne == -1 and nf == -2:
                    cands = [(a, b) for a, b in key.get(tuple(np.round(cen[e, f] / tol * 1e-3).astype(np.int64)), [])
                             if (a, b) != (e, f)]
                    if len(cands) != 1:
                        raise ValueError(f"face {f + 1} of element {e + 1}: {len(cands)} faces at its centroid")
                    ne, nf = cands[0]
                if ne < 0:
                    continue
                if not (0 <= ne < nelt and 0 <= nf < 6):
                    raise ValueError(f"face {f + 1} of element {e + 1} names element {ne + 1} face {nf + 1}")
                P, Q = self._face_xyz(e, f), self._face_xyz(ne, nf)
                shift = cen[ne, nf] - cen[e, f]                       # 0, or the period for 'P' faces
                d = np.linalg.norm((P + shift)[:, None, :] - Q[None, :, :], axis=2)
                b = np.argmin(d, axis=1)
                if np.max(d[np.arange(n2), b]) > 1e-7 * max(size, 1.0) or len(set(b.tolist())) != n2:
                    raise ValueError(f"face {f + 1} of element {e + 1} and face {nf + 1} of element {ne + 1} "
                                     "do not match node for node")
                partner[e, f] = (ne * 6 + nf) * n2 + b + 1
        flat = partner.reshape(-1)
        has = np.nonzero(flat)[0]
        if np.any(flat[flat[has] - 1] != has + 1):
            raise ValueError("face links are not symmetric")
        self.face_partner = flat.astype(np.int32)
        self.pec_face = pec & (partner[:, :, 0] == 0)
        pecslots = np.arange(self.nxzfl).reshape(nelt, 6, n2)[self.pec_face]
        self.cempec = (pecslots.reshape(-1) + 1).astype(np.int32)
        self.ncempec = len(self.cempec)

    # -- the same interface as BoxMesh --------------------------------------
    def coords(self):
        return [c.copy() for c in self._xyz]

    def materials(self, eps=None, mu=None):
        return face_materials(self, eps, mu)

    def dxmin(self):
        """get_dxmin: half the smallest distance between the two neighbours
        of an interior node along an index direction."""
        n1 = self.n1
        X = np.stack([c.reshape(-1, n1, n1, n1) for c in self._xyz])      # (3, e, k, j, i)
        inner = slice(1, n1 - 1)
        d2 = np.inf
        for ax in (4, 3, 2):
            sl_p = [slice(None), slice(None), inner, inner, inner]
            sl_m = list(sl_p)
            sl_p[ax] = slice(2, n1)
            sl_m[ax] = slice(0, n1 - 2)
            diff = X[tuple(sl_p)] - X[tuple(sl_m)]
            d2 = min(d2, float(np.min(np.sum(diff * diff, axis=0))))
        return float(np.sqrt(d2) / 2.0)

    def dt(self, p12):
        if p12 == 0:
            raise ValueError("param(12) must be nonzero")
        return abs(p12) if p12 < 0 else p12 * self.dxmin()

    # -- from NekCEM's case arrays ------------------------------------------
    @classmethod
    def from_corners(cls, N, xc, yc, zc, cbc, bc, curves=(), usrdat2=None):
        """The mesh of corners, curved sides and boundary data as NekCEM's
        readers fill them (rdmesh / rdcurve / rdbdry); usrdat2(x, y, z) may
        move the node coordinates (flat arrays) as a case's usrdat2 does."""
        xc, yc, zc = (np.asarray(a, dtype=np.float64).reshape(-1, 8) for a in (xc, yc, zc))
        nelt = len(xc)
        X, Y, Z = genxyz(xc, yc, zc, int(N) + 1, curves)
        x, y, z = X.reshape(-1), Y.reshape(-1), Z.reshape(-1)
        if usrdat2 is not None:
            x, y, z = usrdat2(x, y, z)
        cbc = np.asarray(cbc).reshape(nelt, 6)
        bc = np.asarray(bc, dtype=np.float64).reshape(nelt, 6, 5)
        links = np.full((nelt, 6, 2), -1, dtype=np.int64)
        pec = np.zeros((nelt, 6), dtype=bool)
        for e in range(nelt):
            for f in range(6):
                code = str(cbc[e, f]).strip().upper()
                if code in ("PEC", "PML"):
                    pec[e, f] = True
                elif code in ("E", "P"):
                    ne, nf = int(round(bc[e, f, 0])) - 1, int(round(bc[e, f, 1])) - 1
                    if ne < 0:
                        if code == "P":
                            raise ValueError(f"periodic face {f + 1} of element {e + 1} has no partner")
                        links[e, f] = (-1, -2)
                    else:
                        links[e, f] = (ne, nf)
                else:
                    raise ValueError(f"boundary code {code!r} (element {e + 1} face {f + 1}) is not on this path")
        return cls(N, x, y, z, links, pec)

    @classmethod
    def from_coords(cls, N, xm1, ym1, zm1, periods=(), walls="PEC"):
        """The mesh of the given node coordinates (flat or (nelt, (N+1)^3)):
        faces whose centroids coincide, directly or after one of the
        periodic translations in periods, face each other; every other face
        is a wall of type walls ('PEC', or None for no partner)."""
        n1 = int(N) + 1
        X = [np.asarray(a, dtype=np.float64).reshape(-1) for a in (xm1, ym1, zm1)]
        nelt = len(X[0]) // n1 ** 3
        fn = skp_face_nodes(n1)
        cen = np.stack([np.stack([c[fn + e * n1 ** 3].mean(axis=1) for c in X], axis=1) for e in range(nelt)])
        scale = max(1.0, float(max(np.max(np.abs(c)) for c in X)))
        q = 1e-7 * scale
        index = {}
        for e in range(nelt):
            for f in range(6):
                index.setdefault(tuple(np.floor(cen[e, f] / q + 0.5).astype(np.int64)), []).append((e, f))
        links = np.full((nelt, 6, 2), -1, dtype=np.int64)
        pec = np.zeros((nelt, 6), dtype=bool)
        shifts = [np.zeros(3)] + [s for T in np.asarray(periods, dtype=np.float64).reshape(-1, 3) for s in (T, -T)]
        for e in range(nelt):
            for f in range(6):
                for T in shifts:
                    other = [ef for ef in index.get(tuple(np.floor((cen[e, f] + T) / q + 0.5).astype(np.int64)), [])
                             if ef != (e, f)]
                    if other:
                        links[e, f] = other[0]
                        break
                else:
                    pec[e, f] = walls == "PEC"
        return cls(N, X[0], X[1], X[2], links, pec)
