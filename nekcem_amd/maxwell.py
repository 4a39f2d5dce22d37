"""Host-side mirror of NekCEM's Maxwell RK driver over the HIP step.

Names follow the reference so a NekCEM user finds the same entry points:

* ``cem_maxwell_init(mesh, eps, mu, ifupwind)`` -- cem_maxwell_init +
  cem_maxwell_materials + acc_copy_all_in (src/cem_maxwell.F:65-191,
  262-325; src/cem_drive.F:397): builds the nkc_desc and calls nkc_create;
* ``MaxwellSolver.cem_maxwell_init_fields(userini)`` -- userini hook
  (src/cem_maxwell.F:193-211);
* ``MaxwellSolver.cem_maxwell_op_rk()`` -- one RK5 step
  (src/cem_maxwell.F:327-345);
* ``MaxwellSolver.time_advancing_pde(nsteps, userchk)`` -- the time loop of
  src/cem_drive.F:568-657 with the ``userchk`` hook called after each step,
  which may stop the run like ``exitt(1)`` does;
* ``cem_error(u, exact, bm1)`` -- mass-weighted L2 and Linf
  (src/cem_common.F:1335-1355).

Field arrays are (3, npts) float64 numpy arrays, H = (Hx, Hy, Hz) and
E = (Ex, Ey, Ez), i.e. hn(lpts1,3) / en(lpts1,3) with lpts1 = npts.
Every step runs in libnekcem_hip.so on the GPU; nothing here computes the
operator.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .mesh import BoxMesh

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)

NLAYOUT = 8   # stage-kernel layouts (NKC_NLAYOUT)


def _d(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _i(a):
    return None if a is None else a.ctypes.data_as(_ip)


class RunStopped(RuntimeError):
    """Raised when userchk asks to stop (the reference calls exitt(1))."""


def box_desc(nel, N, lo=(0.0, 0.0, 0.0), hi=(2 * np.pi,) * 3, bc=("P", "P", "P"),
             eps_elem=None, mu_elem=None, ifupwind=True):
    """nkc_box_desc for a box; the per-element arrays are kept alive on the
    returned structure."""
    d = _lib.NkcBoxDesc()
    d.N, d.ifupwind = int(N), 1 if ifupwind else 0
    d.nel = (C.c_int32 * 3)(*[int(x) for x in nel])
    d.pec = (C.c_int32 * 3)(*[1 if b in ("PEC", "PML") else 0 for b in bc])
    d.lo = (C.c_double * 3)(*[float(x) for x in lo])
    d.hi = (C.c_double * 3)(*[float(x) for x in hi])
    keep = []
    for name, arr in (("eps_elem", eps_elem), ("mu_elem", mu_elem)):
        if arr is not None:
            arr = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
            assert len(arr) == int(np.prod(nel))
            keep.append(arr)
            setattr(d, name, _d(arr))
    d._keep = keep
    return d


# reference face order 1..6 -> (axis, side): y-, x+, y+, x-, z-, z+
_ED_FACES = ((1, 0), (0, 1), (1, 1), (0, 0), (2, 0), (2, 1))


def box_pml_faces(nel, bc):
    """(nelt, 6) int8: 1 on the faces of a box (elements x fastest, faces in
    the reference's order) whose direction's bc is 'PML'."""
    nel = tuple(int(n) for n in nel)
    e = np.arange(int(np.prod(nel)))
    ex = (e % nel[0], (e // nel[0]) % nel[1], e // (nel[0] * nel[1]))
    out = np.zeros((len(e), 6), dtype=np.int8)
    for f, (ax, side) in enumerate(_ED_FACES):
        if bc[ax] == "PML":
            out[:, f] = (ex[ax] == 0) if side == 0 else (ex[ax] == nel[ax] - 1)
    return out


class _BoxMeshInfo:
    """What a box-built solver knows of its mesh (no per-node arrays)."""

    def __init__(self, nel, N, lo, hi, bc):
        from .mesh import box_coords, box_dxmin, gll
        self.nel, self.N, self.lo, self.hi, self.bc = tuple(nel), int(N), tuple(lo), tuple(hi), tuple(bc)
        self.n1 = self.N + 1
        self.nelt = int(np.prod(nel))
        self.npts = self.nelt * self.n1 ** 3
        self._coords, self._dxmin = box_coords, box_dxmin

    def coords(self):
        return self._coords(self.nel, self.N, self.lo, self.hi)

    def pml_faces(self):
        return box_pml_faces(self.nel, self.bc)

    def dxmin(self):
        return self._dxmin(self.nel, self.N, self.lo, self.hi)

    def dt(self, p12):
        return abs(p12) if p12 < 0 else p12 * self.dxmin()


class _HexMeshInfo:
    """What a solver built from element corners knows of its mesh: per
    element lower/upper corners of an axis-aligned box."""

    def __init__(self, N, lo, hi, cbc=None):
        from .mesh import gll
        self.N, self.n1 = int(N), int(N) + 1
        self.lo, self.hi = np.asarray(lo, dtype=np.float64), np.asarray(hi, dtype=np.float64)
        self.cbc = None if cbc is None else np.asarray(cbc).reshape(len(self.lo), 6)
        self.nelt = len(self.lo)
        self.npts = self.nelt * self.n1 ** 3
        self._z, self._w = gll(self.n1)

    def coords(self):
        """xm1, ym1, zm1 in the reference volume order (element-major, i fastest)."""
        n1, z = self.n1, self._z
        t = 0.5 * (1.0 + z)
        out = []
        for a, shape in ((0, (1, 1, n1)), (1, (1, n1, 1)), (2, (n1, 1, 1))):
            lo, hi = self.lo[:, a], self.hi[:, a]
            v = lo[:, None] * (1.0 - t)[None, :] + hi[:, None] * t[None, :]
            v = np.broadcast_to(v.reshape((self.nelt,) + shape), (self.nelt, n1, n1, n1))
            out.append(np.ascontiguousarray(v).reshape(-1))
        return out

    def pml_faces(self):
        """(nelt, 6) int8: 1 where the boundary code is 'PML'."""
        if self.cbc is None:
            return np.zeros((len(self.lo), 6), dtype=np.int8)
        return np.vectorize(lambda c: str(c).strip().upper() == "PML")(self.cbc).astype(np.int8)

    @property
    def bmn(self):
        jac = np.prod(0.5 * (self.hi - self.lo), axis=1)
        w3 = (self._w[None, None, :] * self._w[None, :, None] * self._w[:, None, None]).reshape(-1)
        return (jac[:, None] * w3[None, :]).reshape(-1)

    def dxmin(self):
        """get_dxmin: half the smallest interior-node spacing over two GLL gaps."""
        z = self._z
        d = np.min((z[2:] - z[:-2]))
        return float(0.5 * d * np.min(0.5 * (self.hi - self.lo)))

    def dt(self, p12):
        return abs(p12) if p12 < 0 else p12 * self.dxmin()


def _axis_aligned(xc, yc, zc, tol=1e-9):
    """Whether every element's 8 corners are those of an axis-aligned box
    (corners 1-4 on t = -1 counter-clockwise from (r, s) = (-1, -1), 5-8
    above them, as nkc_create_hex reads them)."""
    side = np.array([[0, 1, 1, 0, 0, 1, 1, 0], [0, 0, 1, 1, 0, 0, 1, 1], [0, 0, 0, 0, 1, 1, 1, 1]], dtype=bool)
    for a, c in enumerate((xc, yc, zc)):
        c = np.asarray(c, dtype=np.float64).reshape(-1, 8)
        lo = c[:, ~side[a]].mean(axis=1)
        hi = c[:, side[a]].mean(axis=1)
        scale = np.maximum(1.0, np.abs(hi - lo))
        if np.any(np.abs(c[:, ~side[a]] - lo[:, None]) > tol * scale[:, None]) or \
                np.any(np.abs(c[:, side[a]] - hi[:, None]) > tol * scale[:, None]):
            return False
    return True


class MaxwellSolver:
    """One nkc handle: the device-resident Maxwell state of a mesh."""

    @classmethod
    def from_corners(cls, N, xc, yc, zc, cbc, bc, eps_elem=None, mu_elem=None, ifupwind=True, graph=True):
        """Handle built by nkc_create_hex from the arrays NekCEM's readers fill:
        xc/yc/zc (nelt, 8), cbc (nelt, 6) 3-character codes, bc (nelt, 6, 5)."""
        L = _lib.load()
        xc, yc, zc = (np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 8) for a in (xc, yc, zc))
        nelt = xc.shape[0]
        codes = "".join(str(c).ljust(3)[:3] for c in np.asarray(cbc).reshape(-1)).encode("ascii")
        assert len(codes) == 18 * nelt
        bcv = np.ascontiguousarray(bc, dtype=np.float64).reshape(nelt * 6 * 5)
        ee = None if eps_elem is None else np.ascontiguousarray(eps_elem, dtype=np.float64)
        mm = None if mu_elem is None else np.ascontiguousarray(mu_elem, dtype=np.float64)
        h = C.c_void_p()
        _lib.check(L.nkc_create_hex(int(N), 1 if ifupwind else 0, nelt, _d(xc), _d(yc), _d(zc), codes, _d(bcv),
                                    _d(ee), _d(mm), C.byref(h)))
        self = cls.__new__(cls)
        lo = np.stack([xc.min(axis=1), yc.min(axis=1), zc.min(axis=1)], axis=1)
        hi = np.stack([xc.max(axis=1), yc.max(axis=1), zc.max(axis=1)], axis=1)
        self.mesh = _HexMeshInfo(N, lo, hi, cbc)
        self.npts = self.mesh.npts
        self._h, self._L = h, L
        self.time, self.istep, self.dt, self.ndrude, self.nlor = 0.0, 0, None, 0, 0
        if not graph:
            _lib.check(L.nkc_set_graph(h, 0))
        return self

    @classmethod
    def from_rea(cls, rea_path, N, usrdat2=None, re2_path=None, usrdat2_nodes=None, **kw):
        """Handle for a NekCEM case file (.rea, with its .re2 when the mesh is
        there).  usrdat2(xc, yc, zc) -> (xc, yc, zc) may move the corners as
        the case's usrdat / usrdat2 move the mesh; usrdat2_nodes(x, y, z) ->
        (x, y, z) the node coordinates.  Axis-aligned box elements go to
        nkc_create_hex; curved sides or other element shapes to the general
        hexahedra (nekcem_amd.hexmesh, nkc_create).  The .rea parameters are
        kept in ``param`` (1-based: param[11] = nsteps, param[12] = dt or
        CFL)."""
        from .rea import read_case
        r = read_case(rea_path, re2_path)
        xc, yc, zc = r["xc"], r["yc"], r["zc"]
        if usrdat2 is not None:
            xc, yc, zc = usrdat2(xc, yc, zc)
        curves = [c for c in r.get("curves", ()) if c[3] not in ("", " ")]
        if curves or usrdat2_nodes is not None or not _axis_aligned(xc, yc, zc):
            from .hexmesh import HexMesh
            mesh = HexMesh.from_corners(N, xc, yc, zc, r["cbc"], r["bc"], curves, usrdat2=usrdat2_nodes)
            self = cls(mesh, ifupwind=kw.get("ifupwind", True), graph=kw.get("graph", True))
        else:
            self = cls.from_corners(N, xc, yc, zc, r["cbc"], r["bc"], **kw)
        self.param = r["param"]
        return self

    @classmethod
    def from_box(cls, nel, N, lo=(0.0, 0.0, 0.0), hi=(2 * np.pi,) * 3, bc=("P", "P", "P"),
                 eps_elem=None, mu_elem=None, ifupwind=True, graph=True):
        """Handle built by nkc_create_box (no per-node reference arrays)."""
        L = _lib.load()
        d = box_desc(nel, N, lo, hi, bc, eps_elem, mu_elem, ifupwind)
        h = C.c_void_p()
        _lib.check(L.nkc_create_box(C.byref(d), C.byref(h)))
        self = cls.__new__(cls)
        self.mesh = _BoxMeshInfo(nel, N, lo, hi, bc)
        self.npts = self.mesh.npts
        self._h, self._L = h, L
        self.time, self.istep, self.dt, self.ndrude, self.nlor = 0.0, 0, None, 0, 0
        if not graph:
            _lib.check(L.nkc_set_graph(h, 0))
        return self

    def __init__(self, mesh: BoxMesh, eps=None, mu=None, ifupwind=True, graph=True):
        L = _lib.load()
        self.mesh = mesh
        self.npts = mesh.npts
        eps, mu, Z0, Z1, Y0, Y1 = mesh.materials(eps, mu)
        self.eps, self.mu = eps, mu
        keep = dict(eps=eps, mu=mu, Z0=Z0, Z1=Z1, Y0=Y0, Y1=Y1)
        d = _lib.NkcDesc()
        d.N = mesh.N
        d.ifupwind = 1 if ifupwind else 0
        d.nelt = mesh.nelt
        for name in ("dxm1", "wxm1", "rxmn", "rymn", "rzmn", "sxmn", "symn", "szmn",
                     "txmn", "tymn", "tzmn", "jacm", "bmn", "unxm", "unym", "unzm", "aream"):
            setattr(d, name, _d(getattr(mesh, name)))
        d.cemface = _i(mesh.cemface)
        d.face_partner = _i(mesh.face_partner)
        d.cempec = _i(mesh.cempec) if mesh.ncempec else None
        d.ncempec = mesh.ncempec
        d.permittivity, d.permeability = _d(eps), _d(mu)
        d.Z_0, d.Z_1, d.Y_0, d.Y_1 = _d(Z0), _d(Z1), _d(Y0), _d(Y1)
        h = C.c_void_p()
        _lib.check(L.nkc_create(C.byref(d), C.byref(h)))
        del keep
        self._h = h
        self._L = L
        self.time = 0.0
        self.istep = 0
        self.dt = None
        self.ndrude = 0
        self.nlor = 0
        if not graph:
            _lib.check(L.nkc_set_graph(h, 0))

    # -- state ------------------------------------------------------------
    def set_state(self, H, E, kH=None, kE=None):
        n = self.npts
        H = np.ascontiguousarray(H, dtype=np.float64).reshape(3, n)
        E = np.ascontiguousarray(E, dtype=np.float64).reshape(3, n)
        kH = None if kH is None else np.ascontiguousarray(kH, dtype=np.float64).reshape(3, n)
        kE = None if kE is None else np.ascontiguousarray(kE, dtype=np.float64).reshape(3, n)
        _lib.check(self._L.nkc_set_state(self._h, _d(H), _d(E), _d(kH), _d(kE), n))

    def get_state(self, with_k=False):
        n = self.npts
        H, E = np.empty((3, n)), np.empty((3, n))
        kH = np.empty((3, n)) if with_k else None
        kE = np.empty((3, n)) if with_k else None
        _lib.check(self._L.nkc_get_state(self._h, _d(H), _d(E), _d(kH), _d(kE), n))
        return (H, E, kH, kE) if with_k else (H, E)

    # -- operator ---------------------------------------------------------
    def residual(self, rktime=0.0):
        n = self.npts
        rH, rE = np.empty((3, n)), np.empty((3, n))
        _lib.check(self._L.nkc_residual(self._h, rktime, _d(rH), _d(rE), n))
        return rH, rE

    def curl(self):
        """(resH, resE) = (-W curl E, W curl H), W = bm1: cem_maxwell's volume
        term alone on the current state (no flux, no inverse mass)."""
        n = self.npts
        rH, rE = np.empty((3, n)), np.empty((3, n))
        _lib.check(self._L.nkc_curl(self._h, _d(rH), _d(rE), n))
        return rH, rE

    def set_keep_k(self, keep=True):
        """With keep, the RK register read back by get_state(with_k=True)
        after a step is rk4_upd's (the last stage writes it); without, it
        is the fourth stage's (the default: the write is skipped)."""
        _lib.check(self._L.nkc_set_keep_k(self._h, 1 if keep else 0))

    def step(self, dt, nsteps=1):
        t = C.c_double(self.time)
        _lib.check(self._L.nkc_step(self._h, C.byref(t), float(dt), int(nsteps)))
        self.time = t.value
        self.istep += int(nsteps)

    def set_source(self, amp, direction=(0.0, 0.0, 1.0), t0=0.0, tw=1.0, omega=None, phase=0.0):
        """usersrc-style E source: resE += amp * dir * g(t) with amp (npts)
        bm1-weighted, g(t) = exp(-((t-t0)/tw)^2), or sin(omega t + phase)
        when omega is given; amp=None removes it."""
        if amp is None:
            _lib.check(self._L.nkc_set_source(self._h, None, None, 0.0, 1.0))
            return
        amp = np.ascontiguousarray(amp, dtype=np.float64).reshape(self.npts)
        d = np.ascontiguousarray(direction, dtype=np.float64)
        self._src_keep = (amp, d)
        _lib.check(self._L.nkc_set_source(self._h, _d(amp), _d(d), float(t0), float(tw)))
        if omega is not None:
            _lib.check(self._L.nkc_set_source_time(self._h, 1, float(omega), float(phase)))

    # -- PML and incident field --------------------------------------------
    def _pml_faces(self):
        if hasattr(self.mesh, "pml_faces"):
            return self.mesh.pml_faces()
        return box_pml_faces(self.mesh.nel, self.mesh.bc)

    def setup_pml(self, thick, order, referr, pml_faces=None):
        """PML layers on the 'PML' walls as cem_maxwell_init sets them up
        (thick, order, referr = param(77..79)); B = D = 0.  Returns
        dict(inner, outer): pmlinner / pmlouter in the order x-, x+, y-, y+,
        z-, z+."""
        self._ensure_coords()
        pf = self._pml_faces() if pml_faces is None else pml_faces
        pf = np.ascontiguousarray(pf, dtype=np.int8).reshape(-1)
        ext = np.empty(12)
        _lib.check(self._L.nkc_setup_pml(self._h, pf.ctypes.data_as(C.POINTER(C.c_int8)), int(thick), float(order),
                                         float(referr), _d(ext)))
        self.pml = dict(inner=ext[:6].copy(), outer=ext[6:].copy(), thick=int(thick), order=float(order),
                        referr=float(referr))
        return self.pml

    def set_pml(self, pmlptr, sigma, B=None, D=None):
        """The reference's PML arrays: pmlptr (0-based element numbers),
        sigma (3, npts), initial B = pmlbn, D = pmldn (3, npts) or None."""
        n = self.npts
        ptr = np.ascontiguousarray(np.asarray(pmlptr) + 1, dtype=np.int32)
        sig = np.ascontiguousarray(sigma, dtype=np.float64).reshape(3, n)
        b = None if B is None else np.ascontiguousarray(B, dtype=np.float64).reshape(3, n)
        d = None if D is None else np.ascontiguousarray(D, dtype=np.float64).reshape(3, n)
        _lib.check(self._L.nkc_set_pml(self._h, len(ptr), _i(ptr), _d(sig), _d(b), _d(d), n))

    def set_pml_fields(self, B, D):
        """pmlbn, pmldn (3, npts) as userini sets them; their RK register := 0."""
        n = self.npts
        b = np.ascontiguousarray(B, dtype=np.float64).reshape(3, n)
        d = np.ascontiguousarray(D, dtype=np.float64).reshape(3, n)
        _lib.check(self._L.nkc_set_pml_fields(self._h, _d(b), _d(d), n))

    _PML_WHICH = dict(state=0, register=1, residual=2, sigma=3)

    def get_pml(self, which="state"):
        """(6, npts) B then D (state, register or last residual), or (3, npts)
        sigma; zero outside the PML elements."""
        w = self._PML_WHICH[which]
        out = np.zeros((3 if w == 3 else 6, self.npts))
        _lib.check(self._L.nkc_get_pml(self._h, w, _d(out), self.npts))
        return out

    def set_incident(self, faces, k, omega, hamp, eamp):
        """Incident plane wave u = cos(k . x - omega t) on the listed faces
        ((element 0-based, face 1..6 in the reference's order) pairs): the
        traces get H += hamp u, E += eamp u before the flux.  faces = [] removes it."""
        self._ensure_coords()
        f = np.ascontiguousarray(np.asarray(faces, dtype=np.int64).reshape(-1, 2) + [1, 0], dtype=np.int32)
        kk, hh, ee = (np.ascontiguousarray(v, dtype=np.float64) for v in (k, hamp, eamp))
        _lib.check(self._L.nkc_set_incident(self._h, len(f), _i(f) if len(f) else None, _d(kk), float(omega),
                                            _d(hh), _d(ee)))

    def set_graphene(self, gindex, params, G=None):
        """Graphene sheet current (userfsrc + cem_3d_graphene_current) on
        the face points gindex (0-based, NekCEM's face-point numbering
        e 6 n2 + (f-1) n2 + i, f in the reference's face order), with the
        twelve sheet parameters per point (params (n, 12) or (12,)) and the
        initial state G (15, n): fjn slots 2..6 x 3 components, row
        3 (slot - 2) + component (None: zero).  gindex = [] removes it."""
        gindex = np.asarray(gindex, dtype=np.int64)
        n = len(gindex)
        if n == 0:
            _lib.check(self._L.nkc_set_graphene(self._h, 0, None, None, None, 0))
            self._gindex = None
            return
        ld = self.mesh.nelt * 6 * (self.mesh.N + 1) ** 2
        par = np.zeros((12, ld))
        par[:, gindex] = np.broadcast_to(np.asarray(params, dtype=np.float64), (n, 12)).T
        fjn = np.zeros((6, 3, ld))
        if G is not None:
            fjn[1:, :, gindex] = np.asarray(G, dtype=np.float64).reshape(5, 3, n)
        gi = np.ascontiguousarray(gindex + 1, dtype=np.int32)
        _lib.check(self._L.nkc_set_graphene(self._h, n, _i(gi), _d(par), _d(fjn), ld))
        self._gindex = gindex

    def get_graphene(self, which="state"):
        """The sheet state (or its RK register, which="register") as
        (15, n) in the order of set_graphene's gindex."""
        w = {"state": 0, "register": 1}[which]
        ld = self.mesh.nelt * 6 * (self.mesh.N + 1) ** 2
        fjn = np.zeros((6, 3, ld))
        _lib.check(self._L.nkc_get_graphene(self._h, w, _d(fjn), ld))
        return fjn[1:, :, self._gindex].reshape(15, -1)

    def set_drude(self, dindex, a, b, J=None, kJ=None):
        """Drude ADE on volume nodes dindex (0-based): a = gamma, b =
        omega_p^2 per listed node; J, kJ (3, ndrude) or None (zero)."""
        dindex = np.asarray(dindex, dtype=np.int64)
        nd = len(dindex)
        if nd == 0:
            _lib.check(self._L.nkc_set_drude(self._h, 0, None, None, None, None))
            self.ndrude = 0
            return
        di = np.ascontiguousarray(dindex + 1, dtype=np.int32)
        par = np.ascontiguousarray(np.stack([np.broadcast_to(a, (nd,)), np.broadcast_to(b, (nd,))], axis=1),
                                   dtype=np.float64)
        J = None if J is None else np.ascontiguousarray(J, dtype=np.float64).reshape(3, nd)
        kJ = None if kJ is None else np.ascontiguousarray(kJ, dtype=np.float64).reshape(3, nd)
        _lib.check(self._L.nkc_set_drude(self._h, nd, _i(di), _d(par), _d(J), _d(kJ)))
        self.ndrude = nd

    def get_drude(self):
        nd = self.ndrude
        J, kJ, rJ = np.empty((3, nd)), np.empty((3, nd)), np.empty((3, nd))
        _lib.check(self._L.nkc_get_drude(self._h, _d(J), _d(kJ), _d(rJ)))
        return J, kJ, rJ

    def set_lorentz(self, lindex, a, b, c, L=None, kL=None):
        """Lorentz ADE on volume nodes lindex (0-based) with per-node a, b, c
        (eps(w) = eps_oo + c/(b - i a w - w^2)); L, kL (6, nlor) = (J, P)
        or None (zero)."""
        lindex = np.asarray(lindex, dtype=np.int64)
        n = len(lindex)
        if n == 0:
            _lib.check(self._L.nkc_set_lorentz(self._h, 0, None, None, None, None))
            self.nlor = 0
            return
        li = np.ascontiguousarray(lindex + 1, dtype=np.int32)
        par = np.ascontiguousarray(np.stack([np.broadcast_to(a, (n,)), np.broadcast_to(b, (n,)),
                                             np.broadcast_to(c, (n,))], axis=1), dtype=np.float64)
        L = None if L is None else np.ascontiguousarray(L, dtype=np.float64).reshape(6, n)
        kL = None if kL is None else np.ascontiguousarray(kL, dtype=np.float64).reshape(6, n)
        _lib.check(self._L.nkc_set_lorentz(self._h, n, _i(li), _d(par), _d(L), _d(kL)))
        self.nlor = n

    def get_lorentz(self):
        """(L, kL, resL), each (6, nlor): J then P."""
        n = self.nlor
        L, kL, rL = np.empty((6, n)), np.empty((6, n)), np.empty((6, n))
        _lib.check(self._L.nkc_get_lorentz(self._h, _d(L), _d(kL), _d(rL)))
        return L, kL, rL

    def set_materials(self, eps, mu=None):
        """Per-node permittivity/permeability (npts each)."""
        eps = np.ascontiguousarray(np.broadcast_to(eps, (self.npts,)), dtype=np.float64)
        mu = np.ones(self.npts) if mu is None else np.ascontiguousarray(np.broadcast_to(mu, (self.npts,)),
                                                                        dtype=np.float64)
        _lib.check(self._L.nkc_set_materials(self._h, _d(eps), _d(mu)))

    def adjoint_forward(self, dt, nsteps, every):
        """nsteps RK steps keeping a checkpoint every `every` steps."""
        t = C.c_double(self.time)
        _lib.check(self._L.nkc_adjoint_forward(self._h, C.byref(t), float(dt), int(nsteps), int(every)))
        self.time = t.value
        self.istep += int(nsteps)

    def adjoint_backward(self, lamH, lamE):
        """dL/deps, dL/dmu per node for a loss with dL/dH = lamH, dL/dE =
        lamE at the end of the last adjoint_forward."""
        n = self.npts
        lamH = np.ascontiguousarray(lamH, dtype=np.float64).reshape(3, n)
        lamE = np.ascontiguousarray(lamE, dtype=np.float64).reshape(3, n)
        ge, gm = np.empty(n), np.empty(n)
        _lib.check(self._L.nkc_adjoint_backward(self._h, _d(lamH), _d(lamE), n, _d(ge), _d(gm)))
        return ge, gm

    # -- on-device diagnostics ------------------------------------------------
    SOLUTIONS = {"3dboxper": 1, "3dboxpec": 2}

    def _ensure_coords(self):
        if getattr(self, "_coords_set", False) or isinstance(self.mesh, _BoxMeshInfo):
            return
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in self.mesh.coords())
        _lib.check(self._L.nkc_set_coords(self._h, _d(x), _d(y), _d(z)))
        self._coords_set = True

    def init_solution(self, sol, t=0.0):
        """userini on the device: the state := analytic solution `sol`
        ('3dboxper' or '3dboxpec') at time t, k := 0."""
        self._ensure_coords()
        _lib.check(self._L.nkc_init_solution(self._h, self.SOLUTIONS[sol], float(t)))
        self.time = float(t)

    def cem_error_device(self, sol, t=None):
        """cem_error of the state against analytic solution `sol` at time t
        (default: the solver's time), on the device: (L2[6], Linf[6]) in
        the order Hx, Hy, Hz, Ex, Ey, Ez."""
        self._ensure_coords()
        out = np.empty(12)
        _lib.check(self._L.nkc_cem_error(self._h, self.SOLUTIONS[sol], float(self.time if t is None else t),
                                         _d(out)))
        return out[:6], out[6:]

    def cem_error_fields(self, H, E):
        """cem_error of the state against given exact fields, on the device."""
        n = self.npts
        H = np.ascontiguousarray(H, dtype=np.float64).reshape(3, n)
        E = np.ascontiguousarray(E, dtype=np.float64).reshape(3, n)
        out = np.empty(12)
        _lib.check(self._L.nkc_cem_error_fields(self._h, _d(H), _d(E), n, _d(out)))
        return out[:6], out[6:]

    def restart_out(self, path, dumpno=0, single=False):
        """cem_restart_out: write the state to a restart file at path, with
        the solver's istep, time and dt in its header."""
        self._ensure_coords()
        _lib.check(self._L.nkc_restart_write(self._h, str(path).encode(), int(self.istep), int(dumpno),
                                             float(self.time), float(self.dt or 0.0), 1 if single else 0))

    def restart_swap(self, path):
        """restart_swap: the state from a restart file (k := 0); the header's
        time is added to the solver's time as the reference does.  Returns
        dict(istep, dumpno, time, dt) of the header."""
        istep, dumpno = C.c_int32(), C.c_int32()
        t, dt = C.c_double(), C.c_double()
        _lib.check(self._L.nkc_restart_read(self._h, str(path).encode(), C.byref(istep), C.byref(dumpno),
                                            C.byref(t), C.byref(dt)))
        self.time = self.time + t.value
        return dict(istep=istep.value, dumpno=dumpno.value, time=t.value, dt=dt.value)

    def set_layout(self, layout):
        """Force the stage-kernel layout (0..NLAYOUT-1), or -1 to re-tune."""
        _lib.check(self._L.nkc_set_layout(self._h, int(layout)))

    def layout(self):
        """dict(us=(tuner times per layout, NaN: not timed), layout) without
        tuning."""
        out = np.empty(NLAYOUT + 1)
        _lib.check(self._L.nkc_get_layout(self._h, _d(out)))
        return dict(us=tuple(float(v) for v in out[:NLAYOUT]), layout=int(out[NLAYOUT]))

    def tune(self):
        """Run the layout autotuner now: dict(us=(per layout), layout)."""
        out = np.empty(NLAYOUT + 1)
        _lib.check(self._L.nkc_tune(self._h, _d(out)))
        return dict(us=tuple(float(v) for v in out[:NLAYOUT]), layout=int(out[NLAYOUT]))

    def adjoint_set_store(self, max_steps):
        """Cap the stage-store window of the adjoint (0: from free memory)."""
        _lib.check(self._L.nkc_adjoint_set_store(self._h, int(max_steps)))

    def adjoint_release(self):
        """Free the adjoint's checkpoints, stage store and gradients."""
        _lib.check(self._L.nkc_adjoint_release(self._h))

    def adjoint_set_grads(self, material=True):
        """Whether the backward sweep computes dL/deps, dL/dmu (default) or
        skips them (Drude gradients only; half the stage store)."""
        _lib.check(self._L.nkc_adjoint_set_grads(self._h, 1 if material else 0))

    def adjoint_info(self):
        """dict(every, segments, window, kept, bytes) of the last adjoint_forward
        (kept: segments whose stages the forward kept)."""
        out = (C.c_int64 * 5)()
        _lib.check(self._L.nkc_adjoint_info(self._h, out))
        return dict(every=out[0], segments=out[1], window=out[2], kept=out[3], bytes=out[4])

    def adjoint_drude_grad(self):
        """(dL/da, dL/db) per listed Drude node (a = gamma, b = omega_p^2),
        in the order given to set_drude, from the last adjoint_backward."""
        nd = self.ndrude
        ga, gb = np.empty(nd), np.empty(nd)
        _lib.check(self._L.nkc_adjoint_get_drude(self._h, _d(ga), _d(gb)))
        return ga, gb

    def adjoint_apply(self, kH, kE):
        """B^T P (kH, kE): transpose of the (source-free) residual operator."""
        n = self.npts
        kH = np.ascontiguousarray(kH, dtype=np.float64).reshape(3, n)
        kE = np.ascontiguousarray(kE, dtype=np.float64).reshape(3, n)
        aH, aE = np.empty((3, n)), np.empty((3, n))
        _lib.check(self._L.nkc_adjoint_apply(self._h, _d(kH), _d(kE), _d(aH), _d(aE), n))
        return aH, aE

    def sync(self):
        _lib.check(self._L.nkc_sync(self._h))

    def check_finite(self):
        m = C.c_double()
        _lib.check(self._L.nkc_check_finite(self._h, C.byref(m)))
        return m.value

    def stream(self):
        s = C.c_void_p()
        _lib.check(self._L.nkc_stream(self._h, C.byref(s)))
        return s.value

    def info(self):
        out = (C.c_int64 * 8)()
        _lib.check(self._L.nkc_info(self._h, out))
        keys = ("npts", "mat_mode", "elems_per_wg", "threads_per_wg", "lds_bytes", "graph", "fs", "nelt")
        return dict(zip(keys, list(out)))

    def close(self):
        if self._h:
            _lib.check(self._L.nkc_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- reference-named driver ------------------------------------------
    def cem_maxwell_init_fields(self, userini):
        """userini(t, xm1, ym1, zm1) -> (H, E); sets the state at self.time."""
        x, y, z = self.mesh.coords()
        H, E = userini(self.time, x, y, z)
        self.set_state(H, E)

    def cem_maxwell_op_rk(self):
        """One 5-stage low-storage RK step of size self.dt."""
        if self.dt is None:
            raise RuntimeError("dt not set (use set_dt)")
        self.step(self.dt, 1)

    def set_dt(self, p12):
        """set_dt semantics: p12 < 0 is dt itself, p12 > 0 a CFL number."""
        self.dt = self.mesh.dt(p12)
        return self.dt

    def time_advancing_pde(self, nsteps, userchk=None, chk_every=None):
        """Run nsteps steps; call userchk(self) after each step for which
        istep <= 10 or istep % chk_every == 0 (the reference's iocomm rule),
        or only at the end when chk_every is None.  userchk returning False
        stops the run with RunStopped (exitt(1))."""
        done = 0
        while done < nsteps:
            if userchk is None:
                self.step(self.dt, nsteps - done)
                return
            if chk_every is None:
                self.step(self.dt, nsteps - done)
                done = nsteps
            else:
                # advance to the next check point
                nxt = self.istep + 1
                if nxt > 10:
                    nxt = ((self.istep // chk_every) + 1) * chk_every
                n = min(nxt - self.istep, nsteps - done)
                self.step(self.dt, n)
                done += n
            if userchk(self) is False:
                raise RunStopped(f"userchk stopped the run at istep={self.istep}")


class BatchSolver:
    """Many independent box cavities advanced together (nkc_batch_*).

    cavities: list of dicts with keys nel, N, lo, hi, bc, eps_elem, mu_elem.
    """

    def __init__(self, cavities, ifupwind=True):
        L = _lib.load()
        self._L = L
        n = len(cavities)
        arr = (_lib.NkcBoxDesc * n)()
        self._keep = []
        self.info_cav = []
        for c, cv in enumerate(cavities):
            d = box_desc(cv["nel"], cv["N"], cv.get("lo", (0, 0, 0)), cv.get("hi", (1, 1, 1)),
                         cv.get("bc", ("PEC",) * 3), cv.get("eps_elem"), cv.get("mu_elem"), ifupwind)
            self._keep.append(d._keep)
            arr[c] = d
            self.info_cav.append(_BoxMeshInfo(cv["nel"], cv["N"], cv.get("lo", (0, 0, 0)),
                                              cv.get("hi", (1, 1, 1)), cv.get("bc", ("PEC",) * 3)))
        b = C.c_void_p()
        _lib.check(L.nkc_batch_create(n, arr, C.byref(b)))
        self._b = b
        self.ncav = n
        self.times = np.zeros(n)

    def set_state(self, cav, H, E):
        n = self.info_cav[cav].npts
        H = np.ascontiguousarray(H, dtype=np.float64).reshape(3, n)
        E = np.ascontiguousarray(E, dtype=np.float64).reshape(3, n)
        _lib.check(self._L.nkc_batch_set_state(self._b, int(cav), _d(H), _d(E), n))

    def get_state(self, cav):
        n = self.info_cav[cav].npts
        H, E = np.empty((3, n)), np.empty((3, n))
        _lib.check(self._L.nkc_batch_get_state(self._b, int(cav), _d(H), _d(E), n))
        return H, E

    def step(self, dt, nsteps=1):
        dt = np.ascontiguousarray(np.broadcast_to(dt, (self.ncav,)), dtype=np.float64)
        t = np.empty(self.ncav)
        _lib.check(self._L.nkc_batch_step(self._b, _d(dt), _d(t), int(nsteps)))
        self.times = t

    def sync(self):
        _lib.check(self._L.nkc_batch_sync(self._b))

    def stream(self):
        s = C.c_void_p()
        _lib.check(self._L.nkc_batch_stream(self._b, C.byref(s)))
        return s.value

    def info(self):
        out = (C.c_int64 * 8)()
        _lib.check(self._L.nkc_batch_info(self._b, out))
        return dict(npts=out[0], ncav=out[1], norders=out[2], nelt=out[3])

    def layout(self):
        """{N: stage-kernel layout} per polynomial order (-1 before tuning)."""
        nb = self.info()["norders"]
        out = (C.c_int32 * (2 * nb))()
        _lib.check(self._L.nkc_batch_layout(self._b, out, 2 * nb))
        return {int(out[2 * i]): int(out[2 * i + 1]) for i in range(nb)}

    def close(self):
        if self._b:
            _lib.check(self._L.nkc_batch_destroy(self._b))
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cem_maxwell_init(mesh: BoxMesh, eps=None, mu=None, ifupwind=True, graph=True):
    return MaxwellSolver(mesh, eps=eps, mu=mu, ifupwind=ifupwind, graph=graph)


def cem_error(u, exact, bm1):
    """(L2, Linf) of exact - u, L2 = sqrt(sum(bm1 err^2) / sum(bm1))."""
    err = np.asarray(exact, dtype=np.float64) - np.asarray(u, dtype=np.float64)
    l2 = float(np.sum(err * bm1 * err) / np.sum(bm1))
    l2 = np.sqrt(l2) if l2 > 0 else l2
    return float(l2), float(np.max(np.abs(err)))
