"""CPU oracle for NekCEM's 3-D DG-SEM Maxwell RK step -- TEST INFRASTRUCTURE.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this module, and only as the checker / the timed CPU baseline.
The product package ``nekcem_amd`` never imports it.

The arithmetic lives in ``nekcem_oracle.c`` (built by ``oracle/Makefile`` into
``oracle/_build/libnekcem_oracle.so``).  This module assembles a box case the
way the reference does for its ``.box``/``.re2`` cases:

* element numbering ``e = ex + nelx*(ey + nely*ez)`` (get_exyz in
  src/nek5_genbox.F:83-145);
* straight-sided elements, so ``xm1`` is the tensor-product linear map of the
  element's vertices (src/nek5_genxyz.F, xyzlin);
* face-node gs ids from the "extended lattice" of src/nek5_connect11.F:
  2179-2347 (setup_dgds2 / iface_vert_int8): every element owns an
  (n1+2)^3 extended grid whose ghost layers coincide with the neighbour's,
  and face node (a, b) of face f takes the id of the ghost-layer point just
  outside the face; periodic directions wrap (gen_gtp_vertex,
  src/nek5_genbox.F:111-117).  Entries sharing an id are the gs partners.
* CBC letters per face: 'P' (periodic pair) or 'PEC' on the box walls.

Parity status of each C function is recorded in tests/test_oracle.py, which
pins it against closed forms, invariants and the reference's own analytic
test cases (tests/3dboxper, tests/3dboxpec, tests/3ddielectric,
tests/3dgraphene) at the reference's tolerances.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libnekcem_oracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_bp = C.POINTER(C.c_int8)


def build():
    """Compile the oracle (gcc, OpenMP) into oracle/_build/."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_zwgll.argtypes = [C.c_int, _dp, _dp]
        L.orc_dgll.argtypes = [C.c_int, _dp, _dp]
        L.orc_pnleg.argtypes = [C.c_double, C.c_int]
        L.orc_pnleg.restype = C.c_double
        L.orc_geom.argtypes = [C.c_int, C.c_int64, _dp, _dp, _dp, _dp, _dp] + [_dp] * 16
        L.orc_set_fc_ptr.argtypes = [C.c_int, C.c_int64, _ip]
        L.orc_get_dxmin.argtypes = [C.c_int, C.c_int64, _dp, _dp, _dp]
        L.orc_get_dxmin.restype = C.c_double
        L.orc_gs_add.argtypes = [_ip, _dp, C.c_int64, C.c_int, C.c_int64]
        L.orc_materials.argtypes = [C.c_int, C.c_int64, _ip, _ip, _bp, _dp, _dp, _dp,
                                    _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_pec_init.argtypes = [C.c_int, C.c_int64, _bp, _ip]
        L.orc_pec_init.restype = C.c_int64
        L.orc_maxwell_op.argtypes = [C.c_void_p, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_rk_run.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                 C.c_double, C.c_int64, _dp]
        L.orc_maxwell_op_ade.argtypes = [C.c_void_p, C.c_double] + [_dp] * 8
        L.orc_rk_run_ade.argtypes = [C.c_void_p] + [_dp] * 12 + [C.c_double, C.c_int64, _dp]
        L.orc_maxwell_op_full.argtypes = [C.c_void_p, C.c_double] + [_dp] * 10
        L.orc_curl.argtypes = [C.c_void_p] + [_dp] * 4
        L.orc_rk_run_full.argtypes = [C.c_void_p] + [_dp] * 15 + [C.c_double, C.c_int64, _dp]
        L.orc_maxwell_op_all.argtypes = [C.c_void_p, C.c_double] + [_dp] * 12
        L.orc_rk_run_all.argtypes = [C.c_void_p] + [_dp] * 18 + [C.c_double, C.c_int64, _dp]
        L.orc_rk_coeffs.argtypes = [_dp, _dp, _dp]
        L.orc_cem_error.argtypes = [_dp, _dp, _dp, C.c_int64, _dp, _dp]
        L.orc_maxwell_vjp.argtypes = [C.c_void_p, C.c_double, _dp, _dp, _bp, _dp, _dp, _dp, _dp,
                                      _dp, _dp, _dp, _dp]
        L.orc_drude_vjp.argtypes = [C.c_void_p] + [_dp] * 10
        L.orc_num_threads.restype = C.c_int
        L.orc_set_num_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    if a.dtype == np.int8:
        return a.ctypes.data_as(_bp)
    raise TypeError(a.dtype)


class _Mesh(C.Structure):
    _fields_ = [
        ("n1", C.c_int32), ("nelt", C.c_int32), ("npts", C.c_int64), ("nxzfl", C.c_int64),
        ("D", _dp), ("w3", _dp),
        ("rx", _dp), ("ry", _dp), ("rz", _dp), ("sx", _dp), ("sy", _dp), ("sz", _dp),
        ("tx", _dp), ("ty", _dp), ("tz", _dp),
        ("bm1", _dp),
        ("unx", _dp), ("uny", _dp), ("unz", _dp), ("area", _dp),
        ("cemface", _ip), ("partner", _ip), ("cempec", _ip), ("ncempec", C.c_int64),
        ("Y0", _dp), ("Y1", _dp), ("Z0", _dp), ("Z1", _dp),
        ("hbm1", _dp), ("ebm1", _dp), ("C0", C.c_double),
        ("ndrude", C.c_int64), ("dindex", _ip), ("drude_a", _dp), ("drude_b", _dp),
        ("src_amp", _dp), ("src_dir", C.c_double * 3), ("src_t0", C.c_double),
        ("src_tw", C.c_double),
        ("srflx", _dp),
        ("nlorentz", C.c_int64), ("lindex", _ip), ("lor_par", _dp),
        ("src_kind", C.c_int32), ("src_omega", C.c_double), ("src_phase", C.c_double),
        ("npml", C.c_int64), ("pmlptr", _ip), ("pmlsigma", _dp), ("eps", _dp), ("mu", _dp),
        ("inc_on", _bp), ("xm1", _dp), ("ym1", _dp), ("zm1", _dp),
        ("inc_k", C.c_double * 3), ("inc_omega", C.c_double), ("inc_h", C.c_double * 3),
        ("inc_e", C.c_double * 3),
        ("ngraph", C.c_int64), ("gindex", _ip), ("gpar", _dp),
    ]


def zwgll(n1):
    z = np.zeros(n1)
    w = np.zeros(n1)
    lib().orc_zwgll(n1, _p(z), _p(w))
    return z, w


def dgll(z):
    n1 = len(z)
    z = np.ascontiguousarray(z, dtype=np.float64)
    D = np.zeros(n1 * n1)
    lib().orc_dgll(n1, _p(z), _p(D))
    return D.reshape(n1, n1, order="F")


def rk_coeffs():
    a, b, c = np.zeros(5), np.zeros(5), np.zeros(6)
    lib().orc_rk_coeffs(_p(a), _p(b), _p(c))
    return a, b, c


# Ed's face order -> (axis, side): 1 y-, 2 x+, 3 y+, 4 x-, 5 z-, 6 z+
_ED_FACES = [(1, 0), (0, 1), (1, 1), (0, 0), (2, 0), (2, 1)]


@dataclass
class BoxSpec:
    nel: tuple          # (nelx, nely, nelz)
    N: int              # polynomial order
    lo: tuple = (0.0, 0.0, 0.0)
    hi: tuple = (2 * np.pi, 2 * np.pi, 2 * np.pi)
    bc: tuple = ("P", "P", "P")   # per direction: 'P' periodic, 'PEC' or 'PML' walls
    upwind: bool = True
    verts: tuple = None  # optional per-axis vertex coordinates (nel+1 each): a graded box


class BoxCase:
    """A reference-layout box case: every array the reference's
    cem_maxwell_init leaves in COMMON, for a box mesh."""

    def __init__(self, spec: BoxSpec, eps=None, mu=None):
        L = lib()
        self.spec = spec
        nelx, nely, nelz = spec.nel
        n1 = spec.N + 1
        self.n1 = n1
        self.nelt = nelt = nelx * nely * nelz
        n2, n3 = n1 * n1, n1 ** 3
        self.npts = npts = n3 * nelt
        self.nxzfl = nxzfl = 6 * n2 * nelt
        self.z, self.w = zwgll(n1)
        self.D = dgll(self.z)
        self._Dflat = np.ascontiguousarray(self.D.reshape(-1, order="F"))

        # coordinates: element (ex,ey,ez) spans vertices lo + (hi-lo)*v/nel
        e = np.arange(nelt)
        ex, ey, ez = e % nelx, (e // nelx) % nely, e // (nelx * nely)
        self.exyz = (ex, ey, ez)
        coords = []
        for ax, (ea, nea) in enumerate(((ex, nelx), (ey, nely), (ez, nelz))):
            if spec.verts is not None:
                vv = np.asarray(spec.verts[ax], dtype=np.float64)
                v0, v1 = vv[ea], vv[ea + 1]
            else:
                lo, hi = spec.lo[ax], spec.hi[ax]
                v0 = lo + (hi - lo) * ea / nea
                v1 = lo + (hi - lo) * (ea + 1) / nea
            h0 = 0.5 * (1.0 - self.z)
            h1 = 0.5 * (1.0 + self.z)
            line = v0[:, None] * h0[None, :] + v1[:, None] * h1[None, :]   # (nelt, n1)
            shape = [nelt, 1, 1, 1]
            shape[3 - ax] = n1
            arr = np.broadcast_to(line.reshape(shape), (nelt, n1, n1, n1))
            coords.append(np.ascontiguousarray(arr).reshape(-1))
        self.xm1, self.ym1, self.zm1 = coords

        g = {k: np.zeros(npts) for k in ("rx", "ry", "rz", "sx", "sy", "sz",
                                          "tx", "ty", "tz", "jac", "bm1")}
        self.w3 = np.zeros(n3)
        fa = {k: np.zeros(nxzfl) for k in ("unx", "uny", "unz", "area")}
        L.orc_geom(n1, nelt, _p(self._Dflat), _p(self.w), _p(self.xm1), _p(self.ym1),
                   _p(self.zm1), *[_p(g[k]) for k in ("rx", "ry", "rz", "sx", "sy", "sz",
                                                       "tx", "ty", "tz", "jac", "bm1")],
                   _p(self.w3), _p(fa["unx"]), _p(fa["uny"]), _p(fa["unz"]), _p(fa["area"]))
        self.geom = g
        self.face = fa
        self.bm1 = g["bm1"]

        self.cemface = np.zeros(nxzfl, dtype=np.int32)
        L.orc_set_fc_ptr(n1, nelt, _p(self.cemface))

        self.gsid = self._face_ids()
        self.partner = self._partner_from_ids(self.gsid)

        self.pec_face = np.zeros(6 * nelt, dtype=np.int8)
        for f, (ax, side) in enumerate(_ED_FACES):
            if spec.bc[ax] in ("PEC", "PML"):   # cem_maxwell_pec_init lists both
                ea, nea = self.exyz[ax], spec.nel[ax]
                on = (ea == 0) if side == 0 else (ea == nea - 1)
                self.pec_face[6 * e[on] + f] = 1
        self.cempec = np.zeros(nxzfl, dtype=np.int32)
        self.ncempec = int(L.orc_pec_init(n1, nelt, _p(self.pec_face), _p(self.cempec)))

        self.set_materials(eps, mu)
        self.dxmin = float(L.orc_get_dxmin(n1, nelt, _p(self.xm1), _p(self.ym1), _p(self.zm1)))
        self.srflx = np.zeros(6 * nxzfl)
        self.src_amp = None
        self.src_dir = (0.0, 0.0, 1.0)
        self.src_t0 = 0.0
        self.src_tw = 1.0
        self.dindex = None
        self.drude_a = None
        self.drude_b = None
        self.lindex = None
        self.lor_par = None
        self.src_kind, self.src_omega, self.src_phase = 0, 0.0, 0.0
        self.pmlptr = None
        self.pmlsigma = None
        self.pmltag = np.zeros(nelt, dtype=np.int64)
        self.inc_on = None
        self.gindex = None
        self.inc = None

    # -- setup pieces ---------------------------------------------------
    def _face_ids(self):
        """gs ids of every face node (Ed order, in-face a fast, b slow)."""
        spec, n1 = self.spec, self.n1
        nelx, nely, nelz = spec.nel
        nelt = self.nelt
        ex, ey, ez = self.exyz
        Lx, Ly, Lz = nelx * (n1 + 1), nely * (n1 + 1), nelz * (n1 + 1)
        a = np.arange(n1)
        A, B = np.meshgrid(a, a, indexing="xy")        # A[b, a] = a, B[b, a] = b
        A = A.reshape(-1)
        B = B.reshape(-1)
        ids = np.zeros((nelt, 6, n1 * n1), dtype=np.int64)
        for f in range(6):
            if f == 0:   ix, iy, iz = A + 1, np.zeros_like(A), B + 1
            elif f == 1: ix, iy, iz = np.full_like(A, n1 + 1), A + 1, B + 1
            elif f == 2: ix, iy, iz = A + 1, np.full_like(A, n1 + 1), B + 1
            elif f == 3: ix, iy, iz = np.zeros_like(A), A + 1, B + 1
            elif f == 4: ix, iy, iz = A + 1, B + 1, np.zeros_like(A)
            else:        ix, iy, iz = A + 1, B + 1, np.full_like(A, n1 + 1)
            GX = ex[:, None] * (n1 + 1) + ix[None, :]
            GY = ey[:, None] * (n1 + 1) + iy[None, :]
            GZ = ez[:, None] * (n1 + 1) + iz[None, :]
            if spec.bc[0] == "P": GX = GX % Lx
            if spec.bc[1] == "P": GY = GY % Ly
            if spec.bc[2] == "P": GZ = GZ % Lz
            ids[:, f, :] = GX + (Lx + 1) * (GY + (Ly + 1) * GZ)
        return ids.reshape(-1)

    @staticmethod
    def _partner_from_ids(ids):
        order = np.argsort(ids, kind="stable")
        s = ids[order]
        partner = np.full(len(ids), -1, dtype=np.int32)
        same = np.nonzero(s[1:] == s[:-1])[0]
        if len(same):
            # at most two entries per id on a conforming hex mesh
            assert not np.any(np.diff(same) == 1), "gs id shared by more than two face nodes"
            i, j = order[same], order[same + 1]
            partner[i] = j
            partner[j] = i
        return partner

    def set_materials(self, eps=None, mu=None):
        npts, nxzfl = self.npts, self.nxzfl
        self.eps = np.ones(npts) if eps is None else np.ascontiguousarray(eps, dtype=np.float64)
        self.mu = np.ones(npts) if mu is None else np.ascontiguousarray(mu, dtype=np.float64)
        assert self.eps.shape == (npts,) and self.mu.shape == (npts,)
        self.Z0, self.Z1, self.Y0, self.Y1 = (np.zeros(nxzfl) for _ in range(4))
        self.ebm1, self.hbm1 = np.zeros(npts), np.zeros(npts)
        lib().orc_materials(self.n1, self.nelt, _p(self.cemface), _p(self.partner),
                            _p(self.pec_face), _p(self.eps), _p(self.mu), _p(self.bm1),
                            _p(self.Z0), _p(self.Z1), _p(self.Y0), _p(self.Y1),
                            _p(self.ebm1), _p(self.hbm1))

    def set_source(self, amp, direction=(0.0, 0.0, 1.0), t0=0.0, tw=1.0, omega=None, phase=0.0):
        """usersrc-style source: resE += amp(x) * dir * g(t), amp already
        bm1-weighted (as in src/cem_maxwell.F:3130-3132); g(t) =
        exp(-((t-t0)/tw)^2), or sin(omega t + phase) when omega is given."""
        self.src_amp = None if amp is None else np.ascontiguousarray(amp, dtype=np.float64)
        self.src_dir = tuple(float(x) for x in direction)
        self.src_t0, self.src_tw = float(t0), float(tw)
        self.src_kind = 0 if omega is None else 1
        self.src_omega, self.src_phase = (0.0 if omega is None else float(omega)), float(phase)

    # -- PML set-up (src/cem_maxwell_pml.F), restated on the face arrays ---
    def _pml_faceary(self, thick):
        """pml_fill_faceary + thick x march_faces (src/cem_maxwell_pml.F:85-190):
        1 on the interior points of every PML face, then per element and axis
        both opposite faces += 1 where their sum is nonzero, and the gs max
        over face partners."""
        n1, nelt = self.n1, self.nelt
        n2 = n1 * n1
        fa = np.zeros((nelt, 6, n2))
        a = np.arange(n1)
        A, B = np.meshgrid(a, a, indexing="xy")
        interior = ((A >= 1) & (A <= n1 - 2) & (B >= 1) & (B <= n1 - 2)).reshape(-1)
        e = np.arange(nelt)
        for f, (ax, side) in enumerate(_ED_FACES):
            if self.spec.bc[ax] != "PML":
                continue
            ea, nea = self.exyz[ax], self.spec.nel[ax]
            on = (ea == 0) if side == 0 else (ea == nea - 1)
            fa[np.ix_(e[on], [f], np.nonzero(interior)[0])] = 1.0
        eface = [4, 2, 1, 3, 5, 6]   # pff/sym face -> Ed face (1-based)
        flat = fa.reshape(-1)
        has = self.partner >= 0
        for _ in range(thick):
            fa = flat.reshape(nelt, 6, n2)
            for axis in range(3):
                pos, neg = eface[1 + 2 * axis] - 1, eface[2 * axis] - 1
                ssum = fa[:, pos, :] + fa[:, neg, :]
                nz = ssum != 0
                fa[:, pos, :][nz] += 1
                fa[:, neg, :][nz] += 1
            flat = fa.reshape(-1)
            other = flat.copy()
            other[has] = flat[self.partner[has]]
            flat = np.maximum(flat, other)
        return flat.reshape(nelt, 6, n2)

    def set_pml(self, thick, order, referr):
        """PML of `thick` layers on the 'PML' walls: pml_extent_and_tags and
        pml_calc_sigma (src/cem_maxwell_pml.F:261-505) for axis-aligned
        elements (dir_local_to_global is the identity).  Sets pmltag,
        pmlinner/pmlouter (pff order x-, x+, y-, y+, z-, z+), pmlptr and
        pmlsigma (npts, 3)."""
        n1, nelt, npts = self.n1, self.nelt, self.npts
        fa = self._pml_faceary(int(thick))
        ind = 1 + 1 * n1                       # ind_ptidx = 1 + 1*lx1 + 1 (1-based)
        eface = [4, 2, 1, 3, 5, 6]
        inf = 1e20
        inner = np.array([-inf, inf] * 3)
        outer = np.array([inf, -inf] * 3)
        tag = np.zeros(nelt, dtype=np.int64)
        X = (self.xm1, self.ym1, self.zm1)
        n3 = n1 ** 3
        for e in range(nelt):
            for axis in range(3):
                face = 2 * axis                 # pff negative face (0-based)
                here = fa[e, eface[face] - 1, ind]
                opp = fa[e, eface[face + 1] - 1, ind]
                if here == 0 or opp == 0:
                    continue
                if here == opp:
                    raise ValueError("no gradient in the PML indicators: PMLs collide")
                c = X[axis][e * n3:(e + 1) * n3]
                lo, hi = c.min(), c.max()
                if -(here - opp) > 0:           # globsign = -1 for the negative face
                    inner[face + 1] = min(inner[face + 1], lo)
                    outer[face + 1] = max(outer[face + 1], hi)
                    tag[e] |= 1 << (face + 1)
                else:
                    inner[face] = max(inner[face], hi)
                    outer[face] = min(outer[face], lo)
                    tag[e] |= 1 << face
        sigma = np.zeros((3, npts))
        ptr = np.nonzero(tag)[0]
        for e in ptr:
            sl = slice(e * n3, (e + 1) * n3)
            for face in range(6):
                if not (tag[e] >> face) & 1:
                    continue
                axis = face // 2
                width = abs(outer[face] - inner[face])
                eta = np.sqrt(self.mu[sl] / self.eps[sl])
                smax = -(order + 1) * np.log(referr) / (2 * eta * width)
                z01 = (X[axis][sl] - inner[face]) / (outer[face] - inner[face])
                sigma[axis, sl] = smax * z01 ** order
        self.pmltag, self.pmlinner, self.pmlouter = tag, inner, outer
        self.pmlptr = np.ascontiguousarray(ptr, dtype=np.int32)
        self.pmlsigma = np.ascontiguousarray(sigma.reshape(-1))
        self.pml_order, self.pml_referr = float(order), float(referr)

    def set_incident(self, on, k, omega, h_amp, e_amp):
        """userinc-style plane wave on the face points where `on` (nxzfl) is
        set: fH += h_amp u, fE += e_amp u, u = cos(k . x - omega t)."""
        self.inc_on = None if on is None else np.ascontiguousarray(on, dtype=np.int8)
        self.inc = (tuple(float(x) for x in k), float(omega), tuple(float(x) for x in h_amp),
                    tuple(float(x) for x in e_amp))

    def set_graphene(self, gindex, params):
        """Graphene sheet current (cem_3d_graphene_current) on the face
        points gindex (0-based nxzfl numbers) with the twelve sheet
        parameters per point (params: (n, 12) or (12,) for all).  Its state
        (fjn slots 2..6) is G: (15, n), block 3 (slot - 2) + component."""
        self.gindex = np.ascontiguousarray(gindex, dtype=np.int32)
        n = len(self.gindex)
        self.gpar = np.ascontiguousarray(np.broadcast_to(np.asarray(params, dtype=np.float64), (n, 12)))

    def set_drude(self, dindex, a, b):
        """Drude ADE over volume nodes dindex with per-node a (gamma) and
        b (omega_p^2)."""
        self.dindex = np.ascontiguousarray(dindex, dtype=np.int32)
        n = len(self.dindex)
        self.drude_a = np.ascontiguousarray(np.broadcast_to(a, (n,)), dtype=np.float64)
        self.drude_b = np.ascontiguousarray(np.broadcast_to(b, (n,)), dtype=np.float64)

    def set_lorentz(self, lindex, a, b, c):
        """Lorentz ADE (cem_maxwell_lorentz) over volume nodes lindex with
        per-node a, b, c: eps(w) = eps_oo + c/(b - i a w - w^2)."""
        self.lindex = np.ascontiguousarray(lindex, dtype=np.int32)
        n = len(self.lindex)
        self.lor_par = np.ascontiguousarray(np.stack([np.broadcast_to(a, (n,)), np.broadcast_to(b, (n,)),
                                                      np.broadcast_to(c, (n,))], axis=1).reshape(-1),
                                            dtype=np.float64)

    def dt_from_cfl(self, cfl):
        """set_dt (src/cem_drive.F:351-395): dt = |p12| if p12 < 0 else p12*dxmin."""
        return abs(cfl) if cfl < 0 else cfl * self.dxmin

    # -- operator ---------------------------------------------------------
    def _mesh(self):
        m = _Mesh()
        m.n1, m.nelt, m.npts, m.nxzfl = self.n1, self.nelt, self.npts, self.nxzfl
        m.D, m.w3 = _p(self._Dflat), _p(self.w3)
        for k in ("rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz", "bm1"):
            setattr(m, k, _p(self.geom[k]))
        for k in ("unx", "uny", "unz", "area"):
            setattr(m, k, _p(self.face[k]))
        m.cemface, m.partner, m.cempec, m.ncempec = (_p(self.cemface), _p(self.partner),
                                                      _p(self.cempec), self.ncempec)
        m.Y0, m.Y1, m.Z0, m.Z1 = _p(self.Y0), _p(self.Y1), _p(self.Z0), _p(self.Z1)
        m.hbm1, m.ebm1 = _p(self.hbm1), _p(self.ebm1)
        m.C0 = 1.0 if self.spec.upwind else 0.0
        if self.dindex is not None and len(self.dindex):
            m.ndrude = len(self.dindex)
            m.dindex, m.drude_a, m.drude_b = _p(self.dindex), _p(self.drude_a), _p(self.drude_b)
        else:
            m.ndrude = 0
        if self.lindex is not None and len(self.lindex):
            m.nlorentz = len(self.lindex)
            m.lindex, m.lor_par = _p(self.lindex), _p(self.lor_par)
        else:
            m.nlorentz = 0
        if self.src_amp is not None:
            m.src_amp = _p(self.src_amp)
        m.src_dir = (C.c_double * 3)(*self.src_dir)
        m.src_t0, m.src_tw = self.src_t0, self.src_tw
        m.src_kind, m.src_omega, m.src_phase = self.src_kind, self.src_omega, self.src_phase
        m.srflx = _p(self.srflx)
        m.eps, m.mu = _p(self.eps), _p(self.mu)
        m.xm1, m.ym1, m.zm1 = _p(self.xm1), _p(self.ym1), _p(self.zm1)
        if self.pmlptr is not None and len(self.pmlptr):
            m.npml, m.pmlptr, m.pmlsigma = len(self.pmlptr), _p(self.pmlptr), _p(self.pmlsigma)
        else:
            m.npml = 0
        if self.inc_on is not None:
            k, om, hh, ee = self.inc
            m.inc_on = _p(self.inc_on)
            m.inc_k = (C.c_double * 3)(*k)
            m.inc_omega = om
            m.inc_h = (C.c_double * 3)(*hh)
            m.inc_e = (C.c_double * 3)(*ee)
        if self.gindex is not None and len(self.gindex):
            m.ngraph, m.gindex, m.gpar = len(self.gindex), _p(self.gindex), _p(self.gpar)
        else:
            m.ngraph = 0
        self._keep = m
        return m

    def residual(self, H, E, t=0.0, J=None, L=None, P=None, G=None):
        """(resH, resE[, resJ][, resL][, resP]) = cem_maxwell_op(H, E) incl.
        inverse mass; J the Drude current (3*ndrude), L the Lorentz state (J
        then P, 6*nlorentz), P the PML state (pmlbn, pmldn: (6, npts))."""
        H = np.ascontiguousarray(H, dtype=np.float64).reshape(3, self.npts)
        E = np.ascontiguousarray(E, dtype=np.float64).reshape(3, self.npts)
        rH, rE = np.zeros_like(H), np.zeros_like(E)
        m = self._mesh()
        rJ = rL = None
        if J is not None:
            J = np.ascontiguousarray(J, dtype=np.float64)
            rJ = np.zeros_like(J)
        if L is not None:
            L = np.ascontiguousarray(L, dtype=np.float64)
            rL = np.zeros_like(L)
        rP = None
        if P is not None:
            P = np.ascontiguousarray(P, dtype=np.float64).reshape(6, self.npts)
            rP = np.zeros_like(P)
        rG = None
        if G is not None:
            G = np.ascontiguousarray(G, dtype=np.float64).reshape(15, -1)
            rG = np.zeros_like(G)
        lib().orc_maxwell_op_all(C.byref(m), t, _p(H), _p(E), _p(rH), _p(rE), _p(J), _p(rJ), _p(L), _p(rL),
                                 _p(P), _p(rP), _p(G), _p(rG))
        out = (rH, rE)
        if J is not None:
            out += (rJ,)
        if L is not None:
            out += (rL,)
        if P is not None:
            out += (rP,)
        if G is not None:
            out += (rG,)
        return out

    def curl(self, H, E):
        """(resH, resE) = (-W curl E, W curl H): cem_maxwell's volume term
        (maxwell_wght_curl with chsign), W = bm1."""
        H = np.ascontiguousarray(H, dtype=np.float64).reshape(3, self.npts)
        E = np.ascontiguousarray(E, dtype=np.float64).reshape(3, self.npts)
        rH, rE = np.zeros_like(H), np.zeros_like(E)
        m = self._mesh()
        lib().orc_curl(C.byref(m), _p(H), _p(E), _p(rH), _p(rE))
        return rH, rE

    def run(self, H, E, dt, nsteps, t=0.0, kH=None, kE=None, J=None, kJ=None, L=None, kL=None, P=None, kP=None,
            G=None, kG=None):
        """nsteps of cem_maxwell_op_rk from (H, E) at time t.  Returns a dict
        with H, E, kH, kE, time (and J, kJ with Drude, L, kL with Lorentz,
        P, kP with a PML)."""
        H = np.array(H, dtype=np.float64).reshape(3, self.npts)
        E = np.array(E, dtype=np.float64).reshape(3, self.npts)
        kH = np.zeros_like(H) if kH is None else np.array(kH, dtype=np.float64)
        kE = np.zeros_like(E) if kE is None else np.array(kE, dtype=np.float64)
        rH, rE = np.zeros_like(H), np.zeros_like(E)
        rJ = None
        if self.dindex is not None and len(self.dindex):
            nd = len(self.dindex)
            J = np.zeros(3 * nd) if J is None else np.array(J, dtype=np.float64)
            kJ = np.zeros(3 * nd) if kJ is None else np.array(kJ, dtype=np.float64)
            rJ = np.zeros(3 * nd)
        rL = None
        if self.lindex is not None and len(self.lindex):
            nl = len(self.lindex)
            L = np.zeros(6 * nl) if L is None else np.array(L, dtype=np.float64)
            kL = np.zeros(6 * nl) if kL is None else np.array(kL, dtype=np.float64)
            rL = np.zeros(6 * nl)
        rP = None
        if self.pmlptr is not None and len(self.pmlptr):
            P = np.zeros((6, self.npts)) if P is None else np.array(P, dtype=np.float64).reshape(6, self.npts)
            kP = np.zeros((6, self.npts)) if kP is None else np.array(kP, dtype=np.float64).reshape(6, self.npts)
            rP = np.zeros((6, self.npts))
        rG = None
        if self.gindex is not None and len(self.gindex):
            ng = len(self.gindex)
            G = np.zeros((15, ng)) if G is None else np.array(G, dtype=np.float64).reshape(15, ng)
            kG = np.zeros((15, ng)) if kG is None else np.array(kG, dtype=np.float64).reshape(15, ng)
            rG = np.zeros((15, ng))
        tt = np.array([t], dtype=np.float64)
        m = self._mesh()
        lib().orc_rk_run_all(C.byref(m), _p(H), _p(E), _p(kH), _p(kE), _p(rH), _p(rE),
                             _p(J), _p(kJ), _p(rJ), _p(L), _p(kL), _p(rL), _p(P), _p(kP), _p(rP),
                             _p(G), _p(kG), _p(rG), dt, nsteps, _p(tt))
        out = dict(H=H, E=E, kH=kH, kE=kE, time=float(tt[0]))
        if J is not None:
            out.update(J=J, kJ=kJ)
        if L is not None:
            out.update(L=L, kL=kL)
        if P is not None:
            out.update(P=P, kP=kP)
        if G is not None:
            out.update(G=G, kG=kG)
        return out

    def vjp(self, H, E, kH, kE, t=0.0, J=None, kJ=None):
        """(aH, aE, geps, gmu): transpose of d residual/d(H,E) applied to
        (kH, kE), and the eps/mu gradients of <(kH,kE), residual>.  With a
        Drude region and (J, kJ) given, also (aJ, ga, gb): the transpose on
        the polarisation current and the gradients with respect to the
        per-node a (gamma) and b (omega_p^2)."""
        n = self.npts
        arr = lambda a: np.ascontiguousarray(a, dtype=np.float64).reshape(3, n)
        H, E, kH, kE = arr(H), arr(E), arr(kH), arr(kE)
        aH, aE = np.zeros((3, n)), np.zeros((3, n))
        ge, gm = np.zeros(n), np.zeros(n)
        m = self._mesh()
        lib().orc_maxwell_vjp(C.byref(m), t, _p(self.eps), _p(self.mu), _p(self.pec_face),
                              _p(H), _p(E), _p(kH), _p(kE), _p(aH), _p(aE), _p(ge), _p(gm))
        if J is None:
            return aH, aE, ge, gm
        nd = len(self.dindex)
        J = np.ascontiguousarray(J, dtype=np.float64).reshape(3 * nd)
        kJ = np.ascontiguousarray(kJ, dtype=np.float64).reshape(3 * nd)
        aJ, ga, gb = np.zeros(3 * nd), np.zeros(nd), np.zeros(nd)
        lib().orc_drude_vjp(C.byref(m), _p(self.eps), _p(E), _p(J), _p(kE), _p(kJ), _p(aE), _p(aJ),
                            _p(ge), _p(ga), _p(gb))
        return aH, aE, ge, gm, aJ, ga, gb

    def loss_and_grad(self, H0, E0, dt, nsteps, probes, target, t0=0.0, drude=False):
        """Probe loss L = sum_p (Ez(T)[p] - target[p])^2 after nsteps RK steps
        from (H0, E0), and its gradient with respect to eps and mu per node by
        the discrete adjoint of the 2N-storage RK (every stage input kept).
        With drude=True the Drude current (zero at t0) is advanced too and
        the gradients with respect to its per-node a and b are returned as
        well: (loss, geps, gmu, (H, E), ga, gb)."""
        a, b, c = rk_coeffs()
        H = np.array(H0, dtype=np.float64).reshape(3, self.npts)
        E = np.array(E0, dtype=np.float64).reshape(3, self.npts)
        kH, kE = np.zeros_like(H), np.zeros_like(E)
        nd = len(self.dindex) if drude else 0
        J, kJ = np.zeros(3 * nd), np.zeros(3 * nd)
        stages = []
        t = t0
        for _ in range(nsteps):
            for i in range(5):
                stages.append((H.copy(), E.copy(), J.copy(), t + dt * c[i]))
                if drude:
                    rH, rE, rJ = self.residual(H, E, t=t + dt * c[i], J=J)
                    kJ = a[i] * kJ + dt * rJ
                    J = J + b[i] * kJ
                else:
                    rH, rE = self.residual(H, E, t=t + dt * c[i])
                kH = a[i] * kH + dt * rH
                kE = a[i] * kE + dt * rE
                H = H + b[i] * kH
                E = E + b[i] * kE
            t = t + dt
        diff = E[2, probes] - target
        loss = float(np.sum(diff * diff))
        lH, lE = np.zeros_like(H), np.zeros_like(E)
        np.add.at(lE[2], probes, 2.0 * diff)
        qH, qE = np.zeros_like(H), np.zeros_like(E)
        lJ, qJ = np.zeros(3 * nd), np.zeros(3 * nd)
        ge, gm = np.zeros(self.npts), np.zeros(self.npts)
        ga, gb = np.zeros(nd), np.zeros(nd)
        for sidx in range(len(stages) - 1, -1, -1):
            i = sidx % 5
            Hs, Es, Js, ts = stages[sidx]
            tH = qH + b[i] * lH
            tE = qE + b[i] * lE
            if drude:
                tJ = qJ + b[i] * lJ
                aH, aE, g1, g2, aJ, g3, g4 = self.vjp(Hs, Es, tH, tE, t=ts, J=Js, kJ=tJ)
                lJ = lJ + dt * aJ
                qJ = a[i] * tJ
                ga += dt * g3
                gb += dt * g4
            else:
                aH, aE, g1, g2 = self.vjp(Hs, Es, tH, tE, t=ts)
            lH = lH + dt * aH
            lE = lE + dt * aE
            qH = a[i] * tH
            qE = a[i] * tE
            ge += dt * g1
            gm += dt * g2
        if drude:
            return loss, ge, gm, (H, E), ga, gb
        return loss, ge, gm, (H, E)

    def cem_error(self, u, exact):
        l2 = np.zeros(1)
        li = np.zeros(1)
        u = np.ascontiguousarray(u, dtype=np.float64)
        exact = np.ascontiguousarray(exact, dtype=np.float64)
        lib().orc_cem_error(_p(u), _p(exact), _p(self.bm1), self.npts, _p(l2), _p(li))
        return float(l2[0]), float(li[0])


def set_num_threads(n):
    lib().orc_set_num_threads(int(n))


def num_threads():
    return int(lib().orc_num_threads())


# ---------------------------------------------------------------------------
# Restart file (parity with the reference's writer unpinned: it needs MPI-IO
# and is not run here; the layout is restated from the source and checked by
# round trips and by the reference's own tests/restart semantics).
# ---------------------------------------------------------------------------

def restart_cells(n1, nelt):
    """(ncell, 8) 0-based node ids of the GLL sub-hexahedra, in the order of
    vtk_dump_header_swap (src/io_dumpvtk.F:660-681): per element eg, level
    j = 1..n1-1 from nfcnt = (j-1) n1^2 + (eg-1) n1^3, the (n1-1)^2 squares
    of the level stepping nfcnt by one and over each row's last node."""
    n2, n3 = n1 * n1, n1 ** 3
    out = []
    for eg in range(1, nelt + 1):
        for j in range(1, n1):
            nfcnt = (j - 1) * n2 + (eg - 1) * n3
            for _ in range((n1 - 1) * (n1 - 1)):
                a = [nfcnt, nfcnt + 1, nfcnt + n1 + 1, nfcnt + n1]
                out.append(a + [v + n2 for v in a])
                nfcnt += 1
                if (nfcnt + 1) % n1 == 0:
                    nfcnt += 1
    return np.array(out, dtype=np.int64).reshape(-1, 8)


def restart_bytes(istep, dumpno, n1, time, dt, xyz, E, H, double=True, np_ranks=1):
    """The bytes cem_restart_out writes for one rank (src/io.F:411-490;
    writeheader4 src/io_co.c:115-148, writenodes4 :152-227, writecells4
    :238-405, writefield4 / writefield4_double :412-554): header lines
    printed with %d and %f, big-endian float32 node coordinates, 9 int32 per
    cell (8 and the node ids), int32 cell types 12, then the E and H vectors
    (big-endian float64, or float32), the last field's trailing " \\n" left
    in the writer's buffer.  xyz, E, H: (3, npts)."""
    xyz, E, H = (np.asarray(a, dtype=np.float64) for a in (xyz, E, H))
    npts = xyz.shape[1]
    cells = restart_cells(n1, npts // n1 ** 3)
    ncell = len(cells)
    b = bytearray()
    b += b"# vtk DataFile Version 3.0 \n"
    b += b"%d %d %d %d %f %f Electromagnetic Field  \n" % (istep, dumpno, n1, np_ranks, time, dt)
    b += b"BINARY \n"
    b += b"DATASET UNSTRUCTURED_GRID \n"
    b += b"POINTS  %d  float  \n" % npts
    b += xyz.T.astype(">f4").tobytes()
    b += b"\nCELLS %d  %d \n" % (ncell, 9 * ncell)
    conn = np.concatenate([np.full((ncell, 1), 8), cells], axis=1)
    b += conn.astype(">i4").tobytes()
    b += b"\nCELL_TYPES %d \n" % ncell
    b += np.full(ncell, 12).astype(">i4").tobytes()
    b += b"\nPOINT_DATA %d \n" % npts
    kind, fmt = (b"double", ">f8") if double else (b"float", ">f4")
    b += b"VECTORS %s  %s \n" % (b"E ", kind)
    b += E.T.astype(fmt).tobytes()
    b += b" \n"
    b += b"VECTORS %s  %s \n" % (b"H ", kind)
    b += H.T.astype(fmt).tobytes()
    return bytes(b)


class GeneralCase(BoxCase):
    """A reference-layout case on a general hexahedral mesh (curved,
    rotated or sheared elements).  The node coordinates are given; the rest
    is derived here as the reference's setup does: geometry by orc_geom
    (GLMAPM1 / AREA3 from the coordinates), cemface by orc_set_fc_ptr, the
    gs pairing as face nodes at equal coordinates on faces whose centroids
    coincide, directly or after one of the given periodic translations
    (the pairing setup_dgds2 produces from the global vertex numbering,
    src/nek5_connect11.F:2179-2347), and the PEC list from pec_face
    (nelt, 6) in the reference's face order (cem_maxwell_pec_init)."""

    def __init__(self, N, xm1, ym1, zm1, pec_face, periods=(), upwind=True, eps=None, mu=None):
        L = lib()
        n1 = N + 1
        self.n1 = n1
        n2, n3 = n1 * n1, n1 ** 3
        self.xm1, self.ym1, self.zm1 = (np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (xm1, ym1, zm1))
        self.npts = npts = len(self.xm1)
        self.nelt = nelt = npts // n3
        self.nxzfl = nxzfl = 6 * n2 * nelt
        self.spec = BoxSpec(nel=(nelt, 1, 1), N=N, bc=("PEC",) * 3, upwind=upwind)
        self.z, self.w = zwgll(n1)
        self.D = dgll(self.z)
        self._Dflat = np.ascontiguousarray(self.D.reshape(-1, order="F"))
        g = {k: np.zeros(npts) for k in ("rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz", "jac", "bm1")}
        self.w3 = np.zeros(n3)
        fa = {k: np.zeros(nxzfl) for k in ("unx", "uny", "unz", "area")}
        L.orc_geom(n1, nelt, _p(self._Dflat), _p(self.w), _p(self.xm1), _p(self.ym1), _p(self.zm1),
                   *[_p(g[k]) for k in ("rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz", "jac", "bm1")],
                   _p(self.w3), _p(fa["unx"]), _p(fa["uny"]), _p(fa["unz"]), _p(fa["area"]))
        self.geom, self.face, self.bm1 = g, fa, g["bm1"]
        self.cemface = np.zeros(nxzfl, dtype=np.int32)
        L.orc_set_fc_ptr(n1, nelt, _p(self.cemface))
        self.partner = self._partner_by_position(np.asarray(periods, dtype=np.float64).reshape(-1, 3))
        self.pec_face = np.ascontiguousarray(np.asarray(pec_face, dtype=np.int8).reshape(-1))
        self.cempec = np.zeros(nxzfl, dtype=np.int32)
        self.ncempec = int(L.orc_pec_init(n1, nelt, _p(self.pec_face), _p(self.cempec)))
        self.set_materials(eps, mu)
        self.dxmin = float(L.orc_get_dxmin(n1, nelt, _p(self.xm1), _p(self.ym1), _p(self.zm1)))
        self.srflx = np.zeros(6 * nxzfl)
        self.src_amp = None
        self.src_dir, self.src_t0, self.src_tw = (0.0, 0.0, 1.0), 0.0, 1.0
        self.dindex = self.drude_a = self.drude_b = None
        self.lindex = self.lor_par = None
        self.src_kind, self.src_omega, self.src_phase = 0, 0.0, 0.0
        self.pmlptr = self.pmlsigma = None
        self.pmltag = np.zeros(nelt, dtype=np.int64)
        self.inc_on = self.gindex = self.inc = None

    def _partner_by_position(self, periods):
        n1, nelt, nxzfl = self.n1, self.nelt, self.nxzfl
        n2 = n1 * n1
        vol = self.cemface.astype(np.int64)                                      # 0-based here
        P = np.stack([self.xm1[vol], self.ym1[vol], self.zm1[vol]], axis=1)      # (nxzfl, 3)
        cen = P.reshape(nelt * 6, n2, 3).mean(axis=1)
        scale = max(1.0, float(np.max(np.abs(P))))
        q = 1e-7 * scale

        def key(c):
            return tuple(np.floor(c / q + 0.5).astype(np.int64))

        faces = {}
        for fi in range(6 * nelt):
            faces.setdefault(key(cen[fi]), []).append(fi)
        partner = np.full(nxzfl, -1, dtype=np.int32)
        shifts = [np.zeros(3)] + [s for T in periods for s in (T, -T)]
        for fi in range(6 * nelt):
            for T in shifts:
                other = [fj for fj in faces.get(key(cen[fi] + T), []) if fj != fi]
                if other:
                    fj = other[0]
                    A, B = P[fi * n2:(fi + 1) * n2] + T, P[fj * n2:(fj + 1) * n2]
                    d = np.linalg.norm(A[:, None, :] - B[None, :, :], axis=2)
                    b = np.argmin(d, axis=1)
                    assert np.max(d[np.arange(n2), b]) < 1e-6 * scale
                    partner[fi * n2:(fi + 1) * n2] = fj * n2 + b
                    break
        return partner
