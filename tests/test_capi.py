"""C-ABI checks that need no GPU: the library loads, exports every entry
point include/nekcem_hip.h declares, and rejects bad input with the right
status before touching a device."""
import ctypes as C
import re

import numpy as np
import pytest

from nekcem_amd import _lib
from nekcem_amd.mesh import BoxMesh


def _declared():
    src = open(_lib.HEADER_PATH).read()
    return sorted(set(re.findall(r"\b(nkc_[a-z_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = _declared()
    assert len(names) >= 12
    for n in names:
        assert hasattr(L, n), n
    assert {s[0] for s in _lib.SIGNATURES} == set(names)


def _desc(mesh, **over):
    eps, mu, Z0, Z1, Y0, Y1 = mesh.materials()
    arrs = dict(eps=eps, mu=mu, Z0=Z0, Z1=Z1, Y0=Y0, Y1=Y1)
    d = _lib.NkcDesc()
    d.N, d.ifupwind, d.nelt = mesh.N, 1, mesh.nelt
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    for name in ("dxm1", "wxm1", "rxmn", "rymn", "rzmn", "sxmn", "symn", "szmn",
                 "txmn", "tymn", "tzmn", "jacm", "bmn", "unxm", "unym", "unzm", "aream"):
        a = over.get(name, getattr(mesh, name))
        arrs[name] = a
        setattr(d, name, dp(a))
    d.cemface, d.face_partner = ip(mesh.cemface), ip(mesh.face_partner)
    d.ncempec = 0
    d.permittivity, d.permeability = dp(eps), dp(mu)
    d.Z_0, d.Z_1, d.Y_0, d.Y_1 = dp(Z0), dp(Z1), dp(Y0), dp(Y1)
    return d, arrs


def test_create_rejects_null():
    L = _lib.load()
    h = C.c_void_p()
    assert L.nkc_create(None, C.byref(h)) == _lib.NKC_ERR_ARG
    assert b"null" in L.nkc_last_error()


def test_create_rejects_unsupported_order():
    L = _lib.load()
    mesh = BoxMesh((2, 2, 2), 3)
    d, keep = _desc(mesh)
    d.N = 12
    h = C.c_void_p()
    assert L.nkc_create(C.byref(d), C.byref(h)) == _lib.NKC_ERR_UNSUPPORTED


def test_create_takes_non_box_geometry_to_the_general_tables():
    """A sheared element (nonzero off-diagonal cofactor) passes the host
    set-up on the general-hexahedra tables (without a GPU nkc_create then
    stops at the device: NKC_ERR_NODEVICE); a face partner that does not
    pair back is refused there."""
    L = _lib.load()
    mesh = BoxMesh((2, 2, 2), 4)
    ry = mesh.rymn.copy()
    ry[: mesh.n1 ** 3] = 0.3 * mesh.rxmn[0]
    d, keep = _desc(mesh, rymn=ry)
    h = C.c_void_p()
    rc = L.nkc_create(C.byref(d), C.byref(h))
    assert rc in (_lib.NKC_OK, _lib.NKC_ERR_NODEVICE), L.nkc_last_error()
    if rc == _lib.NKC_OK:
        L.nkc_destroy(h)
    fp = mesh.face_partner.copy()
    s = int(np.nonzero(fp)[0][0])
    fp[s] = fp[s] + 1 if fp[s] < len(fp) else fp[s] - 1     # the next face node: it pairs elsewhere
    mesh.face_partner = fp
    d, keep = _desc(mesh, rymn=ry)
    assert L.nkc_create(C.byref(d), C.byref(h)) == _lib.NKC_ERR_ARG
    assert b"pair back" in L.nkc_last_error() or b"range" in L.nkc_last_error()


def test_create_rejects_bad_cemface():
    L = _lib.load()
    mesh = BoxMesh((2, 2, 2), 4)
    cf = mesh.cemface.copy()
    cf[[0, 1]] = cf[[1, 0]]
    mesh.cemface = cf
    d, keep = _desc(mesh)
    h = C.c_void_p()
    assert L.nkc_create(C.byref(d), C.byref(h)) == _lib.NKC_ERR_ARG


def test_handle_calls_reject_null():
    L = _lib.load()
    x = np.zeros(8)
    p = x.ctypes.data_as(C.POINTER(C.c_double))
    assert L.nkc_destroy(None) == _lib.NKC_ERR_ARG
    assert L.nkc_step(None, p, 0.1, 1) == _lib.NKC_ERR_ARG
    assert L.nkc_set_state(None, p, p, None, None, 8) == _lib.NKC_ERR_ARG
    assert L.nkc_get_state(None, p, p, None, None, 8) == _lib.NKC_ERR_ARG
    assert L.nkc_residual(None, 0.0, p, p, 8) == _lib.NKC_ERR_ARG
    assert L.nkc_check_finite(None, p) == _lib.NKC_ERR_ARG
    assert L.nkc_sync(None) == _lib.NKC_ERR_ARG
