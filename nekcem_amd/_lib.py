"""ctypes binding of libnekcem_hip.so (the C-ABI in include/nekcem_hip.h).

The library is built in-tree by ``nekcem_amd.build.build()`` into
``nekcem_amd/lib/libnekcem_hip.so``.  There is no fallback: if the library is
missing, ``load()`` raises, so a missing build can never silently turn into a
CPU computation.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# NKC_LIBRARY may point at an alternative build of the same sources (e.g. a
# tuning variant); by default the in-tree build is used.
LIB_PATH = os.environ.get("NKC_LIBRARY") or os.path.join(_HERE, "lib", "libnekcem_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "nekcem_hip.h")

NKC_OK = 0
NKC_ERR_ARG = 1
NKC_ERR_HIP = 2
NKC_ERR_UNSUPPORTED = 3
NKC_ERR_BLOWUP = 4
NKC_ERR_NODEVICE = 5

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class NkcDesc(C.Structure):
    """Mirror of struct nkc_desc."""
    _fields_ = [
        ("N", C.c_int32), ("ifupwind", C.c_int32), ("nelt", C.c_int64),
        ("dxm1", _dp), ("wxm1", _dp),
        ("rxmn", _dp), ("rymn", _dp), ("rzmn", _dp),
        ("sxmn", _dp), ("symn", _dp), ("szmn", _dp),
        ("txmn", _dp), ("tymn", _dp), ("tzmn", _dp),
        ("jacm", _dp), ("bmn", _dp),
        ("unxm", _dp), ("unym", _dp), ("unzm", _dp), ("aream", _dp),
        ("cemface", _ip), ("face_partner", _ip), ("cempec", _ip), ("ncempec", C.c_int64),
        ("permittivity", _dp), ("permeability", _dp),
        ("Z_0", _dp), ("Z_1", _dp), ("Y_0", _dp), ("Y_1", _dp),
    ]


class NkcBoxDesc(C.Structure):
    """Mirror of struct nkc_box_desc."""
    _fields_ = [
        ("N", C.c_int32), ("ifupwind", C.c_int32),
        ("nel", C.c_int32 * 3), ("pec", C.c_int32 * 3),
        ("lo", C.c_double * 3), ("hi", C.c_double * 3),
        ("eps_elem", _dp), ("mu_elem", _dp),
    ]


class NkcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"nekcem_hip error {code}: {msg}")
        self.code = code


_lib = None

# (name, restype, argtypes) for every function in include/nekcem_hip.h
SIGNATURES = [
    ("nkc_create", C.c_int32, [C.POINTER(NkcDesc), C.POINTER(C.c_void_p)]),
    ("nkc_create_box", C.c_int32, [C.POINTER(NkcBoxDesc), C.POINTER(C.c_void_p)]),
    ("nkc_destroy", C.c_int32, [C.c_void_p]),
    ("nkc_set_state", C.c_int32, [C.c_void_p, _dp, _dp, _dp, _dp, C.c_int64]),
    ("nkc_get_state", C.c_int32, [C.c_void_p, _dp, _dp, _dp, _dp, C.c_int64]),
    ("nkc_step", C.c_int32, [C.c_void_p, _dp, C.c_double, C.c_int64]),
    ("nkc_residual", C.c_int32, [C.c_void_p, C.c_double, _dp, _dp, C.c_int64]),
    ("nkc_curl", C.c_int32, [C.c_void_p, _dp, _dp, C.c_int64]),
    ("nkc_set_source", C.c_int32, [C.c_void_p, _dp, _dp, C.c_double, C.c_double]),
    ("nkc_set_drude", C.c_int32, [C.c_void_p, C.c_int64, _ip, _dp, _dp, _dp]),
    ("nkc_get_drude", C.c_int32, [C.c_void_p, _dp, _dp, _dp]),
    ("nkc_set_materials", C.c_int32, [C.c_void_p, _dp, _dp]),
    ("nkc_adjoint_forward", C.c_int32, [C.c_void_p, _dp, C.c_double, C.c_int64, C.c_int64]),
    ("nkc_adjoint_backward", C.c_int32, [C.c_void_p, _dp, _dp, C.c_int64, _dp, _dp]),
    ("nkc_adjoint_get_drude", C.c_int32, [C.c_void_p, _dp, _dp]),
    ("nkc_set_lorentz", C.c_int32, [C.c_void_p, C.c_int64, _ip, _dp, _dp, _dp]),
    ("nkc_get_lorentz", C.c_int32, [C.c_void_p, _dp, _dp, _dp]),
    ("nkc_set_source_time", C.c_int32, [C.c_void_p, C.c_int32, C.c_double, C.c_double]),
    ("nkc_set_pml", C.c_int32, [C.c_void_p, C.c_int64, _ip, _dp, _dp, _dp, C.c_int64]),
    ("nkc_setup_pml", C.c_int32, [C.c_void_p, C.POINTER(C.c_int8), C.c_int32, C.c_double, C.c_double, _dp]),
    ("nkc_set_pml_fields", C.c_int32, [C.c_void_p, _dp, _dp, C.c_int64]),
    ("nkc_get_pml", C.c_int32, [C.c_void_p, C.c_int32, _dp, C.c_int64]),
    ("nkc_set_incident", C.c_int32, [C.c_void_p, C.c_int64, _ip, _dp, C.c_double, _dp, _dp]),
    ("nkc_set_graphene", C.c_int32, [C.c_void_p, C.c_int64, _ip, _dp, _dp, C.c_int64]),
    ("nkc_get_graphene", C.c_int32, [C.c_void_p, C.c_int32, _dp, C.c_int64]),
    ("nkc_set_layout", C.c_int32, [C.c_void_p, C.c_int32]),
    ("nkc_tune", C.c_int32, [C.c_void_p, _dp]),
    ("nkc_get_layout", C.c_int32, [C.c_void_p, _dp]),
    ("nkc_batch_layout", C.c_int32, [C.c_void_p, C.POINTER(C.c_int32), C.c_int64]),
    ("nkc_adjoint_set_store", C.c_int32, [C.c_void_p, C.c_int64]),
    ("nkc_adjoint_set_grads", C.c_int32, [C.c_void_p, C.c_int32]),
    ("nkc_adjoint_release", C.c_int32, [C.c_void_p]),
    ("nkc_adjoint_info", C.c_int32, [C.c_void_p, C.POINTER(C.c_int64)]),
    ("nkc_adjoint_apply", C.c_int32, [C.c_void_p, _dp, _dp, _dp, _dp, C.c_int64]),
    ("nkc_create_hex", C.c_int32, [C.c_int32, C.c_int32, C.c_int64, _dp, _dp, _dp, C.c_char_p, _dp, _dp, _dp,
                                   C.POINTER(C.c_void_p)]),
    ("nkc_set_coords", C.c_int32, [C.c_void_p, _dp, _dp, _dp]),
    ("nkc_init_solution", C.c_int32, [C.c_void_p, C.c_int32, C.c_double]),
    ("nkc_cem_error", C.c_int32, [C.c_void_p, C.c_int32, C.c_double, _dp]),
    ("nkc_cem_error_fields", C.c_int32, [C.c_void_p, _dp, _dp, C.c_int64, _dp]),
    ("nkc_cem_error_noalloc", C.c_int32, [C.c_void_p, C.c_int32, C.c_double, _dp]),
    ("nkc_restart_write", C.c_int32, [C.c_void_p, C.c_char_p, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                      C.c_int32]),
    ("nkc_restart_read", C.c_int32, [C.c_void_p, C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _dp,
                                     _dp]),
    ("nkc_check_finite", C.c_int32, [C.c_void_p, _dp]),
    ("nkc_sync", C.c_int32, [C.c_void_p]),
    ("nkc_stream", C.c_int32, [C.c_void_p, C.POINTER(C.c_void_p)]),
    ("nkc_info", C.c_int32, [C.c_void_p, C.POINTER(C.c_int64)]),
    ("nkc_set_graph", C.c_int32, [C.c_void_p, C.c_int32]),
    ("nkc_set_keep_k", C.c_int32, [C.c_void_p, C.c_int32]),
    ("nkc_batch_create", C.c_int32, [C.c_int32, C.POINTER(NkcBoxDesc), C.POINTER(C.c_void_p)]),
    ("nkc_batch_destroy", C.c_int32, [C.c_void_p]),
    ("nkc_batch_set_state", C.c_int32, [C.c_void_p, C.c_int32, _dp, _dp, C.c_int64]),
    ("nkc_batch_get_state", C.c_int32, [C.c_void_p, C.c_int32, _dp, _dp, C.c_int64]),
    ("nkc_batch_step", C.c_int32, [C.c_void_p, _dp, _dp, C.c_int64]),
    ("nkc_batch_sync", C.c_int32, [C.c_void_p]),
    ("nkc_batch_stream", C.c_int32, [C.c_void_p, C.POINTER(C.c_void_p)]),
    ("nkc_batch_info", C.c_int32, [C.c_void_p, C.POINTER(C.c_int64)]),
    ("nkc_last_error", C.c_char_p, []),
]


def load():
    """Load libnekcem_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with nekcem_amd.build.build() "
                "(hipcc --offload-arch=gfx950); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != NKC_OK:
        msg = load().nkc_last_error().decode(errors="replace")
        raise NkcError(rc, msg)
    return rc
