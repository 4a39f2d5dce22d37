"""The host-side table builders under AddressSanitizer and
UndefinedBehaviorSanitizer (CPU only): tests/sanitize_host_tables.cpp
includes nekcem_amd/csrc/mesh_tables.hpp (compact_reference, build_box,
build_hex, pml_tables) and is built with g++ -fsanitize=address,undefined
-fno-sanitize-recover=all.  It runs on the reference-layout arrays of four
boxes (written here from nekcem_amd/mesh.py), on the element corners and
boundary codes of the reference's own case files (read by nekcem_amd/rea.py),
on restart files (restart_io.hpp's writer against the oracle's restatement
of the layout, byte for byte, and its reader),
on boxes of every order 3..10 with the PML set-up on their walls, and on
malformed inputs that must be refused; any out-of-bounds access, leak or
undefined behaviour aborts the run."""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from nekcem_amd import rea
from nekcem_amd.mesh import BoxMesh
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = os.path.join(HERE, "golden", "cases")
SRC = os.path.join(HERE, "sanitize_host_tables.cpp")

BOXES = [
    dict(nel=(3, 2, 2), N=5, bc=("P", "P", "P"), lo=(0, 0, 0), hi=(1.0, 0.8, 1.2)),
    dict(nel=(2, 2, 3), N=7, bc=("PEC", "PEC", "PEC"), lo=(-1, -1, -1), hi=(1, 1, 1)),
    dict(nel=(1, 1, 1), N=4, bc=("PEC", "P", "P"), lo=(0, 0, 0), hi=(2, 1, 1)),
    dict(nel=(2, 3, 2), N=9, bc=("P", "PEC", "P"), lo=(0, 0, 0), hi=(1, 1, 1)),
]
HEX_CASES = ["3dboxpec", "3dboxper", "3ddielectric", "3dboxpml", "3dgraphene"]


def _write_ref(path, case, seed):
    rng = np.random.default_rng(seed)
    m = BoxMesh(case["nel"], case["N"], case["lo"], case["hi"], case["bc"])
    eps_e = rng.uniform(1, 4, m.nelt)
    eps, mu, Z0, Z1, Y0, Y1 = m.materials(np.repeat(eps_e, m.n1 ** 3), None)
    hdr = np.array([case["N"], 1, m.nelt, m.ncempec, *case["nel"],
                    *[1 if b == "PEC" else 0 for b in case["bc"]]], dtype=np.int64)
    parts = [hdr.tobytes(), np.array([*case["lo"], *case["hi"]], dtype=np.float64).tobytes(),
             eps_e.tobytes()]
    for name in ("dxm1", "wxm1", "rxmn", "rymn", "rzmn", "sxmn", "symn", "szmn", "txmn", "tymn", "tzmn",
                 "jacm", "bmn", "unxm", "unym", "unzm", "aream"):
        parts.append(np.ascontiguousarray(getattr(m, name), dtype=np.float64).tobytes())
    for a in (eps, mu, Z0, Z1, Y0, Y1):
        parts.append(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    parts.append(np.ascontiguousarray(m.cemface, dtype=np.int32).tobytes())
    parts.append(np.ascontiguousarray(m.face_partner, dtype=np.int32).tobytes())
    if m.ncempec:
        parts.append(np.ascontiguousarray(m.cempec, dtype=np.int32).tobytes())
    with open(path, "wb") as f:
        f.write(b"".join(parts))


def _write_hex(path, case):
    r = rea.read_case(os.path.join(CASES, f"{case}.rea"))
    xc, yc, zc = (np.ascontiguousarray(r[k], dtype=np.float64).reshape(-1, 8) for k in ("xc", "yc", "zc"))
    nelt = xc.shape[0]
    codes = "".join(str(c).ljust(3)[:3] for c in np.asarray(r["cbc"]).reshape(-1)).encode("ascii")
    bc = np.ascontiguousarray(r["bc"], dtype=np.float64).reshape(nelt * 30)
    with open(path, "wb") as f:
        f.write(np.array([8, 1, nelt], dtype=np.int64).tobytes())
        for a in (xc, yc, zc, bc):
            f.write(a.tobytes())
        f.write(codes)


def _write_restart(base, n1, nelt, double, seed):
    rng = np.random.default_rng(seed)
    npts = nelt * n1 ** 3
    xyz, E, H = (rng.standard_normal((3, npts)) for _ in range(3))
    istep, dumpno, t, dt = 1234, 7, 0.1234567891, 2.5e-4
    with open(base + ".bin", "wb") as f:
        f.write(np.array([n1, nelt, istep, dumpno, 1 if double else 0], dtype=np.int64).tobytes())
        f.write(np.array([t, dt], dtype=np.float64).tobytes())
        for a in (xyz, E, H):
            f.write(np.ascontiguousarray(a).tobytes())
    with open(base + ".vtk", "wb") as f:
        f.write(O.restart_bytes(istep, dumpno, n1, t, dt, xyz, E, H, double=double))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_tables_under_asan_ubsan():
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "sanitize_host_tables")
        r = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                            "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
                            "-I", os.path.join(os.path.dirname(HERE), "include"), SRC, "-o", exe],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
        args = []
        for i, case in enumerate(BOXES):
            p = os.path.join(tmp, f"ref_{i}.bin")
            _write_ref(p, case, i)
            args.append(p)
        for i, (n1, nelt, dbl) in enumerate(((4, 3, True), (8, 2, True), (6, 1, False))):
            p = os.path.join(tmp, f"rst_{i}")
            _write_restart(p, n1, nelt, dbl, i)
            args.append(p + ".bin")
        for case in HEX_CASES:
            p = os.path.join(tmp, f"hex_{case}.bin")
            _write_hex(p, case)
            args.append(p)
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
                   UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
        r = subprocess.run([exe] + args, capture_output=True, text=True, env=env, timeout=600)
        out = r.stdout + r.stderr
        assert r.returncode == 0, out[-4000:]
        assert "ALL OK" in r.stdout
        assert "runtime error" not in out and "AddressSanitizer" not in out
        n_ok = r.stdout.count("\nok  ") + r.stdout.startswith("ok  ")
        assert n_ok >= 100
