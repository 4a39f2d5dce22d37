"""Build libnekcem_hip.so in-tree with hipcc for gfx950.

``python -m nekcem_amd.build`` or ``nekcem_amd.build.build()``.  The library
lands in ``nekcem_amd/lib/``; the Fortran iso_c_binding module is compiled
alongside when amdflang is present.  The host code and each polynomial
order's kernel instances are separate translation units (``csrc/order_n*.hip``),
compiled in parallel into ``nekcem_amd/_build/`` and linked into one library.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
SRC = os.path.join(CSRC, "nekcem_hip.hip")
ORDERS = (4, 5, 6, 7, 8, 9, 10)
OBJDIR = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "lib", "libnekcem_hip.so")
ARCH = os.environ.get("NKC_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # keep hipcc from forming packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # fuse a*b+c only where one source expression says so: every kernel
    # variant (layouts, batches) then rounds identically
    "-ffp-contract=on",
    "-Wall", "-Wno-unused-function",
]


def hipcc():
    for c in (os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc"), shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def sources():
    return [SRC, os.path.join(CSRC, "geo_inst.hip")] + [os.path.join(CSRC, f"order_n{n}.hip") for n in ORDERS]


def _compile(src, extra):
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    cmd = [hipcc()] + HIPCC_FLAGS + list(extra) + ["-I", os.path.join(ROOT, "include"), "-c", "-o", obj, src]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {os.path.basename(src)}:\n" + r.stderr[-4000:])
    return obj, r.stderr


def build(verbose=False, extra=(), jobs=None):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, min(os.cpu_count() or 1, 8)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        res = list(ex.map(lambda s: _compile(s, extra), srcs))
    if verbose:
        for _, err in res:
            if err:
                print(err)
    tmp = OUT + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + [o for o, _ in res]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    os.replace(tmp, OUT)
    build_fortran()
    return OUT


def build_fortran():
    """Compile the iso_c_binding module and the box driver (Fortran host side)."""
    fc = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "amdflang")
    if not os.path.exists(fc):
        return None
    fdir = os.path.join(HERE, "fortran")
    mk = os.path.join(fdir, "Makefile")
    if not os.path.exists(mk):
        return None
    r = subprocess.run(["make", "-s", "-C", fdir, f"FC={fc}"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("Fortran build failed:\n" + r.stdout[-2000:] + r.stderr[-4000:])
    return fdir


if __name__ == "__main__":
    print(build(verbose=True))
