"""Build libnekcem_hip.so in-tree with hipcc for gfx950.

``python -m nekcem_amd.build`` or ``nekcem_amd.build.build()``.  The library
lands in ``nekcem_amd/lib/``; the Fortran iso_c_binding module is compiled
alongside when amdflang is present.
"""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "nekcem_hip.hip")
OUT = os.path.join(HERE, "lib", "libnekcem_hip.so")
ARCH = os.environ.get("NKC_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
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


def build(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    tmp = OUT + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include"), "-o", tmp, SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stderr[-4000:])
    if verbose and r.stderr:
        print(r.stderr)
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
