"""The Fortran host side: nekcem_amd/fortran/box_driver runs NekCEM's
tests/3dboxper case (4x4x8, N=8, dt=2e-4, 50 steps, userchk at L2 <= 5e-10,
Linf <= 5e-9) through the iso_c_binding module and the C-ABI, then sets a
zero-sigma PML from Fortran arrays (the fields must not change; B and D come
back on the PML elements only) and a Drude-only graphene sheet from the
/usergraph/ arrays (its Drude slot charges up, the other slots stay zero)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FDIR = os.path.join(os.path.dirname(HERE), "nekcem_amd", "fortran")
EXE = os.path.join(FDIR, "_build", "box_driver")


def test_fortran_module_builds():
    if not os.path.exists("/opt/rocm/bin/amdflang"):
        pytest.skip("amdflang not installed")
    assert os.path.exists(EXE), "build() compiles the Fortran driver"
    assert os.path.exists(os.path.join(FDIR, "_build", "nekcem_hip.mod"))


@pytest.mark.gpu
def test_fortran_box_driver_3dboxper():
    print("[case] Fortran box_driver (3dboxper)", flush=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert "3dboxper passed on the HIP path" in r.stdout
    assert "PML entry points ok" in r.stdout
    assert "graphene entry points ok" in r.stdout
