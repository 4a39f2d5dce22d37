"""The Fortran host side: nekcem_amd/fortran/box_driver runs NekCEM's
tests/3dboxper case (4x4x8, N=8, dt=2e-4, 50 steps, userchk at L2 <= 5e-10,
Linf <= 5e-9) through the iso_c_binding module and the C-ABI, then sets a
zero-sigma PML from Fortran arrays (the fields must not change; B and D come
back on the PML elements only) and a Drude-only graphene sheet from the
/usergraph/ arrays (its Drude slot charges up, the other slots stay zero)."""
import os
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FDIR = os.path.join(os.path.dirname(HERE), "nekcem_amd", "fortran")
EXE = os.path.join(FDIR, "_build", "box_driver")
FC = "/opt/rocm/bin/amdflang"
# The reference tree is read here (in the build container) for its
# declaration files only; it is not present where the GPU tests run.
REF = "/root/reference"


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


@pytest.mark.skipif(not (os.path.exists(FC) and os.path.isdir(os.path.join(REF, "src"))),
                    reason="needs amdflang and the reference's declaration files")
@pytest.mark.parametrize("case", ["3dboxper", "3ddielectric"])
def test_drop_in_shim_compiles_against_reference_declarations(case):
    """The drop-in (nekcem_amd/fortran/cem_maxwell_hip.F: cem_maxwell_hip_init,
    cem_maxwell_op_rk, the host-sync and PML routines) passes amdflang
    -fsyntax-only with the reference's own SIZE of a case and its TOTAL /
    EMWAVE / GEOM / INPUT / PML declarations on the include path, with the
    flags of a NekCEM build (-fdefault-real-8, fixed form, -DMPI
    -DMAXWELL): every COMMON name, type and stride it passes across the
    C-ABI is the reference's.  Nothing of the reference is compiled into an
    object, linked or run."""
    with tempfile.TemporaryDirectory() as tmp:
        r = subprocess.run([FC, "-c", os.path.join(FDIR, "nekcem_hip_mod.f90"), "-J", tmp,
                            "-o", os.path.join(tmp, "nekcem_hip_mod.o")], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        r = subprocess.run([FC, "-fsyntax-only", "-cpp", "-DMPI", "-DMPIIO", "-DMAXWELL", "-fdefault-real-8",
                            "-fdefault-double-8", "-ffixed-form", "-I", tmp,
                            "-I", os.path.join(REF, "tests", case), "-I", os.path.join(REF, "src"),
                            os.path.join(FDIR, "cem_maxwell_hip.F")],
                           capture_output=True, text=True, cwd=tmp)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "error" not in (r.stdout + r.stderr).lower()
