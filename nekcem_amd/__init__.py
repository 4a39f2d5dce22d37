"""nekcem_amd -- MI355X-native DG-SEM Maxwell time step for NekCEM.

The product is libnekcem_hip.so (C-ABI in include/nekcem_hip.h, HIP kernels
in nekcem_amd/csrc).  This package holds its host-side mirror of NekCEM's
Maxwell driver (maxwell.py), the box-mesh setup in the reference layout
(mesh.py), the ctypes binding (_lib.py) and the in-tree build (build.py).
"""
from .mesh import BoxMesh  # noqa: F401
from .maxwell import MaxwellSolver, cem_maxwell_init, cem_error  # noqa: F401
