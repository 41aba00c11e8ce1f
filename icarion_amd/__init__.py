"""icarion_amd - MI355X (gfx950) batch ion-step path of the ICARION simulator.

The product is the C-ABI library libicarion_gpu.so (include/icarion_gpu.h):
fused HIP kernels for the engine step (field evaluation, RK4/Boris push,
HSS/EHSS/Langevin collisions, boundary elimination).  This package holds its
sources (csrc/), its build recipe (build.py), a ctypes binding (gpu.py), the
host-side metric definitions (metrics.py), the trajectory buffer
(trajectory.py), the surrogate training driver (surrogate.py) and synthetic
workloads (workloads.py).
"""
from . import _abi  # noqa: F401
from .ensemble import Ensemble  # noqa: F401

__all__ = ["Ensemble"]
