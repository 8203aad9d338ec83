"""hydro3d_amd -- MI355X-native pressure-Poisson multigrid for Hydro3D.

The product is libh3dmg.so (HIP kernels for gfx950 behind the C-ABI of
include/h3dmg.h); this package holds its sources (csrc/), the Fortran
binding (fortran/), the Python mirror of the interface (solver.py) and the
seeded workload recipes (configs.py).
"""
from .solver import MultigridSolver  # noqa: F401

__all__ = ["MultigridSolver"]
