"""MI355X-native implementation of the Euler3D explicit-solver hot path.

The product is libe3d_hip.so (HIP kernels for gfx950 behind the C ABI of include/e3d.h) and the
E3D_Solver executable built on it; this package binds that library for Python callers.
"""
from .native import E3DError, Params, default_params, version  # noqa: F401
from .solver import Batch, EulerSolver, Mesh, SimConfig, STATE_FIELDS  # noqa: F401
