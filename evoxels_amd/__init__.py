"""evoxels_amd: MI355X (gfx950) 'hip' backend for evoxels' time-stepping hot path.

Public names mirror evoxels/__init__.py:3-13.  The kernels live in
libevoxels_hip.so (C ABI: include/evoxels_hip.h); build it with
``python -m evoxels_amd.build``.
"""
from .voxelfields import VoxelFields
from .precompiled_solvers.cahn_hilliard import run_cahn_hilliard_solver
from .precompiled_solvers.allen_cahn import run_allen_cahn_solver
from .inversion import BatchedInversion, InversionModel, SaveAt

__all__ = ["VoxelFields", "run_cahn_hilliard_solver", "run_allen_cahn_solver", "InversionModel"]
