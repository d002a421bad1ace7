"""Precompiled solvers (evoxels/precompiled_solvers)."""
