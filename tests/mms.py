"""Method-of-manufactured-solutions helpers shared by the oracle and GPU tests.

Restates the grid convergence harness of evoxels/utils.py:14-132 (and of
docs/notebooks/10-appendix-convergence-testing.ipynb, cell 7): grids 2^p per
axis on the unit cube (cell_center), the test function sampled on the
``VoxelFields.meshgrid`` points in the working precision, the numerical RHS
compared with the continuum RHS in relative L2, and the log-log slope fitted.
The continuum right-hand sides are the reference's sympy ``rhs_analytic``
forms (problem_definition.py:223-224, 322-326, 408-419).
"""
from __future__ import annotations

import numpy as np
import sympy as sp
import sympy.vector as spv

CS = spv.CoordSys3D("CS")


def rd_analytic(u, D=1.0):
    return D * spv.laplacian(u)


def ch_analytic(c, eps=3.0, D=1.0):
    mu = 18 / eps * c * (1 - c) * (1 - 2 * c) - 2 * eps * spv.laplacian(c)
    return spv.divergence(D * c * (1 - c) * spv.gradient(mu))


def ac_analytic(phi, eps=2.0, gab=1.0, M=1.0, force=0.0, curvature=0.01):
    grad = spv.gradient(phi)
    lap = spv.laplacian(phi)
    ng = sp.sqrt(grad.dot(grad))
    curv = ng * spv.divergence(grad / ng)
    n_lap = lap - (1 - curvature) * curv
    pot = 18 / eps * phi * (1 - phi) * (1 - 2 * phi)
    return M * (gab * (n_lap - pot / (2 * eps)) + 3 / eps * phi * (1 - phi) * force)


def grid_points(n, precision):
    """VoxelFields((n,n,n),(1,1,1)).meshgrid() with vf.precision = precision
    (voxelfields.py:84-86,183-194)."""
    sp_ = 1.0 / n
    ax = np.arange(0, n, dtype=precision) * sp_ + sp_ / 2
    return np.meshgrid(ax, ax, ax, indexing="ij")


def convergence(expr, analytic, rhs_numeric, precision="float64", powers=(3, 4, 5, 6, 7)):
    """Returns (dx, errors, slope).  rhs_numeric(array, n) -> numpy array."""
    f = sp.lambdify((CS.x, CS.y, CS.z), expr, "numpy")
    g = sp.lambdify((CS.x, CS.y, CS.z), analytic, "numpy")
    dx, errs = [], []
    for p in powers:
        n = 2 ** p
        X, Y, Z = grid_points(n, precision)
        init = np.broadcast_to(f(X, Y, Z), X.shape)
        comp = rhs_numeric(np.ascontiguousarray(init), n)
        exact = np.broadcast_to(g(X, Y, Z), X.shape)
        errs.append(np.linalg.norm(comp - exact) / np.linalg.norm(exact))
        dx.append(1.0 / n)
    dx, errs = np.array(dx), np.array(errs)
    slope = np.polyfit(np.log(dx), np.log(errs), 1)[0]
    return dx, errs, slope
