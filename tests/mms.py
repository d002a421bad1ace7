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


def staggered_grid_points(n, precision):
    """VoxelFields((n+1,n,n),(1,1,1),convention='staggered_x').meshgrid():
    x nodes 0, 1/n, ..., 1 and cell-centred y, z (voxelfields.py:87-89,183-194)."""
    sp_ = 1.0 / n
    ax = np.arange(0, n + 1, dtype=precision) * sp_
    ay = np.arange(0, n, dtype=precision) * sp_ + sp_ / 2
    return np.meshgrid(ax, ay, ay, indexing="ij")


def convergence(expr, analytic, rhs_numeric, precision="float64", powers=(3, 4, 5, 6, 7),
                convention="cell_center"):
    """Returns (dx, errors, slope).  rhs_numeric(array, n) -> numpy array; on
    staggered_x grids the array holds the boundary nodes and the result the
    interior, which is compared with the interior of the analytic RHS."""
    f = sp.lambdify((CS.x, CS.y, CS.z), expr, "numpy")
    g = sp.lambdify((CS.x, CS.y, CS.z), analytic, "numpy")
    dx, errs = [], []
    for p in powers:
        n = 2 ** p
        X, Y, Z = grid_points(n, precision) if convention == "cell_center" else staggered_grid_points(n, precision)
        init = np.broadcast_to(f(X, Y, Z), X.shape)
        comp = rhs_numeric(np.ascontiguousarray(init), n)
        exact = np.broadcast_to(g(X, Y, Z), X.shape)
        if convention != "cell_center":
            exact = exact[1:-1]
        errs.append(np.linalg.norm(comp - exact) / np.linalg.norm(exact))
        dx.append(1.0 / n)
    dx, errs = np.array(dx), np.array(errs)
    slope = np.polyfit(np.log(dx), np.log(errs), 1)[0]
    return dx, errs, slope


def coupled_rd_analytic(us, D_A=1.0, D_B=0.5, feed=0.055, kill=0.117):
    """CoupledReactionDiffusion.rhs_analytic (problem_definition.py:605-610)."""
    inter = us[0] * us[1] ** 2
    return (D_A * spv.laplacian(us[0]) - inter + feed * (1 - us[0]),
            D_B * spv.laplacian(us[1]) + inter - kill * us[1])


def multiphase_ac_analytic(phis, eps=3.0, gab=1.0, M=1.0, curvature=1.0):
    """MultiPhaseAllenCahn.rhs_analytic (problem_definition.py:489-506), 'well'."""
    pot_factor = 9 / (2 * eps ** 2)
    sum_sq = sum(p ** 2 for p in phis)
    dfs = []
    for phi in phis:
        lap = spv.laplacian(phi)
        grad_term = -lap
        if curvature != 1:
            grad = spv.gradient(phi)
            ng = sp.sqrt(grad.dot(grad))
            grad_term = -curvature * lap - (1 - curvature) * (lap - ng * spv.divergence(grad / ng))
        dfs.append(grad_term + pot_factor * (3 * phi * (sum_sq - phi ** 2) + phi ** 3 - phi))
    mean = sum(dfs) / len(dfs)
    return tuple(-M * gab * (df - mean) for df in dfs)


def sbm_analytic(u, mask, D=1.0, bc_flux=0.0):
    """ReactionDiffusionSBM.rhs_analytic (problem_definition.py:264-271), f = 0."""
    grad_m = spv.gradient(mask)
    return spv.divergence(D * (spv.gradient(u) - u / mask * grad_m)) + sp.sqrt(grad_m.dot(grad_m)) * bc_flux


def convergence_channels(exprs, analytics, rhs_numeric, precision="float64", powers=(3, 4, 5, 6, 7),
                         mask=None, convention="cell_center"):
    """Multi-channel form of ``convergence`` (utils.py:58-126): the channels
    are stacked, ``rhs_numeric(array (C,n,n,n), n, mask_array)`` returns the
    (C,n,n,n) RHS; returns (dx, errors (C, P), slopes (C,))."""
    fs = [sp.lambdify((CS.x, CS.y, CS.z), e, "numpy") for e in exprs]
    gs = [sp.lambdify((CS.x, CS.y, CS.z), a, "numpy") for a in analytics]
    fm = sp.lambdify((CS.x, CS.y, CS.z), mask, "numpy") if mask is not None else None
    dx, errs = [], []
    for p in powers:
        n = 2 ** p
        X, Y, Z = grid_points(n, precision) if convention == "cell_center" else staggered_grid_points(n, precision)
        init = np.stack([np.broadcast_to(f(X, Y, Z), X.shape) for f in fs])
        m = np.broadcast_to(fm(X, Y, Z), X.shape) if fm is not None else None
        comp = rhs_numeric(np.ascontiguousarray(init), n, m)
        exact = [np.broadcast_to(g(X, Y, Z), X.shape) for g in gs]
        if convention != "cell_center":
            exact = [e[1:-1] for e in exact]
        errs.append([np.linalg.norm(comp[j] - exact[j]) / np.linalg.norm(exact[j]) for j in range(len(exact))])
        dx.append(1.0 / n)
    dx, errs = np.array(dx), np.array(errs).T
    slopes = np.array([np.polyfit(np.log(dx), np.log(e), 1)[0] for e in errs])
    return dx, errs, slopes
