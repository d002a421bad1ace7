"""CPU oracle for the evoxels time-stepping hot path (TEST INFRASTRUCTURE ONLY).

This module is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The ``evoxels_amd`` package never imports it and has no CPU
fallback.

It restates, op for op, the eager torch path of evoxels 1.0.0 so that the
HIP kernels can be compared with the reference's own arithmetic on identical
inputs.  Every function cites the reference file:line it follows.  Operation
order, dtype promotion and the quirks of the reference are kept on purpose:

* ``div_dx = 1/tensor(spacing)`` with the spacing cast to the working
  precision first (voxelgrid.py:29-30);
* the k-space tables come from ``torch.fft.fftfreq`` without a dtype, so they
  are float32 even for float64 fields (voxelgrid.py:84-90,110-114), and the
  IMEX prefactor ``dt/(1-dt*sigma)`` therefore stays float32
  (timesteppers.py:77); torch evaluates ``scalar / tensor`` as
  ``reciprocal(tensor) * scalar``;
* boundary ghosts are made by a periodic pad followed by per-axis overwrites
  in the order x, y, z (boundary_conditions.py:33-59).

Parity pins (each checked in tests/test_oracle_*.py):
  pad_bc            reference KAT test_solvers.py:105-126 plus hand-built ghost
                    arrays for every axis/kind pair and edge composition order
  laplace           discrete Fourier eigenvalues (any size) and MMS slopes of
                    test_laplace.py (periodic, zero-flux)
  ch_rhs            MMS slope of test_rhs.py:13-21, single-mode linearisation,
                    telescoping mass, translation equivariance, constant field
  ac_rhs            MMS slope of test_rhs.py:29-37, notebook error array
                    (10-appendix-convergence-testing.ipynb:384) as a soft KAT,
                    tanh equilibrium of test_solvers.py:70-82
  normal_laplace    closed form on a planar profile (equals the 1-D second
                    derivative), bulk guard
  rd_rhs/fe/rk4     closed forms on single Fourier modes
  k tables/imex     float32 fftfreq identity, single-mode closed form
                    u1 = u0 + prefac(k)*(-D*lambda_h(k))*u0, brute-force DFT
  phi1/etd1         Pade branch against expm1(z)/z in float64 (and the
                    formula branch at the switch point), single-mode closed
                    form u1 = u0*(1 + dt*phi_1(dt*sigma_k)*D*lambda_h(k))
  solve loop        u+1 with two fields (test_solvers.py:14-27), frame times
  multiphase_ac_rhs MMS slope of test_rhs.py:61-79 (Neumann x, float64),
                    phase sum of the rhs is zero, simplex projections agree on
                    points already on the simplex
  coupled_rd_rhs    MMS slope of test_rhs.py:43-52 (float64), spatially
                    uniform state gives the ODE kinetics exactly
  rd_sbm_rhs        with mask == 1 it reduces to rd_rhs + bc_flux; MMS slope
                    of the cell_center variant of test_rhs.py:82-97
  pad_staggered_x   slopes of the reference notebook for the staggered_x
                    Laplacian (10-appendix-convergence-testing.ipynb cells 9,
                    11), test_laplace.py:63-76 and the staggered smoothed-
                    boundary test test_rhs.py:82-97
  spectral_step_bc  closed forms on the eigenmodes of each extension:
                    cos/sin(pi m (i+1/2)/n) (cell_center Neumann/Dirichlet)
                    and sin(pi m i/N) (staggered_x Dirichlet)
  ch_energy         no counterpart in the reference (build-defined discrete
                    functional); pinned by a closed form on a single mode
"""
from __future__ import annotations

import math
import warnings
from dataclasses import dataclass
from typing import Callable, Optional, Sequence, Tuple

import numpy as np
import torch

__all__ = [
    "OracleGrid", "normalize_bc", "pad_periodic", "pad_bc", "laplace",
    "to_face", "grad_face", "grad_center", "gradient_norm_squared",
    "normal_laplace", "rd_rhs", "ch_rhs", "ac_rhs", "multiphase_ac_rhs", "coupled_rd_rhs", "rd_sbm_setup",
    "rd_sbm_rhs", "fft_axes", "rfft_axes", "pad_staggered_x", "pad_for", "trim_boundary_nodes", "pad_fft",
    "rfft_k_squared_nonperiodic", "spectral_step_bc",
    "rfft_k_squared", "rd_symbol", "ch_symbol", "ac_symbol", "imex_prefac",
    "fe_step", "rk4_step", "imex_step", "phi1", "etd1_multiplier", "etd1_step", "ch_energy", "average",
    "run_loop",
    "spacing_for",
]

_DT = {"float32": torch.float32, "float64": torch.float64,
       torch.float32: torch.float32, torch.float64: torch.float64}


def spacing_for(shape, domain_size=(1, 1, 1), convention="cell_center"):
    """Grid spacing and origin (voxelfields.py:84-91)."""
    nx, ny, nz = shape
    if convention == "cell_center":
        sp = (domain_size[0] / nx, domain_size[1] / ny, domain_size[2] / nz)
        org = (sp[0] / 2, sp[1] / 2, sp[2] / 2)
    elif convention == "staggered_x":
        sp = (domain_size[0] / (nx - 1), domain_size[1] / ny, domain_size[2] / nz)
        org = (0, sp[1] / 2, sp[2] / 2)
    else:
        raise ValueError("Chosen convention must be cell_center or staggered_x.")
    return sp, org


@dataclass
class OracleGrid:
    """The numbers a VoxelGridTorch carries (voxelgrid.py:21-36,155-186)."""
    shape: Tuple[int, int, int]
    spacing: Tuple[float, float, float]
    precision: str = "float64"
    convention: str = "cell_center"

    def __post_init__(self):
        self.dtype = _DT[self.precision]
        sp = torch.tensor(np.array(self.spacing), dtype=self.dtype)
        # voxelgrid.py:29-30: cast first, then reciprocal (and square)
        self.div_dx = 1 / sp
        self.div_dx2 = 1 / sp ** 2

    @classmethod
    def from_domain(cls, shape, domain_size=(1, 1, 1), precision="float64",
                    convention="cell_center"):
        sp, _ = spacing_for(shape, domain_size, convention)
        return cls(tuple(int(n) for n in shape), sp, precision, convention)

    def field(self, array) -> torch.Tensor:
        """init_scalar_field (voxelgrid.py:126-130,188-189)."""
        return torch.tensor(np.asarray(array), dtype=self.dtype).unsqueeze(0)


# --------------------------------------------------------------------------
# Boundary conditions
# --------------------------------------------------------------------------
def normalize_bc(bc, convention="cell_center"):
    """BC validation and normalisation (problem_definition.py:60-103).

    Returns a tuple of three ``(kind, values)`` pairs.
    """
    if bc is None or bc == "fully_periodic":
        return (("periodic", None),) * 3
    if len(bc) != 3:
        raise ValueError("bc must contain exactly three axis entries ordered as (x, y, z).")
    out = []
    for axis_bc in bc:
        if isinstance(axis_bc, str):
            if axis_bc == "dirichlet":
                raise ValueError("Dirichlet BCs require explicit values ('dirichlet', (left, right)).")
            if axis_bc not in {"periodic", "neumann"}:
                raise ValueError(f"Unsupported BC type: {axis_bc}")
            out.append((axis_bc, None))
            continue
        if len(axis_bc) != 2:
            raise ValueError(
                "Each axis boundary specification must be either a string or "
                "a tuple like ('dirichlet', (left, right)).")
        kind, values = axis_bc
        if kind not in {"periodic", "dirichlet", "neumann"}:
            raise ValueError(f"Unsupported BC type: {kind}")
        if kind == "dirichlet":
            if values is None or len(values) != 2:
                raise ValueError("Dirichlet BCs require two boundary values.")
            if convention == "cell_center":
                warnings.warn("Applying Dirichlet BCs on a cell_center grid "
                              "reduces the spatial order of convergence to 0.5!")
            out.append((kind, tuple(values)))
        else:
            if values is not None:
                raise ValueError(f"{kind} BCs do not accept boundary values.")
            out.append((kind, None))
    return tuple(out)


def pad_periodic(field: torch.Tensor) -> torch.Tensor:
    """Circular ghost layer of width one (voxelgrid.py:194-195)."""
    return torch.nn.functional.pad(field, (1, 1, 1, 1, 1, 1), mode="circular")


def pad_bc(field: torch.Tensor, bc) -> torch.Tensor:
    """Generic cell-centred ghost padding (boundary_conditions.py:33-59).

    Periodic pad first, then for x, y, z in that order: Neumann ghost equals
    the first interior value, Dirichlet ghost equals ``2*v - u1``.  The
    specialised paths (boundary_conditions.py:16-31) produce identical values.
    """
    padded = pad_periodic(field)
    for axis, (kind, values) in enumerate(bc):
        if kind == "periodic":
            continue
        d = axis + 1
        lg = [slice(None)] * 4; lg[d] = 0
        li = [slice(None)] * 4; li[d] = 1
        rg = [slice(None)] * 4; rg[d] = -1
        ri = [slice(None)] * 4; ri[d] = -2
        lg, li, rg, ri = tuple(lg), tuple(li), tuple(rg), tuple(ri)
        if kind == "neumann":
            padded[lg] = padded[li]
            padded[rg] = padded[ri]
        elif kind == "dirichlet":
            padded[lg] = 2.0 * values[0] - padded[li]
            padded[rg] = 2.0 * values[1] - padded[ri]
        else:
            raise ValueError(f"Unsupported BC type: {kind}")
    return padded


def pad_staggered_x(field: torch.Tensor, kind, values=None) -> torch.Tensor:
    """StaggeredXBCs.pad_dirichlet_periodic / pad_zero_flux_periodic
    (boundary_conditions.py:96-121): periodic pad, then the whole x ghost
    planes become the boundary node values (Dirichlet) or the zero-slope
    extrapolation 4/3 u1 - 1/3 u2, low plane first."""
    padded = pad_periodic(field)
    if kind == "dirichlet":
        padded[:, 0, :, :] = values[0]
        padded[:, -1, :, :] = values[1]
    elif kind == "neumann":
        padded[:, 0, :, :] = 4 / 3 * padded[:, 1, :, :] - 1 / 3 * padded[:, 2, :, :]
        padded[:, -1, :, :] = 4 / 3 * padded[:, -2, :, :] - 1 / 3 * padded[:, -3, :, :]
    else:
        raise ValueError(f"Unsupported staggered_x BC type: {kind}")
    return padded


def pad_for(field: torch.Tensor, bc, convention="cell_center") -> torch.Tensor:
    """ODE.pad_bc as initialize_boundary_conditions wires it
    (problem_definition.py:60-115): cell_center grids use the generic ghost
    rules (the specialised paths give identical values); staggered_x grids
    implement only (dirichlet | neumann, periodic, periodic)."""
    if convention == "cell_center":
        return pad_bc(field, bc)
    kinds = tuple(k for k, _ in bc)
    if kinds == ("dirichlet", "periodic", "periodic"):
        return pad_staggered_x(field, "dirichlet", bc[0][1])
    if kinds == ("neumann", "periodic", "periodic"):
        return pad_staggered_x(field, "neumann")
    raise NotImplementedError(f"staggered_x padding for {kinds}")


def trim_boundary_nodes(field, convention="cell_center"):
    """CellCenteredBCs / StaggeredXBCs.trim_boundary_nodes
    (boundary_conditions.py:73-74,141-148)."""
    return field if convention == "cell_center" else field[:, 1:-1, :, :]


def pad_fft(field, x_kind="periodic", convention="cell_center"):
    """pad_fft_* of boundary_conditions.py:61-71,133-144: identity (periodic),
    even (Neumann) or odd (Dirichlet) mirror extension of x; on staggered_x
    grids the odd extension carries the zero boundary nodes."""
    if x_kind == "periodic":
        if convention != "cell_center":
            raise NotImplementedError
        return field
    if convention == "cell_center":
        sign = 1.0 if x_kind == "neumann" else -1.0
        return torch.cat((field, sign * torch.flip(field, [1])), 1)
    if x_kind != "dirichlet":
        raise NotImplementedError
    z = torch.zeros_like(field[:, 0:1])
    return torch.cat((field, z, -torch.flip(field, [1]), z), 1)


def pad_zeros(field):
    """Zero ghost layer (voxelgrid.py:197-198)."""
    return torch.nn.functional.pad(field, (1, 1, 1, 1, 1, 1), mode="constant", value=0)


# --------------------------------------------------------------------------
# Finite-difference stencils (fd_stencils.py)
# --------------------------------------------------------------------------
_ = slice(None)
_i = slice(1, -1)
CENTER = (_, _i, _i, _i)
LEFT = (_, slice(None, -2), _i, _i)
RIGHT = (_, slice(2, None), _i, _i)
BOTTOM = (_, _i, slice(None, -2), _i)
TOP = (_, _i, slice(2, None), _i)
BACK = (_, _i, _i, slice(None, -2))
FRONT = (_, _i, _i, slice(2, None))


def to_face(field, axis):
    """to_{x,y,z}_face (fd_stencils.py:20-30)."""
    d = axis + 1
    hi = [_] * 4; hi[d] = slice(1, None)
    lo = [_] * 4; lo[d] = slice(None, -1)
    return 0.5 * (field[tuple(hi)] + field[tuple(lo)])


def grad_face(field, axis, div_dx):
    """grad_{x,y,z}_face (fd_stencils.py:32-42)."""
    d = axis + 1
    hi = [_] * 4; hi[d] = slice(1, None)
    lo = [_] * 4; lo[d] = slice(None, -1)
    return (field[tuple(hi)] - field[tuple(lo)]) * div_dx[axis]


def grad_center(field, axis, div_dx):
    """grad_{x,y,z}_center (fd_stencils.py:44-54)."""
    r, l = [(RIGHT, LEFT), (TOP, BOTTOM), (FRONT, BACK)][axis]
    return 0.5 * (field[r] - field[l]) * div_dx[axis]


def gradient_norm_squared(field, div_dx):
    """fd_stencils.py:56-60."""
    return grad_center(field, 0, div_dx) ** 2 + \
        grad_center(field, 1, div_dx) ** 2 + \
        grad_center(field, 2, div_dx) ** 2


def laplace(field, div_dx2):
    """7-point Laplacian in the reference's evaluation order (fd_stencils.py:62-75)."""
    return (field[RIGHT] + field[LEFT]) * div_dx2[0] + \
        (field[TOP] + field[BOTTOM]) * div_dx2[1] + \
        (field[FRONT] + field[BACK]) * div_dx2[2] - \
        2 * field[CENTER] * torch.sum(div_dx2)


def normal_laplace(field, div_dx, div_dx2):
    """19-point normal Laplacian (fd_stencils.py:77-103)."""
    gx = grad_center(field, 0, div_dx)
    gy = grad_center(field, 1, div_dx)
    gz = grad_center(field, 2, div_dx)
    f = field
    n_lap = \
        gx ** 2 * (f[RIGHT] - 2 * f[CENTER] + f[LEFT]) * div_dx2[0] + \
        gy ** 2 * (f[TOP] - 2 * f[CENTER] + f[BOTTOM]) * div_dx2[1] + \
        gz ** 2 * (f[FRONT] - 2 * f[CENTER] + f[BACK]) * div_dx2[2] + \
        0.5 * gx * gy * (f[:, 2:, 2:, 1:-1] + f[:, :-2, :-2, 1:-1] -
                         f[:, :-2, 2:, 1:-1] - f[:, 2:, :-2, 1:-1]) * div_dx[0] * div_dx[1] + \
        0.5 * gx * gz * (f[:, 2:, 1:-1, 2:] + f[:, :-2, 1:-1, :-2] -
                         f[:, :-2, 1:-1, 2:] - f[:, 2:, 1:-1, :-2]) * div_dx[0] * div_dx[2] + \
        0.5 * gy * gz * (f[:, 1:-1, 2:, 2:] + f[:, 1:-1, :-2, :-2] -
                         f[:, 1:-1, :-2, 2:] - f[:, 1:-1, 2:, :-2]) * div_dx[1] * div_dx[2]
    norm2 = gradient_norm_squared(field, div_dx)
    bulk = torch.where(norm2 <= 1e-7)
    norm2[bulk] = 1.0
    return n_lap / norm2


# --------------------------------------------------------------------------
# Right-hand sides (problem_definition.py)
# --------------------------------------------------------------------------
def rd_rhs(u, g: OracleGrid, bc, D: float, f: Optional[Callable] = None, t=0.0):
    """ReactionDiffusion.rhs (problem_definition.py:226-229); default f = 0 (:202-203)."""
    lap = laplace(pad_for(u, bc, g.convention), g.div_dx2)
    fv = 0 if f is None else f(t, u, torch)
    return D * lap + fv


def ch_rhs(c, g: OracleGrid, bc, eps: float = 3.0, D: float = 1.0,
           mu_hom: Optional[Callable] = None):
    """CahnHilliard.rhs (problem_definition.py:328-371) with the default
    mu_hom of problem_definition.py:304-305."""
    if mu_hom is None:
        mu_hom = lambda c_: 18 / eps * c_ * (1 - c_) * (1 - 2 * c_)  # noqa: E731
    c = torch.clip(c, 0, 1)
    c_bc = pad_for(c, bc, g.convention)
    lap = laplace(c_bc, g.div_dx2)
    mu = mu_hom(c) - 2 * eps * lap
    mu = pad_for(mu, bc, g.convention)
    div = grad_face(to_face(c_bc, 0) * (1 - to_face(c_bc, 0)) *
                    grad_face(mu, 0, g.div_dx), 0, g.div_dx)[:, :, 1:-1, 1:-1]
    div += grad_face(to_face(c_bc, 1) * (1 - to_face(c_bc, 1)) *
                     grad_face(mu, 1, g.div_dx), 1, g.div_dx)[:, 1:-1, :, 1:-1]
    div += grad_face(to_face(c_bc, 2) * (1 - to_face(c_bc, 2)) *
                     grad_face(mu, 2, g.div_dx), 2, g.div_dx)[:, 1:-1, 1:-1, :]
    return D * div


def ac_rhs(phi, g: OracleGrid, bc, eps: float = 2.0, gab: float = 1.0,
           M: float = 1.0, force: float = 0.0, curvature: float = 0.01,
           potential: Optional[Callable] = None):
    """TwoPhaseAllenCahn.rhs (problem_definition.py:440-447), default
    potential of problem_definition.py:390-391."""
    if potential is None:
        potential = lambda u: 18 / eps * u * (1 - u) * (1 - 2 * u)  # noqa: E731
    phi = torch.clip(phi, 0, 1)
    pot = potential(phi)
    phi_pad = pad_for(phi, bc, g.convention)
    lap = curvature * laplace(phi_pad, g.div_dx2)
    n_lap = (1 - curvature) * normal_laplace(phi_pad, g.div_dx, g.div_dx2)
    df_dphi = gab * (lap + n_lap - pot / 2 / eps) + 3 / eps * phi * (1 - phi) * force
    return M * df_dphi


def multiphase_ac_rhs(phis, g: OracleGrid, bc, eps=3.0, gab=1.0, M=1.0, curvature=1.0, fast=True):
    """MultiPhaseAllenCahn.rhs (problem_definition.py:508-573), 'well'
    potential; ``fast`` selects the sloppy or the Euclidean simplex
    projection (problem_definition.py:508-530)."""
    if fast:
        phis = torch.clip(phis, min=0.0)
        phis = phis / torch.sum(phis, dim=0, keepdim=True)
    else:
        n = phis.shape[0]
        u, _ = torch.sort(phis, dim=0, descending=True)
        cssv = torch.cumsum(u, dim=0) - 1.0
        k = torch.arange(1, n + 1, dtype=phis.dtype).reshape(n, 1, 1, 1)
        cond = (u - cssv / k) > 0
        n_active = torch.clip(torch.sum(cond, dim=0, keepdim=False), min=1)
        theta = torch.take_along_dim(cssv, (n_active - 1)[None, ...], dim=0) / n_active[None, ...]
        phis = torch.clip(phis - theta, min=0.0)
    phi_pad = pad_for(phis, bc, g.convention)
    dfgrad = -curvature * laplace(phi_pad, g.div_dx2)
    dfgrad -= (1 - curvature) * normal_laplace(phi_pad, g.div_dx, g.div_dx2)
    pot_factor = 9 / (2 * eps ** 2)
    sum_sq = torch.sum(phis ** 2, dim=0, keepdim=True)
    dfpot = pot_factor * (3 * phis * (sum_sq - phis ** 2) + phis ** 3 - phis)
    df = dfgrad + dfpot
    return -M * gab * (df - torch.mean(df, dim=0, keepdim=True))


def coupled_rd_rhs(u, g: OracleGrid, bc, D_A=1.0, D_B=0.5, feed=0.055, kill=0.117):
    """CoupledReactionDiffusion.rhs (problem_definition.py:612-633), default
    interaction u[0]*u[1]^2."""
    interaction = u[0] * u[1] ** 2
    lap = laplace(pad_for(u, bc, g.convention), g.div_dx2)
    dc_A = D_A * lap[0] - interaction + feed * (1 - u[0])
    dc_B = D_B * lap[1] + interaction - kill * u[1]
    return torch.stack((dc_A, dc_B), 0)


def rd_sbm_setup(g: OracleGrid, bc, mask=None):
    """ReactionDiffusionSBM.__post_init__ (problem_definition.py:236-262):
    returns (padded mask, norm, bc).  ``mask`` covers the whole grid (on
    staggered_x grids including the boundary nodes)."""
    if mask is None:
        m = pad_periodic(trim_boundary_nodes(g.field(np.ones(g.shape)), g.convention))
        return m, 1.0, bc
    mf = g.field(mask)
    mask_0, mask_1 = mf[:, 0, :, :], mf[:, -1, :, :]
    m = pad_periodic(trim_boundary_nodes(mf, g.convention))
    if bc[0][0] != "periodic":
        m[:, 0, 1:-1, 1:-1] = mask_0
        m[:, -1, 1:-1, 1:-1] = mask_1
    norm = torch.sqrt(gradient_norm_squared(m, g.div_dx))
    m = torch.clip(m, 1e-4, 1)
    if bc[0][0] == "dirichlet":
        v = bc[0][1]
        bc = (("dirichlet", (v[0] * m[:, 0, :, :], v[1] * m[:, -1, :, :])), bc[1], bc[2])
    return m, norm, bc


def rd_sbm_rhs(u, g: OracleGrid, bc, D, m, norm, bc_flux=0.0, f=None):
    """ReactionDiffusionSBM.rhs (problem_definition.py:273-287); ``f(u)`` is
    the source term (zero when None)."""
    z = pad_for(u, bc, g.convention)
    div = grad_face(grad_face(z, 0, g.div_dx) - to_face(z / m, 0) * grad_face(m, 0, g.div_dx),
                    0, g.div_dx)[:, :, 1:-1, 1:-1]
    div += grad_face(grad_face(z, 1, g.div_dx) - to_face(z / m, 1) * grad_face(m, 1, g.div_dx),
                     1, g.div_dx)[:, 1:-1, :, 1:-1]
    div += grad_face(grad_face(z, 2, g.div_dx) - to_face(z / m, 2) * grad_face(m, 2, g.div_dx),
                     2, g.div_dx)[:, 1:-1, 1:-1, :]
    out = D * div + norm * bc_flux
    if f is not None:
        inner = m[:, 1:-1, 1:-1, 1:-1]
        out = out + inner * f(u / inner)
    return out


# --------------------------------------------------------------------------
# k-space tables and symbols (voxelgrid.py:84-124, problem_definition.py)
# --------------------------------------------------------------------------
def fft_axes(shape, spacing):
    """2*pi*fftfreq per axis, float32 by torch's default (voxelgrid.py:84-86)."""
    return tuple(2 * math.pi * torch.fft.fftfreq(n, d) for n, d in zip(shape, spacing))


def rfft_axes(shape, spacing):
    """2*pi*rfftfreq per axis (voxelgrid.py:88-90)."""
    return tuple(2 * math.pi * torch.fft.rfftfreq(n, d) for n, d in zip(shape, spacing))


def rfft_k_squared(shape, spacing):
    """voxelgrid.py:110-114 (float32 table)."""
    a_x, a_y, _ = fft_axes(shape, spacing)
    _, _, a_z = rfft_axes(shape, spacing)
    kx, ky, kz = torch.meshgrid(a_x, a_y, a_z, indexing="ij")
    return kx ** 2 + ky ** 2 + kz ** 2


def rfft_k_squared_nonperiodic(shape, spacing, convention="cell_center"):
    """voxelgrid.py:116-124: the x axis of the mirror extension has 2*Nx
    points (cell_center) or 2*Nx-2 (staggered_x, Nx nodes); float32 table."""
    n = 2 * shape[0] if convention == "cell_center" else 2 * shape[0] - 2
    a_x = 2 * math.pi * torch.fft.fftfreq(n, spacing[0])
    _, a_y, _ = fft_axes(shape, spacing)
    _, _, a_z = rfft_axes(shape, spacing)
    kx, ky, kz = torch.meshgrid(a_x, a_y, a_z, indexing="ij")
    return kx ** 2 + ky ** 2 + kz ** 2


def rd_symbol(k2, D, A=0.25):
    """problem_definition.py:206."""
    return -D * A * k2


def ch_symbol(k2, eps=3.0, D=1.0, A=0.25):
    """problem_definition.py:303."""
    return -2 * eps * D * A * k2 ** 2


def ac_symbol(k2, M=1.0, gab=1.0):
    """problem_definition.py:389."""
    return -M * gab * k2


def imex_prefac(dt, symbol):
    """timesteppers.py:77; ``dt / tensor`` is ``reciprocal(tensor) * dt`` in torch."""
    return dt / (1 - dt * symbol)


# --------------------------------------------------------------------------
# Time steppers (timesteppers.py)
# --------------------------------------------------------------------------
def fe_step(u, rhs_fn, dt, t=0.0):
    """ForwardEuler.step (timesteppers.py:42-43)."""
    return u + dt * rhs_fn(t, u)


def rk4_step(u, rhs_fn, dt, t=0.0):
    """RungeKutta4.step (timesteppers.py:56-61)."""
    k1 = rhs_fn(t, u)
    k2 = rhs_fn(t + 0.5 * dt, u + 0.5 * dt * k1)
    k3 = rhs_fn(t + 0.5 * dt, u + 0.5 * dt * k2)
    k4 = rhs_fn(t + dt, u + dt * k3)
    return u + (dt / 6) * (k1 + 2 * k2 + 2 * k3 + k4)


_PADE_N = (1, 1 / 26, 5 / 156, 1 / 858, 1 / 5720, 1 / 205920, 1 / 8648640)
_PADE_D = (1, -6 / 13, 5 / 52, -5 / 429, 1 / 1144, -1 / 25740, 1 / 1235520)


def phi1(z):
    """ExponentialEuler.phi1 (timesteppers.py:155-169): (exp(z)-1)/z, with the
    degree-6 Pade approximant (timesteppers.py:171-190, Hochbruck, Lubich,
    Selhofer 1998) where |z| < 0.5."""
    phi = (torch.exp(z) - 1) / z
    idx = torch.abs(z) < 0.5
    zz = z[idx]
    num, den = _PADE_N[6], _PADE_D[6]
    for k in range(5, -1, -1):
        num = num * zz + _PADE_N[k]
        den = den * zz + _PADE_D[k]
    phi[idx] = num / den
    return phi


def etd1_multiplier(dt, symbol):
    """dt * phi_1(dt * sigma) of ExponentialEuler (timesteppers.py:151,197)."""
    return dt * phi1(dt * symbol)


def etd1_step(u, rhs_fn, mult, t=0.0):
    """ExponentialEuler.step for a periodic problem (timesteppers.py:195-199)."""
    dc = rhs_fn(t, u)
    dc_fft = mult * torch.fft.rfftn(dc, s=dc.shape)
    update = torch.fft.irfftn(dc_fft, s=dc.shape)[:, :u.shape[1]]
    return u + update


def imex_step(u, rhs_fn, prefac, t=0.0):
    """PseudoSpectralIMEX.step for a periodic problem (timesteppers.py:85-89,
    voxelgrid.py:203-207; pad_fft_periodic is the identity,
    boundary_conditions.py:61-63).  The 4-D ``s=`` transform of the reference
    also runs over the channel axis; prefac does not depend on the channel
    frequency, so that axis cancels."""
    dc = rhs_fn(t, u)
    dc_fft = prefac * torch.fft.rfftn(dc, s=dc.shape)
    update = torch.fft.irfftn(dc_fft, s=dc.shape)[:, :u.shape[1]]
    return u + update


def spectral_step_bc(u, rhs_fn, mult, x_kind="periodic", convention="cell_center", t=0.0):
    """PseudoSpectralIMEX.step / ExponentialEuler.step with a non-periodic x
    axis (timesteppers.py:85-89,195-199): the RHS is mirror-extended in x,
    multiplied in k-space (``mult`` from rfft_k_squared_nonperiodic) and the
    first Nx planes of the inverse transform are the update."""
    dc = pad_fft(rhs_fn(t, u), x_kind, convention)
    dc_fft = mult * torch.fft.rfftn(dc, s=dc.shape)
    update = torch.fft.irfftn(dc_fft, s=dc.shape)[:, :u.shape[1]]
    return u + update


# --------------------------------------------------------------------------
# Reductions and the solver loop
# --------------------------------------------------------------------------
def average(field):
    """VoxelGrid.average for cell_center (voxelgrid.py:137-141)."""
    return torch.mean(field, (1, 2, 3))


def ch_energy(c, g: OracleGrid, bc, eps=3.0):
    """Discrete Cahn-Hilliard free energy, summed in float64.

    The reference defines no energy function; the functional is
    ``F = integral eps*|grad c|^2 + 9/eps*c^2*(1-c)^2 dV`` (consistent with
    mu = 18/eps*c(1-c)(1-2c) - 2*eps*lap c, problem_definition.py:305,353).
    Gradients are the centred differences of fd_stencils.py:44-60 on the
    BC-padded field; dV is the voxel volume.
    """
    cb = pad_bc(c, bc).to(torch.float64)
    dd = g.div_dx.to(torch.float64)
    g2 = gradient_norm_squared(cb, dd)
    cc = c.to(torch.float64)
    dens = eps * g2 + 9.0 / eps * cc ** 2 * (1 - cc) ** 2
    dv = float(np.prod(np.asarray(g.spacing, dtype=np.float64)))
    return torch.sum(dens, (1, 2, 3)) * dv


def run_loop(u, step, time_increment, frames, max_iters, on_output=None):
    """TimeDependentSolver._run_loop (solvers.py:192-208): output before the
    step whenever ``i % (max_iters // frames) == 0``, plus a final output."""
    n_out = max_iters // frames
    frame = 0
    for i in range(max_iters):
        time = i * time_increment
        if i % n_out == 0:
            if on_output is not None:
                on_output(u, frame, time)
            frame += 1
        u = step(time, u)
    time = max_iters * time_increment
    if on_output is not None:
        on_output(u, frame, time)
    return u
