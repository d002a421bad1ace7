"""Problem definitions for the 'hip' backend (mirror of evoxels.problem_definition).

Constructor arguments, defaults, boundary-condition validation and warnings
follow evoxels/problem_definition.py:13-447.  With the default closures
(``f``, ``mu_hom``, ``potential`` left as None) ``rhs`` runs one fused HIP
kernel that evaluates the whole right-hand side with in-kernel ghost cells;
time steppers can additionally fuse the state update into that kernel
(``launch``).  A user-supplied closure is evaluated as torch code on the
device inside a generic composition of HIP operators (pad, Laplacian,
normal Laplacian), the same composition as the reference.
"""
from __future__ import annotations

import ctypes as C
import warnings
from abc import ABC, abstractmethod
from dataclasses import dataclass, field
from typing import Any, Callable

import numpy as np
import torch

from . import _lib
from .voxelgrid import VoxelGridHip

try:  # analytic right-hand sides for convergence tests (problem_definition.py:223-224,322-326,408-419)
    import sympy as sp
    import sympy.vector as spv
except ImportError:  # pragma: no cover
    sp = spv = None


class ODE(ABC):
    """problem_definition.py:13-120."""

    @property
    @abstractmethod
    def order(self) -> int:
        ...

    @abstractmethod
    def rhs_analytic(self, t, u):
        ...

    @abstractmethod
    def rhs(self, t, u):
        ...

    def pad_bc(self, u):
        return self._pad_bc(u)

    def initialize_boundary_conditions(self):
        """BC normalisation and validation (problem_definition.py:60-115).  All
        specialised paddings of the reference are one HIP kernel here."""
        bc = getattr(self, "bc", None)
        if bc is None or bc == "fully_periodic":
            self.bc = (("periodic", None), ("periodic", None), ("periodic", None))
            if self.vg.convention == "staggered_x":
                raise NotImplementedError("a staggered_x grid carries boundary nodes in x; a fully periodic "
                                          "problem belongs on a cell_center grid (boundary_conditions.py:89-94)")
            self._pad_bc = lambda f: self.vg.bc.pad_bc(f, self.bc)
            return
        if len(bc) != 3:
            raise ValueError("bc must contain exactly three axis entries ordered as (x, y, z).")
        normalized = []
        for axis_bc in bc:
            if isinstance(axis_bc, str):
                if axis_bc == "dirichlet":
                    raise ValueError("Dirichlet BCs require explicit values ('dirichlet', (left, right)).")
                if axis_bc not in {"periodic", "neumann"}:
                    raise ValueError(f"Unsupported BC type: {axis_bc}")
                normalized.append((axis_bc, None))
                continue
            if len(axis_bc) != 2:
                raise ValueError("Each axis boundary specification must be either a string or "
                                 "a tuple like ('dirichlet', (left, right)).")
            kind, values = axis_bc
            if kind not in {"periodic", "dirichlet", "neumann"}:
                raise ValueError(f"Unsupported BC type: {kind}")
            if kind == "dirichlet":
                if values is None or len(values) != 2:
                    raise ValueError("Dirichlet BCs require two boundary values.")
                if self.vg.convention == "cell_center":
                    warnings.warn("Applying Dirichlet BCs on a cell_center grid "
                                  "reduces the spatial order of convergence to 0.5!")
                normalized.append((kind, tuple(values)))
            else:
                if values is not None:
                    raise ValueError(f"{kind} BCs do not accept boundary values.")
                normalized.append((kind, None))
        self.bc = tuple(normalized)
        if self.vg.convention == "staggered_x":
            # the reference's specialised paddings (problem_definition.py:105-115);
            # any other combination reaches StaggeredXBCs.pad_bc, which raises
            if self.bc_type == ("dirichlet", "periodic", "periodic"):
                self._pad_bc = lambda f: self.vg.bc.pad_dirichlet_periodic(f, self.bc[0][1][0], self.bc[0][1][1])
            elif self.bc_type == ("neumann", "periodic", "periodic"):
                self._pad_bc = self.vg.bc.pad_zero_flux_periodic
            elif self.bc_type == ("periodic", "periodic", "periodic"):
                raise NotImplementedError("a staggered_x grid carries boundary nodes in x; a fully periodic "
                                          "problem belongs on a cell_center grid (boundary_conditions.py:89-94)")
            else:
                self._pad_bc = lambda f: self.vg.bc.pad_bc(f, self.bc)
            return
        self._pad_bc = lambda f: self.vg.bc.pad_bc(f, self.bc)

    @property
    def bc_type(self):
        return tuple(axis_bc[0] for axis_bc in self.bc)

    def _kernel_bcs(self) -> bool:
        """Whether the fused kernels implement this problem's ghost rules:
        every cell_center combination; on staggered_x the two x paddings with
        scalar node values and periodic y, z."""
        if self.vg.convention == "cell_center":
            return True
        if self.bc_type == ("neumann", "periodic", "periodic"):
            return True
        return self.bc_type == ("dirichlet", "periodic", "periodic") and \
            not any(isinstance(v, torch.Tensor) for v in self.bc[0][1])

    # ---- fused-kernel protocol used by the time steppers -----------------
    fused = False

    def _own_rhs(self, cls) -> bool:
        """True while ``rhs`` is ``cls.rhs`` itself: neither replaced on the
        instance (the manufactured-solution driver assigns ``problem.rhs``,
        utils.py:285) nor overridden by a subclass.  A fused step evaluates the
        built-in right-hand side in its kernel, so it must not run when the
        user's ``rhs`` is another function."""
        return "rhs" not in vars(self) and type(self).rhs is cls.rhs

    def launch(self, u, out, epilogue):  # pragma: no cover - overridden
        raise NotImplementedError

    def _check_state(self, u):
        if not isinstance(u, torch.Tensor) or u.device.type != "cuda":
            raise ValueError("state must be a torch tensor on the ROCm device")
        # the state of a staggered_x grid excludes the two boundary nodes in x
        nx = self.vg.shape[0] - (2 if self.vg.convention == "staggered_x" else 0)
        if u.dim() != 4 or tuple(u.shape[1:]) != (nx,) + tuple(self.vg.shape[1:]):
            raise ValueError(f"state must have shape (C, {nx}, {self.vg.shape[1]}, "
                             f"{self.vg.shape[2]}), got {tuple(u.shape)}")
        return u.contiguous()

    def _grid(self, u):
        return self.vg._grid_desc(u.shape)


class SemiLinearODE(ODE):
    """problem_definition.py:123-180."""

    @property
    @abstractmethod
    def fourier_symbol(self):
        ...

    def verify_fft_bc_config(self):
        x_bc, _, _ = self.bc_type
        nonperiodic_axes = tuple(axis for axis, kind in zip(("x", "y", "z"), self.bc_type) if kind != "periodic")
        if len(nonperiodic_axes) > 1:
            raise ValueError("FFT-based timesteppers currently support at most one non-periodic axis, "
                             f"got {self.bc_type}.")
        if len(nonperiodic_axes) == 1 and nonperiodic_axes[0] != "x":
            raise NotImplementedError(
                "FFT-based timesteppers currently only implement the single non-periodic axis "
                f"case for x; got {self.bc_type}. Axis permutation is not implemented yet.")
        if x_bc == "periodic":
            self._pad_fft_bc = self.vg.bc.pad_fft_periodic
        elif x_bc == "dirichlet":
            self._pad_fft_bc = self.vg.bc.pad_fft_dirichlet_periodic
        elif x_bc == "neumann":
            self._pad_fft_bc = self.vg.bc.pad_fft_zero_flux_periodic
        else:
            raise ValueError("FFT-based timesteppers only support periodic, dirichlet, or "
                             f"neumann boundary conditions in x, got {x_bc}.")

    def k_squared(self):
        if self.bc_type[0] in {"dirichlet", "neumann"}:
            return self.vg.rfft_k_squared_nonperiodic()
        return self.vg.rfft_k_squared()

    def pad_fft_bc(self, u):
        return self._pad_fft_bc(u)


def _default(fn):
    fn._evx_default = True
    return fn


def _is_default(fn) -> bool:
    return getattr(fn, "_evx_default", False)


@dataclass
class ReactionDiffusion(SemiLinearODE):
    """problem_definition.py:191-229."""
    vg: VoxelGridHip
    D: float
    f: Callable | None = None
    A: float = 0.25
    bc: tuple = ("periodic", "periodic", "periodic")
    _fourier_symbol: Any = field(init=False, repr=False, default=None)

    def __post_init__(self):
        if self.f is None:
            self.f = _default(lambda c=None, t=None, lib=None: 0)
        self.initialize_boundary_conditions()

    @property
    def _kernel_ok(self):
        return _is_default(self.f) and self._kernel_bcs()

    @property
    def fused(self):
        return self._kernel_ok and self._own_rhs(ReactionDiffusion)

    @property
    def order(self):
        return 2

    @property
    def fourier_symbol(self):
        if self._fourier_symbol is None:
            self._fourier_symbol = -self.D * self.A * self.k_squared()
        return self._fourier_symbol

    def _eval_f(self, t, c, lib):
        try:
            return self.f(t, c, lib)
        except TypeError:
            return self.f(t, c)

    def rhs_analytic(self, t, u):
        return self.D * spv.laplacian(u) + self._eval_f(t, u, sp)

    def launch(self, u, out, epilogue):
        g = self._grid(u)
        b = _lib.make_bc(self.bc, self.vg.convention)
        _lib.call("evx_rd_rhs", _lib.dtype_code(u.dtype), _lib.ptr(u), _lib.ptr(out), C.byref(g), C.byref(b),
                  float(self.D), C.byref(epilogue), _lib.stream_handle(u.device))

    def rhs(self, t, u):
        u = self._check_state(u)
        if self._kernel_ok:
            out = torch.empty_like(u)
            self.launch(u, out, _lib.make_epilogue())
            return out
        laplace = self.vg.laplace(self.pad_bc(u))
        return self.D * laplace + self._eval_f(t, u, self.vg.lib)


@dataclass
class CahnHilliard(SemiLinearODE):
    """problem_definition.py:290-371."""
    vg: VoxelGridHip
    eps: float = 3.0
    D: float = 1.0
    mu_hom: Callable | None = None
    A: float = 0.25
    bc: tuple = ("periodic", "periodic", "periodic")
    _fourier_symbol: Any = field(init=False, repr=False, default=None)

    def __post_init__(self):
        self.initialize_boundary_conditions()
        if self.mu_hom is None:
            self.mu_hom = _default(lambda c, lib=None: 18 / self.eps * c * (1 - c) * (1 - 2 * c))

    @property
    def _kernel_ok(self):
        return _is_default(self.mu_hom) and self.vg.convention == "cell_center"

    @property
    def fused(self):
        return self._kernel_ok and self._own_rhs(CahnHilliard)

    @property
    def order(self):
        return 2

    @property
    def fourier_symbol(self):
        if self._fourier_symbol is None:
            self._fourier_symbol = -2 * self.eps * self.D * self.A * self.k_squared() ** 2
        return self._fourier_symbol

    def _eval_mu(self, c, lib):
        try:
            return self.mu_hom(c, lib)
        except TypeError:
            return self.mu_hom(c)

    def rhs_analytic(self, t, c):
        mu = self._eval_mu(c, sp) - 2 * self.eps * spv.laplacian(c)
        fluxes = self.D * c * (1 - c) * spv.gradient(mu)
        return spv.divergence(fluxes)

    def launch(self, u, out, epilogue):
        g = self._grid(u)
        b = _lib.make_bc(self.bc, self.vg.convention)
        _lib.call("evx_ch_rhs", _lib.dtype_code(u.dtype), _lib.ptr(u), _lib.ptr(out), C.byref(g), C.byref(b),
                  float(self.eps), float(self.D), C.byref(epilogue), _lib.stream_handle(u.device))

    def rhs(self, t, c):
        c = self._check_state(c)
        if self._kernel_ok:
            out = torch.empty_like(c)
            self.launch(c, out, _lib.make_epilogue())
            return out
        vg = self.vg
        c = torch.clip(c, 0, 1)
        c_bc = self.pad_bc(c)
        laplace = vg.laplace(c_bc)
        mu = self._eval_mu(c, vg.lib) - 2 * self.eps * laplace
        mu = self.pad_bc(mu)
        div = vg.grad_x_face(vg.to_x_face(c_bc) * (1 - vg.to_x_face(c_bc)) * vg.grad_x_face(mu))[:, :, 1:-1, 1:-1]
        div += vg.grad_y_face(vg.to_y_face(c_bc) * (1 - vg.to_y_face(c_bc)) * vg.grad_y_face(mu))[:, 1:-1, :, 1:-1]
        div += vg.grad_z_face(vg.to_z_face(c_bc) * (1 - vg.to_z_face(c_bc)) * vg.grad_z_face(mu))[:, 1:-1, 1:-1, :]
        return self.D * div


@dataclass
class TwoPhaseAllenCahn(SemiLinearODE):
    """problem_definition.py:374-447."""
    vg: VoxelGridHip
    eps: float = 2.0
    gab: float = 1.0
    M: float = 1.0
    force: float = 0.0
    curvature: float = 0.01
    potential: Callable | None = None
    bc: tuple = ("neumann", "neumann", "neumann")
    _fourier_symbol: Any = field(init=False, repr=False, default=None)

    def __post_init__(self):
        self.initialize_boundary_conditions()
        if self.potential is None:
            self.potential = _default(lambda u, lib=None: 18 / self.eps * u * (1 - u) * (1 - 2 * u))

    @property
    def _kernel_ok(self):
        return _is_default(self.potential) and self._kernel_bcs()

    @property
    def fused(self):
        return self._kernel_ok and self._own_rhs(TwoPhaseAllenCahn)

    @property
    def order(self):
        return 2

    @property
    def fourier_symbol(self):
        if self._fourier_symbol is None:
            self._fourier_symbol = -self.M * self.gab * self.k_squared()
        return self._fourier_symbol

    def _eval_potential(self, phi, lib):
        try:
            return self.potential(phi, lib)
        except TypeError:
            return self.potential(phi)

    def rhs_analytic(self, t, phi):
        grad = spv.gradient(phi)
        laplace = spv.laplacian(phi)
        norm_grad = sp.sqrt(grad.dot(grad))
        unit_normal = grad / norm_grad
        curv = norm_grad * spv.divergence(unit_normal)
        n_laplace = laplace - (1 - self.curvature) * curv
        df_dphi = self.gab * (n_laplace - self._eval_potential(phi, sp) / (2 * self.eps)) \
            + 3 / self.eps * phi * (1 - phi) * self.force
        return self.M * df_dphi

    def launch(self, u, out, epilogue):
        g = self._grid(u)
        b = _lib.make_bc(self.bc, self.vg.convention)
        _lib.call("evx_ac_rhs", _lib.dtype_code(u.dtype), _lib.ptr(u), _lib.ptr(out), C.byref(g), C.byref(b),
                  float(self.eps), float(self.gab), float(self.M), float(self.force), float(self.curvature),
                  C.byref(epilogue), _lib.stream_handle(u.device))

    def rhs(self, t, phi):
        phi = self._check_state(phi)
        if self._kernel_ok:
            out = torch.empty_like(phi)
            self.launch(phi, out, _lib.make_epilogue())
            return out
        vg = self.vg
        phi = torch.clip(phi, 0, 1)
        potential = self._eval_potential(phi, vg.lib)
        phi_pad = self.pad_bc(phi)
        laplace = self.curvature * vg.laplace(phi_pad)
        n_laplace = (1 - self.curvature) * vg.normal_laplace(phi_pad)
        df_dphi = self.gab * (laplace + n_laplace - potential / 2 / self.eps) \
            + 3 / self.eps * phi * (1 - phi) * self.force
        return self.M * df_dphi


@dataclass
class MultiPhaseAllenCahn(SemiLinearODE):
    """problem_definition.py:450-573: N phase fractions in the channels,
    projected onto the Gibbs simplex, double-well potential, curvature-
    controlled gradient term through the HIP Laplacian and normal Laplacian
    on padded phases."""
    vg: VoxelGridHip
    eps: float = 3.0
    gab: float = 1.0
    M: float = 1.0
    force: float = 0.0
    curvature: float = 1.0
    potential: str = "well"
    fast: bool = True
    bc: tuple = ("periodic", "periodic", "periodic")
    _fourier_symbol: Any = field(init=False, repr=False, default=None)

    def __post_init__(self):
        self.initialize_boundary_conditions()
        self._fourier_symbol = None
        if self.potential == "well":
            self.pot_factor = 9 / (2 * self.eps ** 2)
            self.calc_potential_derivatives = self._calc_well_derivatives
        else:
            raise ValueError(f"Unknown potential type: {self.potential}")
        self.project_to_simplex = (self._sloppy_simplex_projection if self.fast
                                   else self._euclidean_simplex_projection)

    @property
    def order(self):
        return 2

    @property
    def fourier_symbol(self):
        if self._fourier_symbol is None:
            self._fourier_symbol = -self.M * self.gab * self.k_squared()
        return self._fourier_symbol

    def rhs_analytic(self, t, phis):
        sum_phi_squared = sum(phi ** 2 for phi in phis)
        df_dphi = []
        for phi in phis:
            grad = spv.gradient(phi)
            laplace = spv.laplacian(phi)
            norm_grad = sp.sqrt(grad.dot(grad))
            curv = norm_grad * spv.divergence(grad / norm_grad)
            grad_term = -self.curvature * laplace - (1 - self.curvature) * (laplace - curv)
            pot_term = self.pot_factor * (3 * phi * (sum_phi_squared - phi ** 2) + phi ** 3 - phi)
            df_dphi.append(grad_term + pot_term)
        mean_df = sum(df_dphi) / len(df_dphi)
        return tuple(-self.M * self.gab * (df - mean_df) for df in df_dphi)

    def _sloppy_simplex_projection(self, phis):
        phis = torch.clip(phis, min=0.0)
        return phis / torch.sum(phis, dim=0, keepdim=True)

    def _euclidean_simplex_projection(self, phis):
        """Euclidean projection onto {phi >= 0, sum phi = 1} per voxel."""
        n = phis.shape[0]
        u = self.vg.sort(phis, dim=0, descending=True)
        cssv = self.vg.cumsum(u, dim=0) - 1.0
        k = self.vg.arange(1, n + 1).reshape(n, 1, 1, 1)
        cond = (u - cssv / k) > 0
        n_active = torch.clip(torch.sum(cond, dim=0, keepdim=False), min=1)
        theta = self.vg.take_along_dim(cssv, (n_active - 1)[None, ...], dim=0) / n_active[None, ...]
        return torch.clip(phis - theta, min=0.0)

    def _calc_well_derivatives(self, phis):
        sum_phi_squared = torch.sum(phis ** 2, dim=0, keepdim=True)
        df_dphi = 3 * phis * (sum_phi_squared - phis ** 2) + phis ** 3 - phis
        return self.pot_factor * df_dphi

    def _calc_obstacle_derivatives(self, phis):
        """problem_definition.py:536-538 (the reference keeps the obstacle
        potential unselectable; kept for API parity)."""
        return -self.pot_factor * phis

    def _kernel_path(self, n_phases):
        """The HIP kernels serve the built-in projections and potential for up
        to eight phases; replaced methods take the tensor path."""
        proj = getattr(self.project_to_simplex, "__func__", None)
        pot = getattr(self.calc_potential_derivatives, "__func__", None)
        return n_phases <= 8 and pot is MultiPhaseAllenCahn._calc_well_derivatives and \
            proj in (MultiPhaseAllenCahn._sloppy_simplex_projection, MultiPhaseAllenCahn._euclidean_simplex_projection)

    def rhs(self, t, phis):
        phis = self._check_state(phis)
        n = phis.shape[0]
        if self._kernel_path(n):
            dt_code = _lib.dtype_code(phis.dtype)
            stream = _lib.stream_handle(phis.device)
            proj = torch.empty_like(phis)
            euclid = getattr(self.project_to_simplex, "__func__", None) is \
                MultiPhaseAllenCahn._euclidean_simplex_projection
            _lib.call("evx_simplex_project", dt_code, _lib.ptr(phis), _lib.ptr(proj), n, phis[0].numel(),
                      int(euclid), stream)
            padded = self.pad_bc(proj).contiguous()
            out = torch.empty_like(phis)
            g = self._grid(proj)
            _lib.call("evx_multiphase_ac_rhs_padded", dt_code, _lib.ptr(proj), _lib.ptr(padded), _lib.ptr(out),
                      C.byref(g), float(self.eps), float(self.gab), float(self.M), float(self.curvature), stream)
            return out
        phis = self.project_to_simplex(phis)
        phi_pad = self.pad_bc(phis)
        dfgrad_dphi = -self.curvature * self.vg.laplace(phi_pad)
        dfgrad_dphi -= (1 - self.curvature) * self.vg.normal_laplace(phi_pad)
        df_dphi = dfgrad_dphi + self.calc_potential_derivatives(phis)
        return -self.M * self.gab * (df_dphi - torch.mean(df_dphi, dim=0, keepdim=True))


@dataclass
class CoupledReactionDiffusion(SemiLinearODE):
    """problem_definition.py:576-633: two species A = u[0], B = u[1] with
    Gray-Scott type kinetics.  The default interaction u[0]*u[1]^2 runs as one
    fused HIP kernel (both species per block, Euler/RK4 in its epilogue); a
    user interaction takes the generic path (HIP pad and Laplacian, device
    tensor algebra)."""
    vg: VoxelGridHip
    D_A: float = 1.0
    D_B: float = 0.5
    feed: float = 0.055
    kill: float = 0.117
    interaction: Callable | None = None
    _fourier_symbol: Any = field(init=False, repr=False, default=None)

    def __post_init__(self):
        self.initialize_boundary_conditions()
        self._fourier_symbol = None
        if self.interaction is None:
            self.interaction = _default(lambda u, lib=None: u[0] * u[1] ** 2)

    @property
    def order(self):
        return 2

    @property
    def fourier_symbol(self):
        if self._fourier_symbol is None:
            self._fourier_symbol = -max(self.D_A, self.D_B) * self.k_squared()
        return self._fourier_symbol

    @property
    def _kernel_ok(self):
        return _is_default(self.interaction) and self._kernel_bcs()

    @property
    def fused(self):
        return self._kernel_ok and self._own_rhs(CoupledReactionDiffusion)

    def _eval_interaction(self, u, lib):
        try:
            return self.interaction(u, lib)
        except TypeError:
            return self.interaction(u)

    def launch(self, u, out, epilogue):
        if u.shape[0] != 2:
            raise ValueError(f"CoupledReactionDiffusion needs two channels (A, B), got {u.shape[0]}")
        g = self._grid(u)
        b = _lib.make_bc(self.bc, self.vg.convention)
        _lib.call("evx_coupled_rd_rhs", _lib.dtype_code(u.dtype), _lib.ptr(u), _lib.ptr(out), C.byref(g),
                  C.byref(b), float(self.D_A), float(self.D_B), float(self.feed), float(self.kill),
                  C.byref(epilogue), _lib.stream_handle(u.device))

    def rhs_analytic(self, t, u):
        interaction = self._eval_interaction(u, sp)
        dc_A = self.D_A * spv.laplacian(u[0]) - interaction + self.feed * (1 - u[0])
        dc_B = self.D_B * spv.laplacian(u[1]) + interaction - self.kill * u[1]
        return (dc_A, dc_B)

    def rhs(self, t, u):
        u = self._check_state(u)
        if u.shape[0] != 2:
            raise ValueError(f"CoupledReactionDiffusion needs two channels (A, B), got {u.shape[0]}")
        if self._kernel_ok:
            out = torch.empty_like(u)
            self.launch(u, out, _lib.make_epilogue())
            return out
        interaction = self._eval_interaction(u, torch)
        laplace = self.vg.laplace(self.pad_bc(u))
        dc_A = self.D_A * laplace[0] - interaction + self.feed * (1 - u[0])
        dc_B = self.D_B * laplace[1] + interaction - self.kill * u[1]
        return torch.stack((dc_A, dc_B), 0)


class SmoothedBoundaryODE(ODE):
    """problem_definition.py:183-188."""

    @property
    @abstractmethod
    def mask(self):
        ...


@dataclass
class ReactionDiffusionSBM(ReactionDiffusion, SmoothedBoundaryODE):
    """problem_definition.py:231-287: reaction-diffusion on a smoothed
    boundary (diffuse domain) given by ``mask``; ``bc_flux`` is the flux
    imposed through the diffuse boundary.  cell_center convention."""
    mask: Any | None = None
    bc_flux: Callable | float = 0.0

    def __post_init__(self):
        super().__post_init__()
        vg = self.vg
        if self.mask is None:
            m = vg.init_scalar_field(np.ones(vg.shape))
            self.mask = vg.pad_periodic(vg.bc.trim_boundary_nodes(m))
            self.norm = 1.0
        else:
            m = vg.init_scalar_field(np.asarray(self.mask))
            mask_0 = m[:, 0, :, :]
            mask_1 = m[:, -1, :, :]
            self.mask = vg.pad_periodic(vg.bc.trim_boundary_nodes(m))
            if self.bc_type[0] != "periodic":
                self.mask[:, 0, 1:-1, 1:-1] = mask_0
                self.mask[:, -1, 1:-1, 1:-1] = mask_1
            self.norm = torch.sqrt(vg.gradient_norm_squared(self.mask))
            self.mask = torch.clip(self.mask, 1e-4, 1)
            x_bc, y_bc, z_bc = self.bc
            if x_bc[0] == "dirichlet":
                self.bc = (("dirichlet", (x_bc[1][0] * self.mask[:, 0, :, :], x_bc[1][1] * self.mask[:, -1, :, :])),
                           y_bc, z_bc)
                self.initialize_boundary_conditions()

    @property
    def _kernel_ok(self):
        return False

    @property
    def fused(self):
        return False

    def rhs_analytic(self, t, u, mask):
        grad_m = spv.gradient(mask)
        norm_grad_m = sp.sqrt(grad_m.dot(grad_m))
        divergence = spv.divergence(self.D * (spv.gradient(u) - u / mask * grad_m))
        return divergence + norm_grad_m * self.bc_flux + mask * self._eval_f(t, u / mask, sp)

    def rhs(self, t, u):
        u = self._check_state(u)
        vg = self.vg
        z = self.pad_bc(u)
        m = self.mask
        if isinstance(self.bc_flux, (int, float)) and not isinstance(self.bc_flux, bool):
            # one HIP launch for the divergence and the boundary flux term
            z = z.contiguous()
            out = torch.empty_like(u)
            g = self._grid(u)
            norm = self.norm
            norm_t = norm.contiguous() if isinstance(norm, torch.Tensor) else None
            norm_s = 0.0 if norm_t is not None else float(norm)
            _lib.call("evx_sbm_rhs_padded", _lib.dtype_code(u.dtype), _lib.ptr(z), _lib.ptr(m.contiguous()),
                      _lib.ptr(norm_t), norm_s, _lib.ptr(out), C.byref(g), float(self.D), float(self.bc_flux),
                      _lib.stream_handle(u.device))
            if not _is_default(self.f):
                inner = m[:, 1:-1, 1:-1, 1:-1]
                out = out + inner * self._eval_f(t, u / inner, torch)
            return out
        divergence = vg.grad_x_face(vg.grad_x_face(z) - vg.to_x_face(z / m) * vg.grad_x_face(m))[:, :, 1:-1, 1:-1]
        divergence += vg.grad_y_face(vg.grad_y_face(z) - vg.to_y_face(z / m) * vg.grad_y_face(m))[:, 1:-1, :, 1:-1]
        divergence += vg.grad_z_face(vg.grad_z_face(z) - vg.to_z_face(z / m) * vg.grad_z_face(m))[:, 1:-1, 1:-1, :]
        inner = m[:, 1:-1, 1:-1, 1:-1]
        return self.D * divergence + self.norm * self.bc_flux + inner * self._eval_f(t, u / inner, torch)
