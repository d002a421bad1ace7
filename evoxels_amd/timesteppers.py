"""Time steppers for the 'hip' backend (mirror of evoxels.timesteppers).

``step(t, u) -> u`` keeps the contract of evoxels/timesteppers.py:9-29.  For a
problem with a fused kernel (default closures) the state update is fused into
the right-hand-side kernel: ForwardEuler is one launch, RungeKutta4 four, and
PseudoSpectralIMEX on a fully periodic Cahn-Hilliard or reaction-diffusion
problem is the five-sweep spectral pipeline of evx_*_imex_step, and with x
non-periodic the mirrored pipeline of evx_*_imex_step_mirror.  Otherwise
the reference's composition runs over the HIP operators.
"""
from __future__ import annotations

import ctypes as C
from abc import ABC, abstractmethod
from dataclasses import dataclass, field
from typing import Any

import torch

from . import _lib
from .problem_definition import ODE, SemiLinearODE, CahnHilliard, ReactionDiffusion

State = Any


class TimeStepper(ABC):
    """timesteppers.py:9-29."""

    @property
    @abstractmethod
    def order(self) -> int:
        ...

    @abstractmethod
    def step(self, t: float, u: State) -> State:
        ...

    #: True when ``step`` issues only HIP launches whose arguments do not
    #: depend on ``t`` (so a sequence of steps may be captured in a hipGraph).
    capturable = False


@dataclass
class ForwardEuler(TimeStepper):
    """u + dt*rhs(t, u) (timesteppers.py:32-43)."""
    problem: ODE
    dt: float

    @property
    def order(self) -> int:
        return 1

    @property
    def capturable(self):
        return self.problem.fused

    def step(self, t: float, u: State) -> State:
        if self.problem.fused:
            u = self.problem._check_state(u)
            out = torch.empty_like(u)
            self.problem.launch(u, out, _lib.make_epilogue(_lib.OUT_EULER, self.dt))
            return out
        return u + self.dt * self.problem.rhs(t, u)


@dataclass
class RungeKutta4(TimeStepper):
    """Classical RK4 (timesteppers.py:46-61); fused: four RHS launches whose
    epilogues form the stage inputs and the running sum k1 + 2k2 + 2k3 (+ k4)."""
    problem: ODE
    dt: float

    @property
    def order(self) -> int:
        return 4

    @property
    def capturable(self):
        return self.problem.fused

    def step(self, t: float, u: State) -> State:
        if not self.problem.fused:
            k1 = self.problem.rhs(t, u)
            k2 = self.problem.rhs(t + 0.5 * self.dt, u + 0.5 * self.dt * k1)
            k3 = self.problem.rhs(t + 0.5 * self.dt, u + 0.5 * self.dt * k2)
            k4 = self.problem.rhs(t + self.dt, u + self.dt * k3)
            return u + (self.dt / 6) * (k1 + 2 * k2 + 2 * k3 + k4)
        p = self.problem
        u = p._check_state(u)
        ya, yb, acc, out = (torch.empty_like(u) for _ in range(4))
        half = 0.5 * self.dt
        p.launch(u, ya, _lib.make_epilogue(_lib.OUT_RK4_FIRST, half, base=u, acc_out=acc))
        p.launch(ya, yb, _lib.make_epilogue(_lib.OUT_RK4_MID, half, base=u, acc_in=acc, acc_out=acc))
        p.launch(yb, ya, _lib.make_epilogue(_lib.OUT_RK4_MID, self.dt, base=u, acc_in=acc, acc_out=acc))
        p.launch(ya, out, _lib.make_epilogue(_lib.OUT_RK4_LAST, self.dt / 6, base=u, acc_in=acc))
        return out


@dataclass
class PseudoSpectralIMEX(TimeStepper):
    """First-order semi-implicit Fourier spectral step (timesteppers.py:64-89;
    Zhu and Chen 1999).  prefac = dt/(1 - dt*sigma(k)) stays float32 like the
    reference's k-table."""
    problem: SemiLinearODE
    dt: float
    _fft_prefac: Any = field(init=False, repr=False, default=None)

    def __post_init__(self):
        self.problem.verify_fft_bc_config()
        self.pad = self.problem.pad_fft_bc
        self._ws = None

    @property
    def _kind(self):
        """'ch' or 'rd' when the fully periodic fused pipeline applies to the
        problem as it stands (its ``rhs`` may be replaced after construction)."""
        p = self.problem
        if p.fused and p.bc_type == ("periodic", "periodic", "periodic"):
            if isinstance(p, CahnHilliard):
                return "ch"
            if isinstance(p, ReactionDiffusion):
                return "rd"
        return None

    @property
    def _mirror(self):
        """'ch' or 'rd' when the mirrored pipeline applies (x non-periodic)."""
        p = self.problem
        if p.fused and p.bc_type[0] != "periodic" and p.bc_type[1:] == ("periodic", "periodic"):
            staggered = p.vg.convention == "staggered_x"
            if isinstance(p, CahnHilliard) and not staggered:
                return "ch"
            if isinstance(p, ReactionDiffusion) and (not staggered or p.bc_type[0] == "dirichlet"):
                return "rd"
        return None

    @property
    def _fft_prefac_generic(self):
        if self._fft_prefac is None:
            self._fft_prefac = self._multiplier()
        return self._fft_prefac

    _FUSED = {"ch": "evx_ch_imex_step", "rd": "evx_rd_imex_step"}

    def _multiplier(self):
        """Spectral multiplier of the generic path (timesteppers.py:77)."""
        return self.dt / (1 - self.dt * self.problem.fourier_symbol)

    @property
    def order(self) -> int:
        return 1

    @property
    def capturable(self):
        return self._kind is not None

    def _params(self):
        prm = _lib.EvxImexParams()
        prm.dt = float(self.dt)
        for a in range(3):
            prm.spacing[a] = float(self.problem.vg.spacing[a])
        p = self.problem
        prm.D = float(p.D)
        prm.A = float(p.A)
        prm.eps = float(p.eps) if "ch" in (self._kind, self._mirror) else 0.0
        return prm

    def workspace(self, u):
        plan = self.problem.vg.plan(u.shape, u.dtype)
        if self._ws is None or self._ws.numel() < plan.workspace_bytes or self._ws.device != u.device:
            self._ws = torch.empty(plan.workspace_bytes, dtype=torch.uint8, device=u.device)
        return plan, self._ws

    def last_step_sums(self, u, nonfinite=None):
        """Compensated float64 sum, as a (1,) device tensor, of the state the
        last fused step wrote (``u``), folded from the per-block records of
        that step's final sweep; NaN / Inf set ``nonfinite`` like
        ``VoxelGridHip.channel_sums``.  None when that step kept no records
        (several channels, or a generic, mirrored or slab step)."""
        if self._kind is None or self._ws is None or u.shape[0] != 1:
            return None
        plan = self.problem.vg.plan(u.shape, u.dtype)
        out = torch.empty(1, dtype=torch.float64, device=u.device)
        st = _lib.load().evx_fft_plan_step_sums(plan.handle, _lib.ptr(self._ws), _lib.ptr(out), _lib.ptr(nonfinite),
                                                _lib.stream_handle(u.device))
        return out if st == _lib.EVX_OK else None

    def step_into(self, u, out):
        """Fused step writing u + update into ``out`` (``out`` may be ``u``)."""
        plan, ws = self.workspace(u)
        prm = self._params()
        fn = self._FUSED[self._kind]
        _lib.call(fn, plan.handle, _lib.ptr(u), _lib.ptr(out), _lib.ptr(ws), C.byref(prm),
                  _lib.stream_handle(u.device))
        return out

    _ETD = 0

    def _mirror_step(self, u):
        """Fused step with x non-periodic: the plan covers the mirror-extended
        x length (2n, or 2n+2 on staggered_x grids)."""
        p = self.problem
        staggered = p.vg.convention == "staggered_x"
        n = u.shape[1]
        ext = (1, 2 * n + 2 if staggered else 2 * n) + tuple(u.shape[2:])
        plan = p.vg.plan(ext, u.dtype)
        if self._ws is None or self._ws.numel() < plan.workspace_bytes or self._ws.device != u.device:
            self._ws = torch.empty(plan.workspace_bytes, dtype=torch.uint8, device=u.device)
        out = torch.empty_like(u)
        prm = self._params()
        bc = _lib.make_bc(p.bc, p.vg.convention)
        fn = "evx_ch_imex_step_mirror" if self._mirror == "ch" else "evx_rd_imex_step_mirror"
        _lib.call(fn, plan.handle, _lib.ptr(u), _lib.ptr(out), _lib.ptr(self._ws), C.byref(prm), C.byref(bc),
                  int(staggered), self._ETD, _lib.stream_handle(u.device))
        return out

    def step(self, t: float, u: State) -> State:
        if self._kind is not None:
            u = self.problem._check_state(u)
            return self.step_into(u, torch.empty_like(u))
        if self._mirror is not None and u.shape[0] == 1 and \
                not any(isinstance(v, torch.Tensor) for v in (self.problem.bc[0][1] or ())):
            return self._mirror_step(self.problem._check_state(u))
        vg = self.problem.vg
        dc = self.pad(self.problem.rhs(t, u))
        dc_fft = vg.rfftn(dc, dc.shape)
        dc_fft.mul_(self._fft_prefac_generic)  # the spectrum is a temporary: multiplied and transformed in place
        update = vg.irfftn(dc_fft, dc.shape, overwrite=True)[:, :u.shape[1]]
        return u + update


@dataclass
class ExponentialEuler(PseudoSpectralIMEX):
    """First-order exponential time differencing (ETD1) for semilinear
    problems (timesteppers.py:136-199; Hochbruck, Lubich, Selhofer 1998):
    u_{n+1} = u_n + dt*phi_1(dt*L)*rhs(u_n), phi_1(z) = (exp(z)-1)/z, with the
    (6,6) Pade form for |z| < 0.5.  The multiplier is built from the same
    float32 symbol as the semi-implicit step.  Fully periodic Cahn-Hilliard
    and reaction-diffusion problems with default closures run the fused
    five-sweep pipeline (evx_*_etd1_step)."""

    _FUSED = {"ch": "evx_ch_etd1_step", "rd": "evx_rd_etd1_step"}
    _ETD = 1
    _PADE_N = (1, 1 / 26, 5 / 156, 1 / 858, 1 / 5720, 1 / 205920, 1 / 8648640)
    _PADE_D = (1, -6 / 13, 5 / 52, -5 / 429, 1 / 1144, -1 / 25740, 1 / 1235520)

    def _multiplier(self):
        return self.dt * self.phi1(self.dt * self.problem.fourier_symbol)

    def phi1(self, z):
        """(exp(z)-1)/z, with the degree-6 Pade approximant where |z| < 0.5
        (timesteppers.py:155-169)."""
        phi = (torch.exp(z) - 1) / z
        idx = torch.abs(z) < 0.5
        phi[idx] = self.phiPade(z[idx], 6, self._PADE_N, self._PADE_D)
        return phi

    @staticmethod
    def phiPade(z, Q, Ncoeff, Dcoeff):
        """(Q,Q)-Pade approximation of phi_1 by Horner's rule
        (timesteppers.py:171-190)."""
        numerator = Ncoeff[Q]
        denominator = Dcoeff[Q]
        for k in range(Q - 1, -1, -1):
            numerator = numerator * z + Ncoeff[k]
            denominator = denominator * z + Dcoeff[k]
        return numerator / denominator
