"""Gradient-based parameter estimation for Cahn-Hilliard on the HIP backend.

The reference fits material parameters by differentiating a JAX/diffrax solve
(``InversionModel``, inversion.py:20-233): a forward solve at the observation
times, residuals ``values - pred[:, 1:]`` over a batch of sequences, positive
parameters handled through ``exp``/``log`` (inversion.py:69-70,93-94,202-207)
and a least-squares optimiser.  This module gives the same capability for the
torch path's semi-implicit Cahn-Hilliard step (``CahnHilliard`` +
``PseudoSpectralIMEX``, problem_definition.py:290-371, timesteppers.py:64-89):

* ``CHImexEngine`` runs forward steps and reverse (vector-Jacobian) steps of a
  batch of independent trajectories, one per channel, each with its own
  ``(eps, D)``, through ``evx_ch_imex_step`` / ``evx_ch_imex_step_vjp``;
* ``loss_and_grad`` evaluates ``J = 1/2 sum_obs |u(t_obs) - obs|^2`` per
  trajectory and ``dJ/d(eps, D)`` with checkpointing: states are kept every
  ``checkpoint_every`` steps and each segment is recomputed once in the
  reverse sweep;
* ``CHImexStep`` / ``CHRhs`` are ``torch.autograd.Function`` wrappers of the
  same kernels for use inside user graphs;
* ``InversionModel`` fits per-trajectory ``(eps, D)`` by gradient descent
  (Adam) on log-parameters.

Only periodic boundaries are supported (the semi-implicit step requires them
for Cahn-Hilliard, problem_definition.py:138-149).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Dict, Optional, Sequence

import numpy as np
import torch

from . import _lib


def _params_tensor(params, channels, device) -> torch.Tensor:
    """(channels, 2) float64 device tensor of (eps, D)."""
    if isinstance(params, dict):
        eps = torch.as_tensor(params["eps"], dtype=torch.float64).reshape(-1)
        D = torch.as_tensor(params["D"], dtype=torch.float64).reshape(-1)
        eps = eps.expand(channels) if eps.numel() == 1 else eps
        D = D.expand(channels) if D.numel() == 1 else D
        p = torch.stack([eps, D], dim=1)
    else:
        p = torch.as_tensor(params, dtype=torch.float64)
    if tuple(p.shape) != (channels, 2):
        raise ValueError(f"parameters must have shape ({channels}, 2) as (eps, D) per channel, got {tuple(p.shape)}")
    return p.to(device).contiguous()


class CHImexEngine:
    """Forward and reverse semi-implicit CH steps for a batch of trajectories
    of shape (channels, nx, ny, nz) on one grid.

    ``exact_symbol=False`` keeps the reference's float32 k-table multiplier
    (timesteppers.py:77, voxelgrid.py:84-114), so forward states equal the
    solver's.  ``exact_symbol=True`` evaluates the multiplier in float64 from
    the exact wavenumbers: the step is then a smooth function of (eps, D) in
    float64, which finite-difference checks and float64 estimation need."""

    def __init__(self, vg, channels: int, dt: float, A: float = 0.25, exact_symbol: bool = False):
        from .voxelgrid import _FFTPlan
        self.vg = vg
        self.dt = float(dt)
        self.A = float(A)
        self.shape = (int(channels),) + tuple(vg.shape)
        self.dtype = vg.precision
        self.device = vg.device
        self.plan = _FFTPlan(_lib.dtype_code(self.dtype), *self.shape)  # own plan: options are per plan
        lib = _lib.load()
        _lib.check(lib.evx_fft_plan_set_option(self.plan.handle, _lib.OPT_EXACT_SYMBOL, int(bool(exact_symbol))),
                   "evx_fft_plan_set_option")
        self.exact_symbol = bool(exact_symbol)
        n = C.c_size_t()
        _lib.check(lib.evx_ch_imex_vjp_workspace_bytes(self.plan.handle, C.byref(n)),
                   "evx_ch_imex_vjp_workspace_bytes")
        # The reverse workspace starts with the forward one (real staging and
        # spectrum), so one buffer serves both directions.
        self.ws = torch.empty(max(n.value, self.plan.workspace_bytes), dtype=torch.uint8, device=self.device)
        self.prm = _lib.EvxImexParams()
        self.prm.dt = self.dt
        for a in range(3):
            self.prm.spacing[a] = float(vg.spacing[a])
        self.prm.A = self.A
        self.prm.eps = 0.0
        self.prm.D = 0.0

    def _check(self, *ts):
        for t in ts:
            if tuple(t.shape) != self.shape or t.dtype != self.dtype or not t.is_contiguous() \
                    or t.device != self.device:
                raise ValueError(f"expected a contiguous {self.dtype} tensor of shape {self.shape} on "
                                 f"{self.device}, got {tuple(t.shape)} {t.dtype} on {t.device}")

    def params(self, params) -> torch.Tensor:
        return _params_tensor(params, self.shape[0], self.device)

    def step(self, u, params, out=None):
        """out = u + irfftn(prefac * rfftn(ch_rhs(u))) per channel."""
        out = torch.empty_like(u) if out is None else out
        self._check(u, out)
        self.prm.chan_params = params.data_ptr()
        _lib.call("evx_ch_imex_step", self.plan.handle, _lib.ptr(u), _lib.ptr(out), _lib.ptr(self.ws),
                  C.byref(self.prm), _lib.stream_handle(u.device))
        return out

    def step_vjp(self, u0, u1, lam1, params, dparams, lam0=None):
        """lam0 = lam1 + (du1/du0)^T lam1 terms; dparams (channels, 2) += dJ/d(eps, D)."""
        lam0 = lam1 if lam0 is None else lam0
        self._check(u0, u1, lam1, lam0)
        self.prm.chan_params = params.data_ptr()
        _lib.call("evx_ch_imex_step_vjp", self.plan.handle, _lib.ptr(u0), _lib.ptr(u1), _lib.ptr(lam1),
                  _lib.ptr(lam0), _lib.ptr(dparams), _lib.ptr(self.ws), C.byref(self.prm),
                  _lib.stream_handle(u0.device))
        return lam0


def ch_rhs_vjp(vg, u, w, eps, D, base=None, params=None):
    """(grad_u, dparams) of r = ch_rhs(u) against cotangent w (periodic)."""
    u = u.contiguous()
    w = w.contiguous()
    c = u.shape[0]
    g = _lib.make_grid(u.shape, vg.div_dx_host, vg.div_dx2_host)
    lib = _lib.load()
    n = C.c_size_t()
    _lib.check(lib.evx_ch_rhs_vjp_workspace_bytes(_lib.dtype_code(u.dtype), C.byref(g), C.byref(n)),
               "evx_ch_rhs_vjp_workspace_bytes")
    ws = torch.empty(n.value, dtype=torch.uint8, device=u.device)
    grad = torch.empty_like(u)
    dparams = torch.zeros((c, 2), dtype=torch.float64, device=u.device)
    _lib.call("evx_ch_rhs_vjp", _lib.dtype_code(u.dtype), _lib.ptr(u), _lib.ptr(w), _lib.ptr(base),
              _lib.ptr(grad), _lib.ptr(dparams), C.byref(g), float(eps), float(D), _lib.ptr(params),
              _lib.ptr(ws), _lib.stream_handle(u.device))
    return grad, dparams


class CHRhs(torch.autograd.Function):
    """r = CahnHilliard.rhs(u) (periodic, default mu_hom) with scalar eps, D
    as 0-dim tensors; differentiable in u, eps and D."""

    @staticmethod
    def forward(ctx, u, eps, D, vg):
        from .problem_definition import CahnHilliard
        prob = CahnHilliard(vg, eps=float(eps), D=float(D))
        ctx.save_for_backward(u, eps, D)
        ctx.vg = vg
        return prob.rhs(0.0, u)

    @staticmethod
    def backward(ctx, w):
        u, eps, D = ctx.saved_tensors
        grad, dp = ch_rhs_vjp(ctx.vg, u, w, float(eps), float(D))
        tot = dp.sum(dim=0)
        return grad, tot[0].to(eps.dtype).reshape(eps.shape), tot[1].to(D.dtype).reshape(D.shape), None


class CHImexStep(torch.autograd.Function):
    """u1 = PseudoSpectralIMEX(CahnHilliard).step(u0) for a batch of
    trajectories with per-channel params (channels, 2) = (eps, D)."""

    @staticmethod
    def forward(ctx, u0, params, engine):
        p = params.detach().to(torch.float64).contiguous()
        u1 = engine.step(u0.contiguous(), p)
        ctx.save_for_backward(u0, u1, p)
        ctx.engine = engine
        ctx.pdtype = params.dtype
        return u1

    @staticmethod
    def backward(ctx, lam1):
        u0, u1, p = ctx.saved_tensors
        dparams = torch.zeros_like(p)
        lam0 = torch.empty_like(u0)
        ctx.engine.step_vjp(u0, u1, lam1.contiguous(), p, dparams, lam0=lam0)
        return lam0, dparams.to(ctx.pdtype), None


def checkpoint_schedule(n_steps: int, every: int):
    """Segments [start, end) of the reverse sweep, last segment first."""
    if n_steps < 0 or every < 1:
        raise ValueError("n_steps must be >= 0 and checkpoint_every >= 1")
    starts = list(range(0, n_steps, every))
    return [(s, min(s + every, n_steps)) for s in reversed(starts)]


def _obs_map(obs_steps: Sequence[int], n_steps: int):
    m = {}
    for i, s in enumerate(obs_steps):
        s = int(s)
        if not 0 <= s <= n_steps:
            raise ValueError(f"observation step {s} outside [0, {n_steps}]")
        if s in m:
            raise ValueError(f"observation step {s} given twice")
        m[s] = i
    return m


def _misfit(u, obs):
    d = (u - obs).to(torch.float64)
    return 0.5 * (d * d).sum(dim=(1, 2, 3))


def solve(engine: CHImexEngine, u0, params, n_steps: int, save_steps: Sequence[int]):
    """Forward solve; returns the states at ``save_steps`` (list of tensors)."""
    p = engine.params(params)
    want = _obs_map(save_steps, n_steps)
    saved = [None] * len(want)
    u = u0.clone()
    if 0 in want:
        saved[want[0]] = u.clone()
    nxt = torch.empty_like(u)
    for k in range(n_steps):
        engine.step(u, p, out=nxt)
        u, nxt = nxt, u
        if k + 1 in want:
            saved[want[k + 1]] = u.clone()
    return saved


def loss_and_grad(engine: CHImexEngine, u0, params, n_steps: int, obs_steps: Sequence[int], obs: Sequence,
                  checkpoint_every: int = 20):
    """J_c = 1/2 sum_obs sum_voxels (u_c(t_obs) - obs_c)^2 per trajectory c
    and dJ_c/d(eps_c, D_c), by a forward sweep with checkpoints and a reverse
    sweep that recomputes each segment once.  Returns (J (channels,),
    grad (channels, 2)) as float64 device tensors."""
    p = engine.params(params)
    om = _obs_map(obs_steps, n_steps)
    if len(obs) != len(om):
        raise ValueError("one observation per observation step is required")
    u0 = u0.contiguous()
    engine._check(u0)
    ce = int(checkpoint_every)
    checkpoints = {0: u0.clone()}
    loss = torch.zeros(engine.shape[0], dtype=torch.float64, device=u0.device)
    if 0 in om:
        loss += _misfit(u0, obs[om[0]])
    u = u0.clone()
    nxt = torch.empty_like(u)
    for k in range(n_steps):
        engine.step(u, p, out=nxt)
        u, nxt = nxt, u
        if k + 1 in om:
            loss += _misfit(u, obs[om[k + 1]])
        if (k + 1) % ce == 0 and k + 1 < n_steps:
            checkpoints[k + 1] = u.clone()
    del nxt
    grad = torch.zeros_like(p)
    lam = torch.zeros_like(u0)
    states = [torch.empty_like(u0) for _ in range(min(ce, max(n_steps, 1)) + 1)]
    for s, e in checkpoint_schedule(n_steps, ce):
        states[0].copy_(checkpoints[s])
        for k in range(s, e):
            engine.step(states[k - s], p, out=states[k - s + 1])
        for k in range(e - 1, s - 1, -1):
            if k + 1 in om:
                lam += states[k + 1 - s] - obs[om[k + 1]]
            engine.step_vjp(states[k - s], states[k + 1 - s], lam, p, grad)
    # an observation of the initial state has no parameter dependence
    return loss, grad


@dataclass
class InversionModel:
    """Per-trajectory (eps, D) estimation for periodic Cahn-Hilliard with the
    semi-implicit step.  ``pos_params`` are optimised in log space like the
    reference's (inversion.py:69-70,202-207)."""
    vg: object
    channels: int
    dt: float
    A: float = 0.25
    pos_params: Sequence[str] = ("eps", "D")
    exact_symbol: bool = False
    engine: CHImexEngine = field(init=False, repr=False)

    def __post_init__(self):
        self.engine = CHImexEngine(self.vg, self.channels, self.dt, self.A, self.exact_symbol)

    def residuals(self, params, u0, obs_steps, values):
        """values - predictions at obs_steps (inversion.py:121-124)."""
        pred = solve(self.engine, u0, params, max(obs_steps), obs_steps)
        return [v - q for v, q in zip(values, pred)]

    def train(self, initial_parameters: Dict[str, object], u0, obs_steps, values, max_steps: int = 50,
              lr: float = 0.05, checkpoint_every: int = 20, tol: float = 0.0, verbose: bool = False):
        """Adam on (log) parameters; returns (params dict, loss history)."""
        names = ("eps", "D")
        p = self.engine.params(initial_parameters)
        logm = torch.tensor([n in self.pos_params for n in names], device=p.device)
        if bool(((p <= 0) & logm).any()):
            raise ValueError("positive parameters must be > 0")
        theta = torch.where(logm, torch.log(p), p)
        m = torch.zeros_like(theta)
        v = torch.zeros_like(theta)
        b1, b2, eps_adam = 0.9, 0.999, 1e-12
        hist = []
        n_steps = max(obs_steps)
        for it in range(1, max_steps + 1):
            p = torch.where(logm, torch.exp(theta), theta)
            loss, g = loss_and_grad(self.engine, u0, p, n_steps, obs_steps, values, checkpoint_every)
            g = torch.where(logm, g * p, g)  # chain rule through exp
            hist.append(loss.detach().cpu().numpy())
            if verbose:
                print(f"step {it}: loss {float(loss.sum()):.6e}")
            if tol and float(loss.sum()) < tol:
                break
            m = b1 * m + (1 - b1) * g
            v = b2 * v + (1 - b2) * g * g
            mh = m / (1 - b1 ** it)
            vh = v / (1 - b2 ** it)
            theta = theta - lr * mh / (torch.sqrt(vh) + eps_adam)
        p = torch.where(logm, torch.exp(theta), theta)
        return {"eps": p[:, 0].clone(), "D": p[:, 1].clone()}, np.array(hist)
