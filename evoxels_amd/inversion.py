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
* ``BatchedInversion`` fits per-trajectory ``(eps, D)`` by gradient descent
  (Adam) on log-parameters;
* ``InversionModel`` keeps the reference's interface (constructor, ``solve``,
  ``forward_solve``, ``residuals``, ``train`` with ``data``/``inds``) for the
  'hip' backend, with one parameter set shared by all sequences.

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
class BatchedInversion:
    """Per-trajectory (eps, D) estimation for a batch of independent periodic
    Cahn-Hilliard trajectories with the semi-implicit step (one trajectory per
    channel, each with its own parameters), by Adam on log-parameters.
    ``pos_params`` are optimised in log space like the reference's
    (inversion.py:69-70,202-207).  The reference has no batched form: its
    ``InversionModel`` shares one parameter set across sequences."""
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
              lr: float = 0.05, checkpoint_every: int = 20, tol: float = 0.0, verbose: bool = False,
              method: str = "bfgs", rtol: float = 1e-6):
        """Fits every trajectory's (eps, D); returns (params dict, loss
        history (iterations, channels)).  ``method="bfgs"``: per-trajectory
        BFGS on the (log-)parameters with backtracking line searches, all
        trajectories evaluated in one batched forward + adjoint pass per trial
        (each trajectory's loss depends on its own parameters only); stops when
        every step is below ``rtol`` in the (log-)parameters.  ``"adam"``:
        Adam with learning rate ``lr``."""
        names = ("eps", "D")
        p = self.engine.params(initial_parameters)
        logm = torch.tensor([n in self.pos_params for n in names], device=p.device)
        if bool(((p <= 0) & logm).any()):
            raise ValueError("positive parameters must be > 0")
        theta = torch.where(logm, torch.log(p), p)
        n_steps = max(obs_steps)
        hist = []

        def value_grad(th):
            pp = torch.where(logm, torch.exp(th), th)
            loss, g = loss_and_grad(self.engine, u0, pp, n_steps, obs_steps, values, checkpoint_every)
            return loss, torch.where(logm, g * pp, g)  # chain rule through exp

        if method == "adam":
            m = torch.zeros_like(theta)
            v = torch.zeros_like(theta)
            b1, b2, eps_adam = 0.9, 0.999, 1e-12
            for it in range(1, max_steps + 1):
                loss, g = value_grad(theta)
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
        elif method == "bfgs":
            B = theta.shape[0]
            eye = torch.eye(2, dtype=theta.dtype, device=theta.device).expand(B, 2, 2)
            f, g = value_grad(theta)
            hist.append(f.detach().cpu().numpy())
            # first step: steepest descent of length lr in the (log-)parameters
            H = eye * (lr / g.norm(dim=1).clamp_min(1e-300))[:, None, None]
            active = torch.ones(B, dtype=torch.bool, device=theta.device)
            for it in range(max_steps):
                d = -(H @ g[:, :, None])[:, :, 0]
                slope = (g * d).sum(dim=1)
                bad = slope >= 0  # not a descent direction: restart from steepest descent
                if bool(bad.any()):
                    H[bad] = eye[bad] * (lr / g[bad].norm(dim=1).clamp_min(1e-300))[:, None, None]
                    d = -(H @ g[:, :, None])[:, :, 0]
                    slope = (g * d).sum(dim=1)
                alpha = torch.where(active, torch.ones_like(f), torch.zeros_like(f))
                done = ~active
                th_new, f_new, g_new = theta.clone(), f.clone(), g.clone()
                for _ in range(12):  # backtracking, every trajectory at its own step length
                    th_t = theta + alpha[:, None] * d
                    f_t, g_t = value_grad(th_t)
                    ok = (f_t <= f + 1e-4 * alpha * slope) & ~done
                    th_new[ok], f_new[ok], g_new[ok] = th_t[ok], f_t[ok], g_t[ok]
                    done |= ok
                    if bool(done.all()):
                        break
                    alpha = torch.where(done, alpha, alpha * 0.5)
                s_vec, y_vec = th_new - theta, g_new - g
                sy = (s_vec * y_vec).sum(dim=1)
                upd = sy > 1e-300
                if bool(upd.any()):  # BFGS update of each inverse Hessian
                    rho = torch.where(upd, 1.0 / sy.clamp_min(1e-300), torch.zeros_like(sy))[:, None, None]
                    A1 = eye - rho * (s_vec[:, :, None] @ y_vec[:, None, :])
                    A2 = eye - rho * (y_vec[:, :, None] @ s_vec[:, None, :])
                    Hn = A1 @ H @ A2 + rho * (s_vec[:, :, None] @ s_vec[:, None, :])
                    H = torch.where(upd[:, None, None], Hn, H)
                theta, f, g = th_new, f_new, g_new
                hist.append(f.detach().cpu().numpy())
                if verbose:
                    print(f"iteration {it}: loss {float(f.sum()):.6e}")
                active &= (s_vec.abs() > rtol * (1.0 + theta.abs())).any(dim=1)
                if not bool(active.any()):
                    break
        else:
            raise ValueError(f"unknown method {method!r}")
        p = torch.where(logm, torch.exp(theta), theta)
        return {"eps": p[:, 0].clone(), "D": p[:, 1].clone()}, np.array(hist)


class SaveAt:
    """Output times of a solve, standing in for ``diffrax.SaveAt(ts=...)``
    (jax and diffrax are not part of this backend): ``ts`` and ``subs.ts``
    hold the times, as the reference's notebooks read them."""

    def __init__(self, ts):
        self.ts = np.asarray(ts, dtype=np.float64).reshape(-1)
        self.subs = self

    @staticmethod
    def times(saveat):
        if isinstance(saveat, SaveAt):
            return saveat.ts
        ts = getattr(getattr(saveat, "subs", None), "ts", None)
        if ts is None:
            ts = getattr(saveat, "ts", saveat)
        return np.asarray(ts, dtype=np.float64).reshape(-1)


def validate_inds(inds):
    """Sequence checks of InversionModel.train (inversion.py:170-192): at least
    two indices per sequence, equal lengths and equal spacing."""
    ref_len = len(inds[0])
    if ref_len < 2:
        raise ValueError("Each sequence in inds must have at least 2 elements")
    ref_spacing = [inds[0][i + 1] - inds[0][i] for i in range(ref_len - 1)]
    for i, sequence in enumerate(inds):
        if len(sequence) != ref_len:
            raise ValueError(f"Sequence {i} has different length than first sequence")
        spacing = [sequence[j + 1] - sequence[j] for j in range(len(sequence) - 1)]
        if spacing != ref_spacing:
            raise ValueError(f"Sequence {i} has different spacing than first sequence")


def steps_of(ts, dt0):
    """Fixed-step counts of output times measured from ts[0] (the reference
    integrates with constant dt0 from ts[0], inversion.py:77-88)."""
    rel = np.asarray(ts, dtype=np.float64) - float(ts[0])
    n = np.rint(rel / dt0).astype(np.int64)
    if np.any(np.abs(n * dt0 - rel) > 1e-6 * np.maximum(1.0, np.abs(rel))) or np.any(np.diff(n) < 0):
        raise ValueError(f"output times {ts.tolist()} are not increasing multiples of dt0 = {dt0} from ts[0]")
    return [int(v) for v in n]


@dataclass
class InversionModel:
    """Parameter estimation with the reference's interface (inversion.py:20-233)
    on the 'hip' backend: ``InversionModel(vf, problem_cls, pos_params,
    problem_kwargs, backend)``, ``solve``, ``forward_solve``, ``residuals`` and
    ``train(initial_parameters, data, inds, ...)`` returning the fitted
    parameters.  The problem is ``CahnHilliard`` with the default ``mu_hom``
    and periodic boundaries; the parameters are any of ``eps`` and ``D``,
    shared by all sequences as in the reference's vmapped solve
    (inversion.py:121-123).  The forward solve is the semi-implicit step with
    constant ``dt0`` (PseudoSpectralIMEX_dfx, timesteppers.py:94-123); the
    gradient comes from the adjoint kernels (``loss_and_grad``), and the
    optimiser is BFGS on the (log-)parameters with a backtracking line
    search instead of the reference's Levenberg-Marquardt.  ``adjoint`` is
    accepted for signature compatibility and ignored."""
    vf: object
    problem_cls: type
    pos_params: Optional[Sequence[str]] = None
    problem_kwargs: Optional[Dict[str, object]] = None
    backend: str = "hip"
    device: str = "cuda"

    FIT = ("eps", "D")

    def __post_init__(self):
        from .problem_definition import CahnHilliard
        from .profiler import HipMemoryProfiler
        from .voxelgrid import VoxelGridHip
        self.problem_kwargs = dict(self.problem_kwargs or {})
        if self.backend != "hip":
            raise ValueError(f"Unsupported backend: {self.backend}")
        if not (isinstance(self.problem_cls, type) and issubclass(self.problem_cls, CahnHilliard)):
            raise NotImplementedError("the 'hip' inversion differentiates CahnHilliard (default mu_hom, periodic)")
        if self.problem_kwargs.get("mu_hom") is not None:
            raise NotImplementedError("the 'hip' inversion needs the default mu_hom")
        if self.problem_kwargs.get("bc") not in (None, "fully_periodic", ("periodic", "periodic", "periodic")):
            raise NotImplementedError("the 'hip' inversion needs periodic boundaries")
        self.vg = VoxelGridHip(self.vf.grid_info(), precision=self.vf.precision, device=self.device)
        self.profiler = HipMemoryProfiler(self.vg.device)
        self._engines = {}

    # ------------------------------------------------------------ helpers
    def _full_params(self, parameters):
        """(eps, D, A) from problem_kwargs, defaults and ``parameters`` (which
        already carry the exp of the positive ones)."""
        kw = {"eps": 3.0, "D": 1.0, "A": 0.25}
        kw.update({k: v for k, v in self.problem_kwargs.items() if k in kw})
        for k, v in parameters.items():
            if k not in self.FIT:
                raise NotImplementedError(f"parameter {k!r}: the 'hip' inversion fits eps and D")
            kw[k] = float(np.asarray(v).reshape(-1)[0])
        return kw

    def _engine(self, channels, dt0, A):
        key = (int(channels), float(dt0), float(A))
        if key not in self._engines:
            self._engines[key] = CHImexEngine(self.vg, channels, dt0, A)
        return self._engines[key]

    def _field(self, y):
        t = torch.as_tensor(np.asarray(y) if not isinstance(y, torch.Tensor) else y, dtype=self.vg.precision,
                            device=self.vg.device)
        return t.reshape((-1,) + tuple(self.vg.shape)).contiguous()

    def _exp(self, parameters):
        pos = self.pos_params or ()
        return {k: (np.exp(np.asarray(v, dtype=np.float64)) if k in pos else v) for k, v in parameters.items()}

    # ------------------------------------------------------------ solves
    def solve(self, parameters, y0, saveat, adjoint=None, dt0=0.1):
        """States at the output times (inversion.py:53-92): an array of shape
        (len(ts), Nx, Ny, Nz); positive parameters are given as logs."""
        kw = self._full_params(self._exp(parameters))
        ts = SaveAt.times(saveat)
        n = steps_of(ts, dt0)
        eng = self._engine(1, dt0, kw["A"])
        states = solve(eng, self._field(y0), [[kw["eps"], kw["D"]]], n[-1], sorted(set(n)))
        by_step = dict(zip(sorted(set(n)), states))
        return torch.stack([by_step[k][0] for k in n])

    def forward_solve(self, parameters, fieldname, saveat, dt0=0.1, verbose=True):
        """inversion.py:94-108: solves from vf.fields[fieldname], stores the
        last state back into it and returns all saved states."""
        from timeit import default_timer as timer
        start = timer()
        pos = self.pos_params or ()
        logp = {k: (np.log(np.asarray(v, dtype=np.float64)) if k in pos else v) for k, v in parameters.items()}
        sol = self.solve(logp, self.vf.fields[fieldname], saveat, dt0=dt0)
        torch.cuda.synchronize(self.vg.device)
        end = timer()
        self.vf.fields[fieldname] = sol[-1].cpu().numpy()
        if verbose:
            ts = SaveAt.times(saveat)
            self.profiler.print_memory_stats(start, end, int((ts[-1] - ts[0]) // dt0))
        return sol

    def residuals(self, parameters, y0s__values__saveat, adjoint=None, dt0=0.1):
        """values - predictions at the output times after the first
        (inversion.py:110-124), all sequences in one batched solve."""
        y0s, values, saveat = y0s__values__saveat
        kw = self._full_params(self._exp(parameters))
        ts = SaveAt.times(saveat)
        n = steps_of(ts, dt0)
        u0 = self._field(y0s)
        eng = self._engine(u0.shape[0], dt0, kw["A"])
        states = solve(eng, u0, [[kw["eps"], kw["D"]]] * u0.shape[0], n[-1], sorted(set(n)))
        by_step = dict(zip(sorted(set(n)), states))
        pred = torch.stack([by_step[k] for k in n[1:]], dim=1)
        return torch.as_tensor(np.asarray(values) if not isinstance(values, torch.Tensor) else values,
                               dtype=pred.dtype, device=pred.device) - pred

    # ------------------------------------------------------------ fitting
    def train(self, initial_parameters, data, inds, adjoint=None, rtol=1e-6, atol=1e-6, verbose=True,
              max_steps=1000, dt0=0.1):
        """Fits the parameters of ``initial_parameters`` to the sequences
        ``inds`` of ``data`` = {"ts", "ys"} (inversion.py:126-233) and returns
        them (positive ones exponentiated back)."""
        validate_inds(inds)
        names = [k for k in self.FIT if k in initial_parameters]
        if not names or len(names) != len(initial_parameters):
            raise NotImplementedError("the 'hip' inversion fits eps and/or D")
        pos = set(self.pos_params or ())
        for key in pos:
            if key in initial_parameters and np.any(np.asarray(initial_parameters[key]) <= 0):
                raise ValueError(f"Parameter {key} must be positive")
        ys, ts = data["ys"], np.asarray(data["ts"], dtype=np.float64)
        y0s = self._field(np.stack([np.asarray(ys[ind[0]]) for ind in inds]))
        obs_t = [ts[inds[0][i]] - ts[inds[0][0]] for i in range(len(inds[0]))]
        n = steps_of(np.asarray(obs_t), dt0)
        if len(set(n)) != len(n):
            raise ValueError("two observations of a sequence fall on the same step")
        obs = [self._field(np.stack([np.asarray(ys[ind[i]]) for ind in inds])) for i in range(1, len(inds[0]))]
        B = y0s.shape[0]
        base = self._full_params({})
        eng = self._engine(B, dt0, base["A"])
        shapes = {k: np.shape(initial_parameters[k]) for k in names}

        def unpack(theta):
            p = dict(base)
            for i, k in enumerate(names):
                p[k] = float(np.exp(theta[i])) if k in pos else float(theta[i])
            return p

        def value_grad(theta):
            p = unpack(theta)
            prm = [[p["eps"], p["D"]]] * B
            loss, g = loss_and_grad(eng, y0s, prm, n[-1], n[1:], obs, checkpoint_every=20)
            g = g.sum(dim=0).cpu().numpy()  # shared parameters: sum over sequences
            gt = np.array([g[self.FIT.index(k)] * (p[k] if k in pos else 1.0) for k in names])
            return float(loss.sum()), gt

        theta = np.array([np.log(float(np.asarray(initial_parameters[k]).reshape(-1)[0])) if k in pos
                          else float(np.asarray(initial_parameters[k]).reshape(-1)[0]) for k in names])
        f, g = value_grad(theta)
        H = np.eye(len(names)) / max(np.linalg.norm(g), 1e-300)  # first step of unit length in theta
        self.history = [f]
        for it in range(max_steps):
            d = -H @ g
            if g @ d >= 0:  # not a descent direction: restart from steepest descent
                H = np.eye(len(names)) / max(np.linalg.norm(g), 1e-300)
                d = -H @ g
            step = 1.0
            while True:
                theta_new = theta + step * d
                f_new, g_new = value_grad(theta_new)
                if f_new <= f + 1e-4 * step * (g @ d) or step < 1e-8:
                    break
                step *= 0.5
            s_vec, y_vec = theta_new - theta, g_new - g
            if verbose:
                print(f"Step: {it}, Loss: {f_new:.9g}, Step size: {step:.3g}, "
                      f"Params: {unpack(theta_new)}")
            # converged when the step in the (log-)parameters is below the
            # tolerances, or when the line search can no longer decrease the loss
            converged = bool(np.all(np.abs(s_vec) <= atol + rtol * np.abs(theta_new))) or \
                (step < 1e-8 and f_new >= f)
            theta, f, g = theta_new, f_new, g_new
            self.history.append(f)
            sy = s_vec @ y_vec
            if sy > 1e-300:  # BFGS update of the inverse Hessian
                rho = 1.0 / sy
                I = np.eye(len(names))
                H = (I - rho * np.outer(s_vec, y_vec)) @ H @ (I - rho * np.outer(y_vec, s_vec)) + \
                    rho * np.outer(s_vec, s_vec)
            if converged:
                break
        p = unpack(theta)
        return {k: np.full(shapes[k], p[k]) if shapes[k] else p[k] for k in names}
