"""Reverse mode of the semi-implicit Cahn-Hilliard step: the HIP adjoint
kernels against torch autograd through the CPU oracle (float64), against
central finite differences, and per-channel parameters of the batched forward
step against the oracle run channel by channel."""
import numpy as np
import pytest
import torch

import evoxels_amd as evx
from evoxels_amd.inversion import (BatchedInversion, CHImexEngine, CHImexStep, CHRhs, InversionModel, SaveAt, ch_rhs_vjp,
                                   loss_and_grad, solve)
from evoxels_amd.problem_definition import CahnHilliard
from evoxels_amd.voxelgrid import VoxelGridHip
from oracle import evoxels_oracle as O

pytestmark = pytest.mark.gpu

P3 = (("periodic", None),) * 3


def grid_for(shape, precision, spacing=(1.0, 1.0, 1.0)):
    vf = evx.VoxelFields(shape, tuple(s * n for s, n in zip(spacing, shape)))
    vf.precision = precision
    return vf, VoxelGridHip(vf.grid_info(), precision=precision, device="cuda")


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def field_with_clip_cases(rng, shape):
    u = 0.5 + 0.15 * rng.standard_normal(shape)
    flat = u.reshape(-1)
    # values on and beyond the clip bounds: torch.clip passes the gradient on
    # the closed interval [0, 1]
    flat[::97] = 0.0
    flat[5::89] = 1.0
    flat[11::83] = -0.05
    flat[17::79] = 1.04
    return u


@pytest.mark.parametrize("precision", ["float64", "float32"])
@pytest.mark.parametrize("shape,spacing", [((16, 16, 16), (1.0, 1.0, 1.0)), ((12, 10, 14), (0.8, 1.2, 1.0)),
                                           ((8, 6, 70), (1.0, 1.0, 1.0))])
def test_ch_rhs_vjp_matches_oracle_autograd(shape, spacing, precision):
    vf, vg = grid_for(shape, precision, spacing)
    rng = np.random.default_rng(7)
    arr = np.stack([field_with_clip_cases(rng, shape) for _ in range(2)])
    w = rng.standard_normal((2,) + shape)
    eps, D = 2.5, 1.3
    ut = torch.tensor(arr, dtype=vg.precision, device="cuda")
    wt = torch.tensor(w, dtype=vg.precision, device="cuda")
    base = torch.tensor(rng.standard_normal((2,) + shape), dtype=vg.precision, device="cuda")
    grad, dp = ch_rhs_vjp(vg, ut, wt, eps, D, base=base)
    og = O.OracleGrid(shape, vf.spacing, "float64")
    u64 = torch.tensor(arr, dtype=torch.float64, requires_grad=True)
    e64 = torch.tensor(eps, dtype=torch.float64, requires_grad=True)
    d64 = torch.tensor(D, dtype=torch.float64, requires_grad=True)
    dpe, dpd = [], []
    gu = torch.zeros_like(u64)
    for c in range(2):
        r = O.ch_rhs(u64[c:c + 1], og, P3, e64, d64)
        ge, gd, gc = torch.autograd.grad((r * torch.tensor(w[c:c + 1])).sum(), (e64, d64, u64))
        dpe.append(float(ge))
        dpd.append(float(gd))
        gu += gc
    want_u = gu.numpy() + base.cpu().double().numpy()
    tol_u, tol_p = (1e-12, 1e-11) if precision == "float64" else (2e-5, 2e-5)
    assert rel(grad.cpu(), want_u) < tol_u
    assert rel(dp[:, 0].cpu(), dpe) < tol_p
    assert rel(dp[:, 1].cpu(), dpd) < tol_p


def test_ch_rhs_autograd_function_scalar_params():
    shape = (10, 12, 16)
    vf, vg = grid_for(shape, "float64")
    rng = np.random.default_rng(3)
    u = torch.tensor(0.5 + 0.1 * rng.standard_normal((1,) + shape), device="cuda", requires_grad=True)
    eps = torch.tensor(3.0, dtype=torch.float64, device="cuda", requires_grad=True)
    D = torch.tensor(0.7, dtype=torch.float64, device="cuda", requires_grad=True)
    r = CHRhs.apply(u, eps, D, vg)
    w = torch.tensor(rng.standard_normal((1,) + shape), device="cuda")
    (r * w).sum().backward()
    og = O.OracleGrid(shape, vf.spacing, "float64")
    u64 = u.detach().cpu().requires_grad_()
    e64 = torch.tensor(3.0, dtype=torch.float64, requires_grad=True)
    d64 = torch.tensor(0.7, dtype=torch.float64, requires_grad=True)
    r64 = O.ch_rhs(u64, og, P3, e64, d64)
    assert rel(r.detach().cpu(), r64.detach()) < 1e-13
    (r64 * w.cpu()).sum().backward()
    assert rel(u.grad.cpu(), u64.grad) < 1e-12
    assert abs(float(eps.grad) - float(e64.grad)) <= 1e-11 * abs(float(e64.grad))
    assert abs(float(D.grad) - float(d64.grad)) <= 1e-11 * abs(float(d64.grad))


def oracle_step_autograd(arr_c, shape, spacing, dt, e64, d64):
    og = O.OracleGrid(shape, spacing, "float64")
    k2 = O.rfft_k_squared(shape, spacing)
    pre = O.imex_prefac(dt, O.ch_symbol(k2, e64, d64))
    return O.imex_step(arr_c, lambda t, v: O.ch_rhs(v, og, P3, e64, d64), pre)


def test_batched_forward_per_channel_params_matches_oracle():
    shape, dt = (12, 16, 20), 0.1
    vf, vg = grid_for(shape, "float64", (1.0, 0.9, 1.1))
    rng = np.random.default_rng(11)
    arr = 0.5 + 0.05 * rng.standard_normal((3,) + shape)
    prm = np.array([[3.0, 1.0], [2.2, 0.6], [4.1, 1.7]])
    eng = CHImexEngine(vg, 3, dt)
    u = torch.tensor(arr, device="cuda")
    got = eng.step(u, eng.params(prm)).cpu().numpy()
    for c in range(3):
        ref = oracle_step_autograd(torch.tensor(arr[c:c + 1]), shape, vf.spacing, dt,
                                   torch.tensor(prm[c, 0]), torch.tensor(prm[c, 1])).numpy()
        upd = got[c] - arr[c]
        upd_ref = ref[0] - arr[c]
        assert np.linalg.norm(upd - upd_ref) / np.linalg.norm(upd_ref) < 1e-10


def test_batched_forward_per_channel_params_512_x_lines():
    """Three float32 channels of a (512, 5, 6) grid: 60 x lines of the
    512-point persistent x sweep in four 16-line bricks, two of which hold
    lines of two channels; each channel against its float64 oracle step."""
    shape, dt = (512, 5, 6), 0.1
    vf, vg = grid_for(shape, "float32", (1.0, 1.1, 0.9))
    rng = np.random.default_rng(12)
    arr = (0.5 + 0.05 * rng.standard_normal((3,) + shape)).astype(np.float32)
    prm = np.array([[3.0, 1.0], [2.2, 0.6], [4.1, 1.7]])
    eng = CHImexEngine(vg, 3, dt)
    u = torch.tensor(arr, device="cuda")
    got = eng.step(u, eng.params(prm)).cpu().numpy().astype(np.float64)
    for c in range(3):
        a64 = arr[c:c + 1].astype(np.float64)
        ref = oracle_step_autograd(torch.tensor(a64), shape, vf.spacing, dt,
                                   torch.tensor(prm[c, 0]), torch.tensor(prm[c, 1])).numpy()
        upd, upd_ref = got[c] - a64[0], ref[0] - a64[0]
        err = np.linalg.norm(upd - upd_ref) / np.linalg.norm(upd_ref)
        print(f"channel {c}: update rel-L2 {err:.3e}", flush=True)
        assert err < 5e-4


@pytest.mark.parametrize("shape,spacing", [((16, 16, 16), (1.0, 1.0, 1.0)), ((12, 10, 18), (0.9, 1.1, 1.0))])
def test_imex_step_vjp_matches_oracle_autograd(shape, spacing):
    dt = 0.1
    vf, vg = grid_for(shape, "float64", spacing)
    rng = np.random.default_rng(5)
    arr = np.stack([field_with_clip_cases(rng, shape) for _ in range(2)])
    lam1 = rng.standard_normal((2,) + shape)
    prm = np.array([[3.0, 1.0], [2.0, 1.4]])
    eng = CHImexEngine(vg, 2, dt)
    p = eng.params(prm)
    u0 = torch.tensor(arr, device="cuda")
    u1 = eng.step(u0, p)
    dparams = torch.zeros_like(p)
    lam0 = eng.step_vjp(u0, u1, torch.tensor(lam1, device="cuda"), p, dparams, lam0=torch.empty_like(u0))
    for c in range(2):
        u64 = torch.tensor(arr[c:c + 1], requires_grad=True)
        e64 = torch.tensor(prm[c, 0], requires_grad=True)
        d64 = torch.tensor(prm[c, 1], requires_grad=True)
        out = oracle_step_autograd(u64, shape, vf.spacing, dt, e64, d64)
        ge, gd, gu = torch.autograd.grad((out * torch.tensor(lam1[c:c + 1])).sum(), (e64, d64, u64))
        assert rel(lam0[c].cpu(), gu[0]) < 1e-12
        # autograd differentiates the float32 prefactor in float32 arithmetic
        assert rel(dparams[c, 0].cpu(), ge) < 2e-5
        assert rel(dparams[c, 1].cpu(), gd) < 2e-5


def make_inverse_problem(shape, precision, n_steps, obs_steps, dt=0.1, seed=1, exact=False):
    vf, vg = grid_for(shape, precision)
    rng = np.random.default_rng(seed)
    u0 = torch.tensor(0.5 + 0.05 * rng.standard_normal((2,) + shape), dtype=vg.precision, device="cuda")
    eng = CHImexEngine(vg, 2, dt, exact_symbol=exact)
    truth = np.array([[3.0, 1.0], [2.4, 1.3]])
    obs = solve(eng, u0, truth, n_steps, obs_steps)
    return vg, eng, u0, truth, obs


def central_differences(eng, u0, prm, n_steps, obs_steps, obs, rel_h):
    fd = np.zeros_like(prm)
    for c in range(prm.shape[0]):
        for j in range(2):
            h = rel_h * prm[c, j]
            pp, pm = prm.copy(), prm.copy()
            pp[c, j] += h
            pm[c, j] -= h
            lp = solve_loss(eng, u0, pp, n_steps, obs_steps, obs)
            lm = solve_loss(eng, u0, pm, n_steps, obs_steps, obs)
            fd[c, j] = (lp[c] - lm[c]) / (2 * h)
    return fd


def solve_loss(eng, u0, prm, n_steps, obs_steps, obs):
    states = solve(eng, u0, prm, n_steps, obs_steps)
    return [sum(0.5 * float(((s[c].double() - o[c].double()) ** 2).sum()) for s, o in zip(states, obs))
            for c in range(u0.shape[0])]


def test_trajectory_gradient_matches_central_differences():
    """float64 with the exact-symbol multiplier, 2 trajectories, 12 steps,
    observations at 4, 8, 12, checkpoints every 5 steps (a short last
    segment): dJ/d(eps, D) against central differences with a relative step
    of 1e-4, to 1e-4 of the largest gradient component of the trajectory."""
    n_steps, obs_steps = 12, [4, 8, 12]
    vg, eng, u0, truth, obs = make_inverse_problem((16, 16, 16), "float64", n_steps, obs_steps, exact=True)
    prm = truth * np.array([[1.1, 0.92], [0.9, 1.07]])
    loss, grad = loss_and_grad(eng, u0, prm, n_steps, obs_steps, obs, checkpoint_every=5)
    grad = grad.cpu().numpy()
    fd = central_differences(eng, u0, prm, n_steps, obs_steps, obs, 1e-4)
    err = np.abs(grad - fd) / np.abs(fd).max(axis=1, keepdims=True)
    print("adjoint", grad, "fd", fd, "rel err", err)
    assert np.all(err <= 1e-4)
    # the checkpoint interval does not change the gradient
    _, g1 = loss_and_grad(eng, u0, prm, n_steps, obs_steps, obs, checkpoint_every=1)
    _, g20 = loss_and_grad(eng, u0, prm, n_steps, obs_steps, obs, checkpoint_every=20)
    assert rel(g1.cpu(), grad) < 1e-12 and rel(g20.cpu(), grad) < 1e-12


def test_trajectory_gradient_reference_multiplier_central_differences():
    """The reference's float32 multiplier is piecewise constant in (eps, D) at
    the 1e-7 level, so differences need a larger step: 1e-3 relative, to 1e-3
    of the largest gradient component."""
    n_steps, obs_steps = 12, [4, 8, 12]
    vg, eng, u0, truth, obs = make_inverse_problem((16, 16, 16), "float64", n_steps, obs_steps)
    prm = truth * np.array([[1.1, 0.92], [0.9, 1.07]])
    _, grad = loss_and_grad(eng, u0, prm, n_steps, obs_steps, obs, checkpoint_every=5)
    grad = grad.cpu().numpy()
    fd = central_differences(eng, u0, prm, n_steps, obs_steps, obs, 1e-3)
    err = np.abs(grad - fd) / np.abs(fd).max(axis=1, keepdims=True)
    print("adjoint", grad, "fd", fd, "rel err", err)
    assert np.all(err <= 1e-3)


def test_trajectory_gradient_matches_oracle_autograd_and_autograd_function():
    n_steps, obs_steps = 6, [3, 6]
    shape = (12, 12, 16)
    vg, eng, u0, truth, obs = make_inverse_problem(shape, "float64", n_steps, obs_steps)
    prm = truth * 0.95
    loss, grad = loss_and_grad(eng, u0, prm, n_steps, obs_steps, obs, checkpoint_every=4)
    # CHImexStep chained under torch autograd
    p = torch.tensor(prm, device="cuda", requires_grad=True)
    u = u0.clone()
    J = 0
    for k in range(n_steps):
        u = CHImexStep.apply(u, p, eng)
        if k + 1 in obs_steps:
            J = J + 0.5 * ((u - obs[obs_steps.index(k + 1)]) ** 2).sum()
    J.backward()
    assert rel(p.grad.cpu(), grad.cpu()) < 1e-12
    assert abs(float(J.detach()) - float(loss.sum())) < 1e-12 * float(loss.sum())
    # the oracle differentiated by torch autograd on the CPU, channel by channel
    spacing = (1.0, 1.0, 1.0)
    for c in range(2):
        e64 = torch.tensor(prm[c, 0], requires_grad=True)
        d64 = torch.tensor(prm[c, 1], requires_grad=True)
        v = u0[c:c + 1].cpu()
        Jc = 0
        for k in range(n_steps):
            v = oracle_step_autograd(v, shape, spacing, 0.1, e64, d64)
            if k + 1 in obs_steps:
                Jc = Jc + 0.5 * ((v - obs[obs_steps.index(k + 1)][c:c + 1].cpu()) ** 2).sum()
        ge, gd = torch.autograd.grad(Jc, (e64, d64))
        assert rel(float(loss[c]), float(Jc.detach())) < 1e-10
        assert rel(grad[c, 0].cpu(), ge) < 2e-5
        assert rel(grad[c, 1].cpu(), gd) < 2e-5


def test_float32_trajectory_gradient_close_to_float64():
    n_steps, obs_steps = 10, [5, 10]
    shape = (16, 16, 16)
    vg64, e64, u64, truth, obs64 = make_inverse_problem(shape, "float64", n_steps, obs_steps)
    vg32, e32, u32, _, obs32 = make_inverse_problem(shape, "float32", n_steps, obs_steps)
    prm = truth * 1.05
    _, g64 = loss_and_grad(e64, u64, prm, n_steps, obs_steps, obs64)
    _, g32 = loss_and_grad(e32, u32, prm, n_steps, obs_steps, obs32)
    assert rel(g32.cpu(), g64.cpu()) < 2e-3


def test_inversion_model_reduces_misfit():
    n_steps, obs_steps = 10, [5, 10]
    vg, eng, u0, truth, obs = make_inverse_problem((16, 16, 16), "float64", n_steps, obs_steps)
    model = BatchedInversion(vg, channels=2, dt=0.1)
    start = {"eps": torch.tensor([3.6, 2.0]), "D": torch.tensor([0.8, 1.6])}
    err0 = np.abs(np.log(np.array([[3.6, 0.8], [2.0, 1.6]]) / truth))
    p, hist = model.train(start, u0, obs_steps, obs, max_steps=40, lr=0.05, method="adam")
    assert hist[-1].sum() < 0.05 * hist[0].sum()
    err1 = np.abs(np.log(np.stack([p["eps"].cpu().numpy(), p["D"].cpu().numpy()], 1) / truth))
    assert err1.sum() < 0.5 * err0.sum()
    # per-trajectory BFGS recovers both trajectories' parameters
    p, hist = model.train(start, u0, obs_steps, obs, max_steps=40)
    err2 = np.abs(np.log(np.stack([p["eps"].cpu().numpy(), p["D"].cpu().numpy()], 1) / truth))
    print("BFGS iterations", len(hist) - 1, "log errors", err2, flush=True)
    assert err2.max() < 1e-4


def config5_draw(n=128, B=16):
    """The inversion workload: truth (M, kappa) ~ U[0.5, 2] per trajectory,
    D = M and eps = kappa, the evaluation point truth * U[0.8, 1.25], and
    c0 = 0.5 + 0.05*N(0,1) per trajectory, numpy default_rng(4) in this
    order."""
    rng = np.random.default_rng(4)
    truth_mk = rng.uniform(0.5, 2.0, (B, 2))
    truth = np.stack([truth_mk[:, 1], truth_mk[:, 0]], 1)
    guess = truth * rng.uniform(0.8, 1.25, (B, 2))
    c0 = 0.5 + 0.05 * rng.standard_normal((B, n, n, n))
    return truth, guess, c0


def exact_k2(n, d=1.0):
    """|k|^2 on the rfft grid from float64 wavenumbers 2*pi*fftfreq(n, d),
    the exact-symbol multiplier's table."""
    k = [2 * np.pi * torch.fft.fftfreq(n, d=d, dtype=torch.float64) for _ in range(2)]
    k.append(2 * np.pi * torch.fft.rfftfreq(n, d=d, dtype=torch.float64))
    kx, ky, kz = torch.meshgrid(*k, indexing="ij")
    return kx ** 2 + ky ** 2 + kz ** 2


def test_config5_two_trajectories_128_adjoint_matches_oracle_autograd():
    """Trajectories 0 and 1 of the inversion workload (trajectory 1 has the
    smallest eps of the draw, 0.57, where the clip of CahnHilliard.rhs is
    active) on 128^3 in float64 with the exact-symbol multiplier: 40 steps,
    observations every 8 steps, checkpoints every 20.  The HIP adjoint
    gradient against torch autograd of the CPU oracle (the same discrete
    function, clip included, float64 wavenumbers), recomputed segment by
    segment with torch.utils.checkpoint; relative to each trajectory's
    largest gradient component, <= 1e-8."""
    import torch.utils.checkpoint as tuc
    n, dt, steps, ce = 128, 0.01, 40, 20
    obs_steps = [8, 16, 24, 32, 40]
    truth, guess, c0 = config5_draw(n)
    pick = [0, 1]
    truth, guess, c0 = truth[pick], guess[pick], c0[pick]
    assert guess[1, 0] < 0.6
    vf, vg = grid_for((n, n, n), "float64")
    eng = CHImexEngine(vg, 2, dt, exact_symbol=True)
    u0 = torch.tensor(c0, dtype=torch.float64, device="cuda")
    obs = solve(eng, u0, truth, steps, obs_steps)
    states = solve(eng, u0, guess, steps, obs_steps)
    clipped = [int(((s < 0) | (s > 1)).sum(dim=(1, 2, 3))[1]) for s in states]
    print(f"trajectory 1 (eps {guess[1, 0]:.3f}): voxels outside [0, 1] at the observations {clipped}", flush=True)
    loss, grad = loss_and_grad(eng, u0, guess, steps, obs_steps, obs, checkpoint_every=ce)
    grad = grad.cpu().numpy()
    k2 = exact_k2(n)
    og = O.OracleGrid((n, n, n), (1.0, 1.0, 1.0), "float64")
    for c in range(2):
        e64 = torch.tensor(guess[c, 0], dtype=torch.float64, requires_grad=True)
        d64 = torch.tensor(guess[c, 1], dtype=torch.float64, requires_grad=True)
        ob = [o[c:c + 1].cpu() for o in obs]

        def segment(v, e, d):
            pre = O.imex_prefac(dt, O.ch_symbol(k2, e, d))
            for _ in range(8):
                v = O.imex_step(v, lambda t, w: O.ch_rhs(w, og, P3, e, d), pre)
            return v

        v = torch.tensor(c0[c:c + 1], dtype=torch.float64)
        J = 0
        for s in range(0, steps, 8):  # checkpointed segments of 8 steps: one per observation
            v = tuc.checkpoint(segment, v, e64, d64, use_reentrant=False)
            J = J + 0.5 * ((v - ob[obs_steps.index(s + 8)]) ** 2).sum()
        ge, gd = torch.autograd.grad(J, (e64, d64))
        want = np.array([float(ge), float(gd)])
        err = np.abs(grad[c] - want).max() / np.abs(want).max()
        print(f"trajectory {c}: adjoint {grad[c]}, oracle autograd {want}, rel {err:.3e}; "
              f"loss {float(loss[c]):.6e} vs {float(J.detach()):.6e}", flush=True)
        assert abs(float(loss[c]) - float(J.detach())) <= 1e-10 * float(J.detach())
        assert err <= 1e-8


def test_reference_interface_forward_solve_constant_solution():
    """test_inversion.py:36-47 on the 'hip' backend: a constant field stays
    constant through forward_solve, which returns every saved state."""
    vf = evx.VoxelFields((4, 4, 4))
    vf.add_field("c", np.full((4, 4, 4), 0.5, dtype=np.float32))
    model = InversionModel(vf, CahnHilliard, {"eps": 3.0})
    sol = model.forward_solve({"D": 1.0}, "c", SaveAt(ts=[0.0, 0.1, 0.2]), dt0=0.1, verbose=False)
    assert tuple(sol.shape) == (3, 4, 4, 4)
    np.testing.assert_allclose(sol.cpu().numpy(), 0.5, atol=1e-6)
    assert vf.fields["c"].shape == (4, 4, 4)


def test_reference_interface_recovers_shared_parameters():
    """The workflow of docs/notebooks/04-parameter-estimation.ipynb at 16^3 in
    float64: data from forward_solve with D = 1, eps = 3 (dt0 0.1, 11 saved
    times), three sequences of three frames, train from D = 1.6, eps = 2.2
    with both parameters positive; the fit recovers the truth."""
    n = 16
    vf = evx.VoxelFields((n, n, n), (n, n, n))
    vf.precision = "float64"
    vf.add_field("c", 0.5 + 0.1 * (0.5 - np.random.default_rng(9).random((n, n, n))))
    model = InversionModel(vf, CahnHilliard, ["D", "eps"], {"mu_hom": None})
    saveat = SaveAt(ts=np.linspace(0, 2, 11))
    res = model.forward_solve({"D": 1.0, "eps": 3.0}, "c", saveat, dt0=0.1, verbose=False)
    data = {"ts": saveat.subs.ts, "ys": res.cpu().numpy()}
    inds = [[1, 2, 3], [4, 5, 6], [7, 8, 9]]
    r0 = model.residuals({"D": np.log(1.0), "eps": np.log(3.0)},
                         (data["ys"][[1, 4, 7]], data["ys"][[[2, 3], [5, 6], [8, 9]]], SaveAt(ts=[0.0, 0.2, 0.4])))
    assert float(r0.abs().max()) < 1e-12
    fit = model.train({"D": np.array([1.6]), "eps": np.array([2.2])}, data, inds, max_steps=60, verbose=False)
    print("fitted", fit, "losses", model.history[:3], "...", model.history[-1], flush=True)
    assert abs(float(fit["D"][0]) - 1.0) < 1e-4 and abs(float(fit["eps"][0]) - 3.0) < 1e-4
    with pytest.raises(ValueError):
        model.train({"D": 1.0}, data, [[0, 1, 2], [0, 1]])
