"""Benchmark: semi-implicit spectral Cahn-Hilliard step, 512^3 float32, one MI355X.

Workload (evoxels' run_cahn_hilliard_solver physics, CahnHilliard +
PseudoSpectralIMEX): periodic 512^3 grid, domain 512 (dx = 1), eps = 3, D = 1,
A = 0.25, dt = 0.05, c0 = 0.5 + 0.05*N(0,1) drawn in float64 with
numpy.random.default_rng(2) and cast to float32.  A step is one full
semi-implicit update u <- u + irfftn(prefac * rfftn(rhs(u))) of the resident
field.

Printed JSON line (last line of stdout):
  value          voxel-steps/s over the timed steps (inputs resident in HBM)
  roofline       dominant kernel: algorithmic bytes / mean kernel time (HIP
                 events around each kernel inside the timed region) vs 8 TB/s
  step_roofline  whole step against the 5-sweep algorithmic minimum
                 s*(3 + 8*(1+2/Nz)) = 44.125 B/voxel
  kernels        per-kernel mean time, algorithmic bytes and GB/s
  cpu_baseline   the CPU oracle (op-faithful torch eager restatement of the
                 reference) timed on this host for a bounded sample of the
                 same workload
"""
from __future__ import annotations

import argparse
import ctypes
import gc
import json
import os
import sys
import time
from pathlib import Path

import numpy as np
import torch

METRIC = "voxel-steps/s + HBM GB/s (CH 512^3 FP32); fwd+adjoint step vs torch autograd"
PEAK_GBPS = 8000.0
ROOT = Path(__file__).resolve().parent


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def stage_bytes(n, s=4):
    """(kernel, algorithmic bytes per voxel of its sweep) by the plan's trace
    kinds: 0 RHS, 1 z R2C, 2 y forward, 3 x forward*prefac*inverse,
    4 y inverse, 5 z C2R + add, 6 fused RHS + z R2C."""
    nh = n // 2 + 1
    spec = 2 * s * nh / n  # one half-spectrum array, per voxel
    return [
        ("ch_rhs_kernel", 2 * s),                    # read c, write rhs
        ("z_r2c_fast", s + spec),                    # read rhs, write spectrum
        ("strided_fast<y,fwd>", 2 * spec),           # read + write spectrum
        ("x_sweep<fwd*prefac*inv>", 2 * spec),
        ("strided_fast<y,inv>", 2 * spec),
        ("z_c2r_fast", spec + 2 * s),                # read spectrum + u, write u
        ("ch_rhs_z_r2c_512", s + spec),              # read c, write spectrum
    ]


def traffic_from_profiles(kernel):
    f = ROOT / "profiles" / "traffic_ch_imex_512_fp32.json"
    if not f.exists():
        return None
    d = json.loads(f.read_text())
    k = d.get("kernels", {}).get(kernel)
    return None if k is None else k.get("hbm_bytes_per_launch")



def energy_decay_leg(out, vg, stepper, c0):
    """Free energy and mass over 500 steps of the headline workload, frames
    every 50 steps: F must not increase and the mass must stay put."""
    import torch
    u = vg.init_scalar_field(c0).contiguous()
    energies, masses = [float(vg.ch_energy(u, 3.0)[0])], [float(vg.channel_sums(u)[0])]
    for k in range(1, 501):
        stepper.step_into(u, u)
        if k % 50 == 0:
            energies.append(float(vg.ch_energy(u, 3.0)[0]))
            masses.append(float(vg.channel_sums(u)[0]))
    de = np.diff(energies)
    out["energy_decay"] = {"steps": 500, "frame_every": 50, "free_energy": [round(e, 3) for e in energies],
                           "non_increasing": bool(np.all(de <= 0)), "largest_frame_change": float(de.max()),
                           "mass_rel_drift": float(np.abs(np.array(masses) - masses[0]).max() / abs(masses[0])),
                           "functional": "int eps|grad c|^2 + 9/eps c^2(1-c)^2 dV, central gradients, float64 sums"}
    log(f"[bench] energy decay: {out['energy_decay']}")
    del u
    torch.cuda.empty_cache()


def cpu_baseline(c0, steps, gpu_step1=None):
    """Times the CPU oracle on `steps` steps (after one warm-up step).  With
    ``gpu_step1`` (the HIP step of c0, on the host) one float64 oracle step
    from the same float32 values checks it: field and update relative L2."""
    import torch
    sys.path.insert(0, str(ROOT))
    from oracle import evoxels_oracle as O
    n = c0.shape[0]
    P3 = (("periodic", None),) * 3
    parity = None
    if gpu_step1 is not None:
        g64 = O.OracleGrid((n, n, n), (1.0, 1.0, 1.0), "float64")
        pre64 = O.imex_prefac(0.05, O.ch_symbol(O.rfft_k_squared((n, n, n), g64.spacing), 3.0, 1.0))
        u0 = g64.field(c0.astype(np.float64))
        ref = O.imex_step(u0, lambda t, v: O.ch_rhs(v, g64, P3, 3.0, 1.0), pre64)
        got = torch.as_tensor(gpu_step1, dtype=torch.float64).reshape(ref.shape)
        parity = {"field_rel_l2": float(torch.linalg.norm(got - ref) / torch.linalg.norm(ref)),
                  "update_rel_l2": float(torch.linalg.norm((got - u0) - (ref - u0)) / torch.linalg.norm(ref - u0)),
                  "tolerance": {"field": 1e-5, "update": 5e-4},
                  "what": "one float32 HIP step from c0 against one float64 oracle step from the same values"}
        del g64, pre64, u0, ref, got
    g = O.OracleGrid((n, n, n), (1.0, 1.0, 1.0), "float32")
    pre = O.imex_prefac(0.05, O.ch_symbol(O.rfft_k_squared((n, n, n), g.spacing), 3.0, 1.0))
    rhs = lambda t, v: O.ch_rhs(v, g, P3, 3.0, 1.0)  # noqa: E731
    u = g.field(c0)
    u = O.imex_step(u, rhs, pre)
    t0 = time.perf_counter()
    for _ in range(steps):
        u = O.imex_step(u, rhs, pre)
    dt = time.perf_counter() - t0
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": n ** 3 * steps / dt, "unit": "voxel-steps/s", "cores": torch.get_num_threads(),
            "kind": "port", "s_per_step": dt / steps, "cpu_model": model, "os_cpu_count": os.cpu_count(),
            "rel_l2_vs_fp64_oracle": parity,
            "sample": f"{steps} timed steps (after 1 warm-up) of the same 512^3 float32 workload, "
                      f"torch {torch.__version__} eager on CPU, os.cpu_count()={os.cpu_count()}"}


def _pad_periodic(x):
    """Periodic ghost layer on the three spatial axes of a (B, nx, ny, nz) tensor."""
    x = torch.cat([x[:, -1:], x, x[:, :1]], 1)
    x = torch.cat([x[:, :, -1:], x, x[:, :, :1]], 2)
    return torch.cat([x[:, :, :, -1:], x, x[:, :, :, :1]], 3)


def _torch_ch_rhs(c, eps, D, idx, idx2):
    """Eager torch form of CahnHilliard.rhs (problem_definition.py:328-371) in
    the reference's operator structure (pad, face averages on the padded
    array, face gradients, trim), with per-trajectory eps, D of shape
    (B, 1, 1, 1).  Used only as the torch-autograd comparator."""
    c = torch.clip(c, 0, 1)
    cb = _pad_periodic(c)
    sum2 = (idx2[0] + idx2[1]) + idx2[2]
    lap = (((cb[:, 2:, 1:-1, 1:-1] + cb[:, :-2, 1:-1, 1:-1]) * idx2[0]
            + (cb[:, 1:-1, 2:, 1:-1] + cb[:, 1:-1, :-2, 1:-1]) * idx2[1])
           + (cb[:, 1:-1, 1:-1, 2:] + cb[:, 1:-1, 1:-1, :-2]) * idx2[2]) - 2 * c * sum2
    mu = 18 / eps * c * (1 - c) * (1 - 2 * c) - 2 * eps * lap
    mb = _pad_periodic(mu)
    div = 0
    for a in range(3):
        ax = a + 1
        n = cb.shape[ax]
        cf = 0.5 * (cb.narrow(ax, 1, n - 1) + cb.narrow(ax, 0, n - 1))
        flux = cf * (1 - cf) * ((mb.narrow(ax, 1, n - 1) - mb.narrow(ax, 0, n - 1)) * idx[a])
        d = (flux.narrow(ax, 1, n - 2) - flux.narrow(ax, 0, n - 2)) * idx[a]
        sl = [slice(None)] + [slice(None) if b == a else slice(1, -1) for b in range(3)]
        div = div + d[tuple(sl)]
    return D * div


def inversion_leg(steps, reps, torch_reps, rec_iters=15):
    """Config: 16 trajectories of 128^3 float32, truth (M, kappa) ~ U[0.5, 2]
    (numpy default_rng(4)), D = M and eps = kappa, kappa being the
    coefficient of |grad c|^2 in the reference's free energy
    F = int eps|grad c|^2 + 9/eps c^2(1-c)^2; dt = 0.01, domain 128,
    c0 = 0.5 + 0.05*N(0,1), observations at 5 frames, evaluated at a guess
    within a factor 1.25 of the truth.  Times one gradient evaluation
    J, dJ/d(eps, D) (forward sweep with checkpoints every 20 steps, segment
    recompute, adjoint sweep) of the HIP kernels and of torch autograd with
    torch.utils.checkpoint every 20 steps on the same GPU, and reports the
    accuracy of both against a float64 reference and central differences."""
    import torch.utils.checkpoint as tuc
    from evoxels_amd.inversion import CHImexEngine, loss_and_grad, solve
    from evoxels_amd.voxelgrid import VoxelGridHip
    import evoxels_amd as evx
    B, n, dt, A, ce = 16, 128, 0.01, 0.25, 20
    obs_steps = [steps * (i + 1) // 5 for i in range(5)]
    rng = np.random.default_rng(4)
    truth_mk = rng.uniform(0.5, 2.0, (B, 2))
    truth = np.stack([truth_mk[:, 1], truth_mk[:, 0]], 1)  # (eps, D) = (kappa, M)
    guess = truth * rng.uniform(0.8, 1.25, (B, 2))
    c0 = 0.5 + 0.05 * rng.standard_normal((B, n, n, n))
    vf = evx.VoxelFields((n, n, n), (n, n, n))
    vg = VoxelGridHip(vf.grid_info(), precision="float32", device="cuda")
    eng = CHImexEngine(vg, B, dt, A)
    u0 = torch.tensor(c0, dtype=torch.float32, device="cuda")
    obs = solve(eng, u0, truth, steps, obs_steps)
    log(f"[bench] inversion: {B}x{n}^3, {steps} steps, obs at {obs_steps}")
    loss, grad = loss_and_grad(eng, u0, guess, steps, obs_steps, obs, ce)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        loss, grad = loss_and_grad(eng, u0, guess, steps, obs_steps, obs, ce)
    torch.cuda.synchronize()
    t_hip = (time.perf_counter() - t0) / reps
    log(f"[bench] inversion: hip {t_hip:.3f} s/gradient")

    idx = [1.0, 1.0, 1.0]
    k = [2 * np.pi * torch.fft.fftfreq(n, d=1.0) for _ in range(2)] + [2 * np.pi * torch.fft.rfftfreq(n, d=1.0)]
    kx, ky, kz = torch.meshgrid(*k, indexing="ij")
    k2 = (kx ** 2 + ky ** 2 + kz ** 2).to("cuda")  # float32, as fftfreq returns
    obs_t = [o.detach() for o in obs]

    def run_torch():
        eps = torch.tensor(guess[:, 0], dtype=torch.float32, device="cuda").reshape(B, 1, 1, 1).requires_grad_()
        D = torch.tensor(guess[:, 1], dtype=torch.float32, device="cuda").reshape(B, 1, 1, 1).requires_grad_()

        def segment(u, e, d):
            pre = dt / (1 - dt * (-2 * e * d * A * k2 ** 2))
            for _ in range(ce):
                r = _torch_ch_rhs(u, e, d, idx, idx)
                u = u + torch.fft.irfftn(pre * torch.fft.rfftn(r, dim=(1, 2, 3)), s=(n, n, n), dim=(1, 2, 3))
            return u

        u = u0
        J = 0
        for s in range(0, steps, ce):
            u = tuc.checkpoint(segment, u, eps, D, use_reentrant=False)
            if s + ce in obs_steps:
                J = J + 0.5 * ((u - obs_t[obs_steps.index(s + ce)]) ** 2).sum(dim=(1, 2, 3), dtype=torch.float64)
        J.sum().backward()
        return J.detach(), torch.stack([eps.grad.reshape(B), D.grad.reshape(B)], 1).double()

    Jt, gt = run_torch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(torch_reps):
        Jt, gt = run_torch()
    torch.cuda.synchronize()
    t_torch = (time.perf_counter() - t0) / torch_reps
    log(f"[bench] inversion: torch autograd {t_torch:.3f} s/gradient")
    g_h = grad.cpu().numpy()
    g_t = gt.cpu().numpy()

    # recovery: Adam on log(eps), log(D) per trajectory from the guess, each
    # iteration one forward + adjoint gradient evaluation of all 16 trajectories
    from evoxels_amd.inversion import BatchedInversion
    inv = BatchedInversion(vg, B, dt, A)
    t0 = time.perf_counter()
    fit, hist = inv.train({"eps": guess[:, 0], "D": guess[:, 1]}, u0, obs_steps, obs, max_steps=rec_iters,
                          lr=0.05, checkpoint_every=ce, method="bfgs")
    torch.cuda.synchronize()
    t_rec = time.perf_counter() - t0
    rec = np.stack([fit["eps"].cpu().numpy(), fit["D"].cpu().numpy()], 1)
    recovery = {"iterations": len(hist) - 1, "max_iterations": rec_iters,
                "optimizer": "per-trajectory BFGS on log-parameters with backtracking line searches, first step "
                             "0.05 in log-parameters", "seconds": round(t_rec, 2),
                "loss_ratio_last_over_first": float(hist[-1].sum() / hist[0].sum()),
                "start_rel_err_max": float(np.abs(guess / truth - 1).max()),
                "kappa_rel_err_per_traj": [float(f"{v:.2e}") for v in np.abs(rec[:, 0] / truth[:, 0] - 1)],
                "M_rel_err_per_traj": [float(f"{v:.2e}") for v in np.abs(rec[:, 1] / truth[:, 1] - 1)],
                "recovered_kappa_M": [[round(float(a), 4), round(float(b), 4)] for a, b in rec],
                "truth_kappa_M": [[round(float(a), 4), round(float(b), 4)] for a, b in truth]}
    log(f"[bench] inversion recovery: {recovery['seconds']} s, max rel err kappa "
        f"{max(recovery['kappa_rel_err_per_traj']):.2e}, M {max(recovery['M_rel_err_per_traj']):.2e}")
    del inv
    loss_rel = float(np.abs(loss.cpu().numpy() - Jt.cpu().numpy()).max() / np.abs(Jt.cpu().numpy()).max())
    del Jt, gt
    torch.cuda.empty_cache()

    # Accuracy: the same problem in float64 (reference multiplier) gives the
    # reference gradient; with the exact-symbol multiplier the float64
    # gradient is checked against central differences (relative step 1e-4;
    # trajectories are independent, so one perturbed solve serves all 16).
    vg64 = VoxelGridHip(vf.grid_info(), precision="float64", device="cuda")
    u64 = torch.tensor(c0, dtype=torch.float64, device="cuda")

    def traj_loss(e, prm, ob):
        st = solve(e, u64, prm, steps, obs_steps)
        return sum(0.5 * ((x - o) ** 2).sum(dim=(1, 2, 3)) for x, o in zip(st, ob)).cpu().numpy()

    eng64 = CHImexEngine(vg64, B, dt, A)
    obs64 = solve(eng64, u64, truth, steps, obs_steps)
    _, g64 = loss_and_grad(eng64, u64, guess, steps, obs_steps, obs64, ce)
    g64 = g64.cpu().numpy()
    del eng64, obs64
    engx = CHImexEngine(vg64, B, dt, A, exact_symbol=True)
    obsx = solve(engx, u64, truth, steps, obs_steps)
    _, gx = loss_and_grad(engx, u64, guess, steps, obs_steps, obsx, ce)
    gx = gx.cpu().numpy()

    def central(h):
        out = np.zeros_like(gx)
        for j in range(2):
            pp, pm = guess.copy(), guess.copy()
            pp[:, j] *= 1 + h
            pm[:, j] *= 1 - h
            out[:, j] = (traj_loss(engx, pp, obsx) - traj_loss(engx, pm, obsx)) / (2 * h * guess[:, j])
        return out

    def clip_flips(h):
        """Per trajectory, the (voxel, step) pairs whose clip state (c < 0 or
        c > 1 entering the right-hand side) differs between the runs at
        eps*(1+h) and eps*(1-h): each is a kink of the loss that the central
        difference straddles."""
        pp, pm = guess.copy(), guess.copy()
        pp[:, 0] *= 1 + h
        pm[:, 0] *= 1 - h
        a, b = u64.clone(), u64.clone()
        ta, tb = torch.empty_like(a), torch.empty_like(b)
        pa, pb = engx.params(pp), engx.params(pm)
        flips = torch.zeros(B, dtype=torch.int64, device="cuda")
        for _ in range(steps):
            flips += (((a < 0) | (a > 1)) ^ ((b < 0) | (b > 1))).sum(dim=(1, 2, 3))
            engx.step(a, pa, out=ta)
            engx.step(b, pb, out=tb)
            a, ta, b, tb = ta, a, tb, b
        return flips.cpu().numpy()

    fd = central(1e-4)
    flips = clip_flips(1e-4)
    fd_small = central(1e-10)
    del engx, obsx
    torch.cuda.empty_cache()

    def rel_traj(a, ref):  # per trajectory, against its largest gradient component
        return (np.abs(a - ref) / np.abs(ref).max(axis=1, keepdims=True)).max(axis=1)

    def rel_err(a, ref):
        return float(rel_traj(a, ref).max())

    fd_traj = rel_traj(gx, fd)
    fd_small_traj = rel_traj(gx, fd_small)

    return {"workload": f"ch_inversion_{B}x{n}^3_fp32_{steps}_steps", "trajectories": B, "grid": [n, n, n],
            "steps": steps, "dt": dt, "checkpoint_every": ce, "obs_steps": obs_steps,
            "params": "per trajectory (eps, D); truth (M, kappa) ~ U[0.5,2], D = M, eps = kappa, default_rng(4)",
            "hip_s_per_gradient": round(t_hip, 4), "torch_autograd_s_per_gradient": round(t_torch, 4),
            "speedup_vs_torch_autograd": round(t_torch / t_hip, 2),
            "voxel_steps_per_s": B * n ** 3 * steps / t_hip,
            "grad_rel_diff_vs_torch_autograd": rel_err(g_h, g_t),
            "accuracy": {"hip_fp32_vs_fp64_reference": rel_err(g_h, g64),
                         "torch_autograd_fp32_vs_fp64_reference": rel_err(g_t, g64),
                         "hip_fp32_vs_fp64_per_trajectory": [float(f"{v:.2e}") for v in rel_traj(g_h, g64)],
                         "torch_fp32_vs_fp64_per_trajectory": [float(f"{v:.2e}") for v in rel_traj(g_t, g64)],
                         "fp64_exact_symbol_vs_central_differences": rel_err(gx, fd),
                         "fd_per_trajectory": [float(f"{v:.2e}") for v in fd_traj],
                         "trajectories_within_1e-4_of_fd": int((fd_traj <= 1e-4).sum()),
                         "eps_per_trajectory": [round(float(v), 3) for v in guess[:, 0]],
                         "clip_flips_between_fd_runs_h_1e-4": [int(v) for v in flips],
                         "fp64_exact_symbol_vs_central_differences_h_1e-10": rel_err(gx, fd_small),
                         "fd_h_1e-10_per_trajectory": [float(f"{v:.2e}") for v in fd_small_traj],
                         "trajectories_within_1e-4_of_fd_h_1e-10": int((fd_small_traj <= 1e-4).sum()),
                         "note": "relative to each trajectory's largest gradient component; central "
                                 "differences in float64 with relative steps 1e-4 and 1e-10. Where the clip "
                                 "of CahnHilliard.rhs is active (small eps: interfaces overshoot [0,1]) the "
                                 "loss has kinks in eps (clip-state flips between the two runs) and, between "
                                 "them, curvature of order 1e15 in eps, so the 1e-4 difference is a secant; "
                                 "the 1e-10 difference resolves the derivative"},
            "loss_rel_diff_vs_torch": loss_rel, "recovery": recovery}


def batched_leg(steps, seq_steps, cpu_steps):
    """48 periodic CH grids, edges drawn from {32,48,64,96,128} per axis with
    numpy default_rng(3), c0 = U[0.3,0.7] + 0.05*N(0,1) per grid, float32,
    dt = 0.1, eps = 3, D = 1 (run_cahn_hilliard_solver defaults), unit spacing.
    The autotuned batched step is captured in a graph of `steps` steps and
    replayed; compared with the single-grid fused step called grid by grid
    and with the CPU oracle run grid by grid."""
    from evoxels_amd.batched import BatchedCahnHilliard
    from evoxels_amd.problem_definition import CahnHilliard
    from evoxels_amd.timesteppers import PseudoSpectralIMEX
    from evoxels_amd.voxelgrid import VoxelGridHip
    import evoxels_amd as evx
    rng = np.random.default_rng(3)
    shapes = [tuple(int(v) for v in s) for s in rng.choice([32, 48, 64, 96, 128], (48, 3))]
    arrs = [rng.uniform(0.3, 0.7) + 0.05 * rng.standard_normal(s) for s in shapes]
    vox = int(sum(np.prod(s) for s in shapes))
    b_alg = sum(4 * (3 + 8 * (1 + 2 / s[2])) * np.prod(s) for s in shapes)  # five-sweep minimum
    b = BatchedCahnHilliard(shapes, 0.1, eps=3.0, D=1.0)
    state = b.pack(arrs)
    log(f"[bench] batched: {len(shapes)} grids, {vox} voxels; autotuning")
    sched = b.autotune(state, steps=5)
    log(f"[bench] batched: schedule {sched}")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def graph_ms():
        g = b.capture(state, steps)
        g.replay()  # warm-up replay
        torch.cuda.synchronize()
        e0.record()
        g.replay()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    ms = graph_ms()
    # the other launch schedule at the chosen tile shapes, for comparison
    b.set_tiles(sched["xchunk"], fft_group=sched["fft_group"], one_launch=not sched["one_launch"])
    ms_other = graph_ms()
    b.set_tiles(sched["xchunk"], fft_group=sched["fft_group"], one_launch=sched["one_launch"])
    by_launches = {("one_launch" if sched["one_launch"] else "six_launches"): round(ms, 4),
                   ("six_launches" if sched["one_launch"] else "one_launch"): round(ms_other, 4)}
    # grid-by-grid single-grid fused steps on the same GPU
    singles = []
    for s, a in zip(shapes, arrs):
        vf = evx.VoxelFields(s, s)
        vg = VoxelGridHip(vf.grid_info(), precision="float32", device="cuda")
        st = PseudoSpectralIMEX(CahnHilliard(vg, eps=3.0, D=1.0), 0.1)
        singles.append((st, vg.init_scalar_field(a).contiguous()))
    for st, u in singles:
        st.step_into(u, u)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(seq_steps):
        for st, u in singles:
            st.step_into(u, u)
    e1.record()
    torch.cuda.synchronize()
    ms_seq = e0.elapsed_time(e1) / seq_steps
    out = {"workload": "ch_batched_48_uneven_fp32", "grids": len(shapes), "voxels": vox, "steps": steps,
           "schedule": sched, "ms_per_step": round(ms, 4), "ms_per_step_by_launch_schedule": by_launches,
           "voxel_steps_per_s": vox / (ms * 1e-3),
           "alg_gbps": round(b_alg / (ms * 1e-3) / 1e9, 1),
           "note": "every grid fits the 256 MB last-level cache, so the GB/s figure is not an HBM fraction",
           "single_grid_loop_ms_per_step": round(ms_seq, 4), "speedup_vs_single_grid_loop": round(ms_seq / ms, 2)}
    if cpu_steps > 0:
        sys.path.insert(0, str(ROOT))
        from oracle import evoxels_oracle as O
        P3 = (("periodic", None),) * 3
        fields = []
        for s, a in zip(shapes, arrs):
            og = O.OracleGrid(s, (1.0, 1.0, 1.0), "float32")
            pre = O.imex_prefac(0.1, O.ch_symbol(O.rfft_k_squared(s, og.spacing), 3.0, 1.0))
            fields.append((og, pre, og.field(a)))
        t0 = time.perf_counter()
        for _ in range(cpu_steps):
            fields = [(og, pre, O.imex_step(u, lambda t, x, og=og: O.ch_rhs(x, og, P3, 3.0, 1.0), pre))
                      for og, pre, u in fields]
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": vox * cpu_steps / dt, "unit": "voxel-steps/s",
                               "cores": torch.get_num_threads(), "kind": "port",
                               "sample": f"{cpu_steps} steps of all 48 grids, grid by grid, CPU oracle"}
    return out


def _time_steps(fn, steps, reps=1):
    """Mean ms per call of fn over `steps` calls (after one warm-up call),
    bracketed by events on the current stream."""
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    best = None
    for _ in range(reps):
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        best = ms if best is None else min(best, ms)
    return best


def copy_ceiling():
    """Measured copy ceiling: torch copy_ of a 1 GiB float32 buffer, read +
    write bytes over the kernel time."""
    a = torch.rand(2 ** 28, device="cuda")
    b = torch.empty_like(a)
    ms = _time_steps(lambda: b.copy_(a), 10, reps=3)
    return {"bytes_per_copy": 2 * 4 * 2 ** 28, "ms": round(ms, 4), "gbps": round(2 * 4 * 2 ** 28 / (ms * 1e-3) / 1e9, 1)}


def configs_leg(cpu):
    """The other workloads of the scope: config 1 (100^3 FP64 CH semi-implicit,
    dt 0.1, 1000 steps, graph-replayed, parity spot-checked against the FP64
    oracle), config 2 (256^3 FP32 RD forward Euler with bc (Neumann,
    Dirichlet(0,1), periodic), dt = 0.1*dx^2/6, 2000 steps, plus the
    Allen-Cahn variant), and one line per next-row path: ETD1 at 512^3, the
    staggered_x Dirichlet RD step, the fused two-species step, the
    non-periodic (Neumann x) semi-implicit step and the generic multiphase
    and smoothed-boundary right-hand sides.  Bytes are algorithmic: 8 B per
    voxel for one-field explicit steps, 16 B for two species."""
    import warnings
    import evoxels_amd as evx
    from evoxels_amd.problem_definition import (CahnHilliard, CoupledReactionDiffusion, MultiPhaseAllenCahn,
                                                ReactionDiffusion, ReactionDiffusionSBM, TwoPhaseAllenCahn)
    from evoxels_amd.timesteppers import ExponentialEuler, ForwardEuler, PseudoSpectralIMEX
    from evoxels_amd.voxelgrid import VoxelGridHip
    sys.path.insert(0, str(ROOT))
    from oracle import evoxels_oracle as O
    out = {}
    warnings.simplefilter("ignore")

    def grid(shape, precision, domain=None, convention="cell_center"):
        vf = evx.VoxelFields(shape, domain or shape, convention=convention)
        vf.precision = precision
        return vf, VoxelGridHip(vf.grid_info(), precision=precision, device="cuda")

    # config 1: 100^3 FP64 CH IMEX, seed 0
    n = 100
    vf, vg = grid((n, n, n), "float64")
    c0 = 0.5 + 0.1 * np.random.default_rng(0).random((n, n, n))
    st = PseudoSpectralIMEX(CahnHilliard(vg, eps=3.0, D=1.0), 0.1)
    u = vg.init_scalar_field(c0).contiguous()
    m0 = float(vg.channel_sums(u)[0])
    st.step_into(u, u)  # plan and workspace
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    ug = u.clone()
    with torch.cuda.graph(g, stream=side):
        for _ in range(100):
            st.step_into(ug, ug)
    u1 = vg.init_scalar_field(c0).contiguous()
    ug.copy_(u1)
    ms = _time_steps(lambda: g.replay(), 10) / 100  # 1000 steps
    m1 = float(vg.channel_sums(ug)[0])
    # parity spot-check: 3 steps from c0 against the FP64 oracle
    chk = vg.init_scalar_field(c0).contiguous()
    og = O.OracleGrid((n, n, n), vf.spacing, "float64")
    pre = O.imex_prefac(0.1, O.ch_symbol(O.rfft_k_squared((n, n, n), og.spacing), 3.0, 1.0))
    ref = og.field(c0)
    P3 = (("periodic", None),) * 3
    worst = 0.0
    for _ in range(3):
        st.step_into(chk, chk)
        ref = O.imex_step(ref, lambda t, x: O.ch_rhs(x, og, P3, 3.0, 1.0), pre)
        got = chk.cpu().numpy()
        worst = max(worst, float(np.linalg.norm(got - ref.numpy()) / np.linalg.norm(ref.numpy())))
    out["config1_ch_imex_100cubed_fp64"] = {
        "steps": 1000, "ms_per_step": round(ms, 4), "voxel_steps_per_s": n ** 3 / (ms * 1e-3),
        "alg_gbps": round(89.28 * n ** 3 / (ms * 1e-3) / 1e9, 1), "mass_rel_drift": abs(m1 - m0) / abs(m0),
        "rel_l2_vs_fp64_oracle_3_steps": worst,
        "note": "24 MB working set (state, right-hand-side staging, half spectrum): resident in the 256 MB "
                "last-level cache, so not an HBM fraction"}
    # free-energy decay over the 1000 steps, frames every 100
    ue = vg.init_scalar_field(c0).contiguous()
    en = [float(vg.ch_energy(ue, 3.0)[0])]
    for k in range(1, 1001):
        st.step_into(ue, ue)
        if k % 100 == 0:
            en.append(float(vg.ch_energy(ue, 3.0)[0]))
    out["config1_ch_imex_100cubed_fp64"]["energy_decay"] = {
        "free_energy": [round(e, 4) for e in en], "non_increasing": bool(np.all(np.diff(en) <= 0)),
        "largest_frame_change": float(np.diff(en).max())}
    del ue
    if cpu:
        t0 = time.perf_counter()
        r = og.field(c0)
        for _ in range(5):
            r = O.imex_step(r, lambda t, x: O.ch_rhs(x, og, P3, 3.0, 1.0), pre)
        dt = time.perf_counter() - t0
        out["config1_ch_imex_100cubed_fp64"]["cpu_baseline"] = {
            "value": 5 * n ** 3 / dt, "unit": "voxel-steps/s", "cores": torch.get_num_threads(), "kind": "port",
            "sample": "5 steps of the CPU oracle"}
    del ug, u, u1, chk, g
    log(f"[bench] config 1: {out['config1_ch_imex_100cubed_fp64']}")

    # config 2: 256^3 FP32 RD forward Euler, bc (N, D(0,1), P), seed 1; AC variant
    n = 256
    vf, vg = grid((n, n, n), "float32", (1, 1, 1))
    phi0 = np.random.default_rng(1).random((n, n, n))
    dt = 0.1 * vf.spacing[0] ** 2 / 6
    bc = ("neumann", ("dirichlet", (0.0, 1.0)), "periodic")
    for name, prob in [("rd", ReactionDiffusion(vg, D=1.0, bc=bc)),
                       ("ac", TwoPhaseAllenCahn(vg, eps=2.0, curvature=0.5, bc=bc))]:
        fe = ForwardEuler(prob, dt)
        state = {"u": vg.init_scalar_field(phi0).contiguous()}

        def one():
            state["u"] = fe.step(0.0, state["u"])
        ms = _time_steps(one, 2000, reps=2)
        out[f"config2_{name}_fe_256cubed_fp32"] = {
            "steps": 2000, "ms_per_step": round(ms, 5), "voxel_steps_per_s": n ** 3 / (ms * 1e-3),
            "gbps_at_8B": round(8 * n ** 3 / (ms * 1e-3) / 1e9, 1),
            "note": "134 MB per step: fits the 256 MB last-level cache, so GB/s can exceed HBM peak"}
        if cpu and name == "rd":
            og = O.OracleGrid((n, n, n), vf.spacing, "float32")
            bcn = O.normalize_bc(bc)
            r = og.field(phi0)
            t0 = time.perf_counter()
            for _ in range(3):
                r = O.fe_step(r, lambda t, x: O.rd_rhs(x, og, bcn, 1.0), dt)
            tc = time.perf_counter() - t0
            out[f"config2_{name}_fe_256cubed_fp32"]["cpu_baseline"] = {
                "value": 3 * n ** 3 / tc, "unit": "voxel-steps/s", "cores": torch.get_num_threads(), "kind": "port",
                "sample": "3 forward-Euler steps of the CPU oracle"}
        log(f"[bench] config 2 {name}: {out[f'config2_{name}_fe_256cubed_fp32']}")
        del state

    # next rows
    n = 512
    vf, vg = grid((n, n, n), "float32")
    c0 = 0.5 + 0.05 * np.random.default_rng(2).standard_normal((n, n, n))
    etd = ExponentialEuler(CahnHilliard(vg, eps=3.0, D=1.0), 0.05)
    u = vg.init_scalar_field(c0).contiguous()
    ms = _time_steps(lambda: etd.step_into(u, u), 20)
    out["f2_ch_etd1_512cubed_fp32"] = {"ms_per_step": round(ms, 4), "voxel_steps_per_s": n ** 3 / (ms * 1e-3),
                                      "step_roofline_frac": round(44.125 * n ** 3 / (ms * 1e-3) / 1e9 / PEAK_GBPS, 4)}
    del u, etd
    torch.cuda.empty_cache()
    n = 256
    vf, vg = grid((n + 1, n, n), "float32", (1, 1, 1), "staggered_x")
    rd = ReactionDiffusion(vg, D=1.0, bc=(("dirichlet", (0.0, 1.0)), "periodic", "periodic"))
    fe = ForwardEuler(rd, 0.1 * vf.spacing[0] ** 2 / 6)
    state = {"u": torch.rand((1, n - 1, n, n), device="cuda")}

    def one_s():
        state["u"] = fe.step(0.0, state["u"])
    ms = _time_steps(one_s, 500, reps=3)
    vox = (n - 1) * n * n
    out["f1_staggered_rd_fe_256cubed_fp32"] = {"ms_per_step": round(ms, 5), "voxel_steps_per_s": vox / (ms * 1e-3),
                                              "gbps_at_8B": round(8 * vox / (ms * 1e-3) / 1e9, 1)}
    vf, vg = grid((n, n, n), "float32", (1, 1, 1))
    imex = PseudoSpectralIMEX(ReactionDiffusion(vg, D=1.0, bc=("neumann", "periodic", "periodic")), 1e-4)
    state = {"u": torch.rand((1, n, n, n), device="cuda")}

    def one_n():
        state["u"] = imex.step(0.0, state["u"])
    ms = _time_steps(one_n, 50, reps=2)
    out["f1_rd_imex_neumann_x_256cubed_fp32"] = {
        "ms_per_step": round(ms, 4), "voxel_steps_per_s": n ** 3 / (ms * 1e-3),
        "note": "mirrored pipeline: RHS with the Neumann ghosts, extension to 512 x 256 x 256 in the staging "
                "buffer, five sweeps over the extended grid, inverse sweeps on the 256 physical planes only"}
    gs = CoupledReactionDiffusion(vg)
    fe = ForwardEuler(gs, 0.1)
    state = {"u": torch.rand((2, n, n, n), device="cuda")}

    def one_g():
        state["u"] = fe.step(0.0, state["u"])
    ms = _time_steps(one_g, 500, reps=3)
    out["f4_gray_scott_fe_256cubed_fp32"] = {"ms_per_step": round(ms, 5), "voxel_steps_per_s": n ** 3 / (ms * 1e-3),
                                            "gbps_at_16B": round(16 * n ** 3 / (ms * 1e-3) / 1e9, 1)}
    n = 128
    vf, vg = grid((n, n, n), "float32")
    mp = MultiPhaseAllenCahn(vg, bc=("neumann", "periodic", "periodic"))
    ph = torch.rand((3, n, n, n), device="cuda")
    ms = _time_steps(lambda: mp.rhs(0.0, ph), 50)
    out["f4_multiphase_ac_rhs_3phases_128cubed_fp32"] = {"ms_per_call": round(ms, 4),
                                                         "voxel_calls_per_s": n ** 3 / (ms * 1e-3),
                                                         "note": "projection kernel, ghost pad, padded RHS kernel"}
    x, y, z = np.meshgrid(*[np.arange(n) / n] * 3, indexing="ij")
    sbm = ReactionDiffusionSBM(vg, D=1.0, mask=0.5 + 0.3 * np.cos(4 * np.pi * x) * np.cos(2 * np.pi * y),
                               bc=(("dirichlet", (0.0, 1.0)), "periodic", "periodic"), bc_flux=1.0)
    us = torch.rand((1, n, n, n), device="cuda")
    ms = _time_steps(lambda: sbm.rhs(0.0, us), 50)
    out["f4_smoothed_boundary_rhs_128cubed_fp32"] = {"ms_per_call": round(ms, 4),
                                                     "voxel_calls_per_s": n ** 3 / (ms * 1e-3),
                                                     "note": "ghost pad (mask-scaled node values) and one RHS kernel"}
    log(f"[bench] next rows: { {k: v for k, v in out.items() if k[:1] == 'f'} }")
    return out


def frames_leg(n, steps, every):
    """Frame output cost: run_cahn_hilliard_solver settings (eps 3, D 1,
    PseudoSpectralIMEX, dt 0.1, jit) on an n^3 float32 grid for `steps` steps
    with an output frame every `every` steps, exported synchronously (copy and
    NaN check at each frame, as the reference loop does) and asynchronously
    (pinned side-stream copies overlapped with the next steps); the run with
    one output at each end gives the step cost alone."""
    import evoxels_amd as evx
    from evoxels_amd.problem_definition import CahnHilliard
    from evoxels_amd.solvers import TimeDependentSolver
    from evoxels_amd.timesteppers import PseudoSpectralIMEX
    c0 = (0.5 + 0.05 * np.random.default_rng(5).standard_normal((n, n, n))).astype(np.float32)

    def run(frames, async_output):
        gc.collect()  # outside the timed solve: the capture's own collection then finds little
        vf = evx.VoxelFields((n, n, n), (n, n, n))
        vf.add_field("c", c0.copy())
        s = TimeDependentSolver(vf, "c", backend="hip", problem_cls=CahnHilliard, timestepper_cls=PseudoSpectralIMEX,
                                device="cuda", async_output=async_output)
        s.solve(time_increment=0.1, frames=frames, max_iters=steps, problem_kwargs={"eps": 3.0, "D": 1.0},
                jit=True, verbose=False)
        return s.computation_time, vf.fields["c"]

    run(1, False)  # plans, graph capture and pinned buffers warm up
    t_none, _ = run(1, False)
    t_sync, f_sync = run(steps // every, False)
    t_async, f_async = run(steps // every, True)
    nf = steps // every + 1
    out = {"workload": f"ch_imex_{n}cubed_fp32_frames", "steps": steps, "frames_written": nf,
           "frame_mb": round(4 * n ** 3 / 2 ** 20, 1),
           "ms_per_step_no_frames": round(1e3 * t_none / steps, 4),
           "ms_per_step_sync_frames": round(1e3 * t_sync / steps, 4),
           "ms_per_step_async_frames": round(1e3 * t_async / steps, 4),
           "ms_per_frame_sync": round(1e3 * (t_sync - t_none) / (nf - 2), 3),
           "ms_per_frame_async": round(1e3 * (t_async - t_none) / (nf - 2), 3),
           "final_frames_equal": bool(np.array_equal(f_sync, f_async))}
    log(f"[bench] frames: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--slabs", type=int, default=0,
                    help="x-slabs of the semi-implicit step (0: whole-grid sweeps)")
    ap.add_argument("--no-inversion", action="store_true", help="skip the forward+adjoint inversion leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched uneven-grid leg")
    ap.add_argument("--no-frames", action="store_true", help="skip the frame-output leg")
    ap.add_argument("--no-energy", action="store_true", help="skip the 500-step free-energy leg")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip configs 1-2, the next-row paths and the copy ceiling")
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--inv-steps", type=int, default=200)
    ap.add_argument("--inv-reps", type=int, default=3)
    ap.add_argument("--inv-torch-reps", type=int, default=1)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 device copies of 1 GiB (calibrates rocprof FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4, "launches": 5}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the single-GPU kernels are the product; multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    import torch
    import evoxels_amd as evx
    from evoxels_amd import _lib
    from evoxels_amd.problem_definition import CahnHilliard
    from evoxels_amd.timesteppers import PseudoSpectralIMEX
    from evoxels_amd.voxelgrid import VoxelGridHip

    n, K, W = args.n, args.steps, args.warmup
    log(f"[bench] generating c0 = 0.5 + 0.05*N(0,1), {n}^3, seed 2")
    c0 = 0.5 + 0.05 * np.random.default_rng(2).standard_normal((n, n, n))
    vf = evx.VoxelFields((n, n, n), (n, n, n))
    vg = VoxelGridHip(vf.grid_info(), precision="float32", device="cuda")
    prob = CahnHilliard(vg, eps=3.0, D=1.0)
    stepper = PseudoSpectralIMEX(prob, 0.05)
    u = vg.init_scalar_field(c0).contiguous()
    plan, _ = stepper.workspace(u)
    lib = _lib.load()
    _lib.check(lib.evx_fft_plan_set_option(plan.handle, _lib.OPT_SLABS, args.slabs), "slabs")

    mass0 = float(vg.channel_sums(u)[0])
    # one step from c0, kept on the host for the float64 oracle check of the CPU leg
    step1 = None if args.no_cpu_baseline else stepper.step(0.0, u).cpu().numpy()
    log(f"[bench] warmup {W} steps")
    for _ in range(W):
        stepper.step_into(u, u)
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream()
    NEV = 64  # events per step: one before the first kernel, one after each kernel
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(NEV)] for _ in range(K)]
    for row in ev:  # materialise the underlying hipEvent_t handles
        for e in row:
            e.record(stream)
    torch.cuda.synchronize()
    handles = [(ctypes.c_void_p * NEV)(*[e.cuda_event for e in row]) for row in ev]
    log(f"[bench] timing {K} steps")
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    t_start.record(stream)
    for k in range(K):
        _lib.check(lib.evx_fft_plan_trace_next_step(plan.handle, handles[k], NEV), "trace")
        stepper.step_into(u, u)
    t_end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms_step = t_start.elapsed_time(t_end) / K
    mass1 = float(vg.channel_sums(u)[0])
    log(f"[bench] {ms_step:.3f} ms/step (events), wall {1e3 * wall / K:.3f} ms/step")
    # the standalone compensated float64 mass + NaN/Inf pass (evx_sum_f64),
    # which frames use when no step records are available
    sum_ms = _time_steps(lambda: vg.channel_sums(u), 20, reps=2)
    sum_f64 = {"ms": round(sum_ms, 4), "gbps": round(4 * n ** 3 / (sum_ms * 1e-3) / 1e9, 1),
               "what": f"evx_sum_f64 of the {n}^3 float32 field (compensated float64 sum and NaN/Inf flags: "
                       "partial and fold kernels), best mean of 2 x 20 calls"}
    nint = ctypes.c_int32()
    _lib.check(lib.evx_fft_plan_trace_kinds(plan.handle, None, ctypes.byref(nint)), "trace kinds")
    kinds = (ctypes.c_int32 * max(1, nint.value))()
    _lib.check(lib.evx_fft_plan_trace_kinds(plan.handle, kinds, ctypes.byref(nint)), "trace kinds")
    kinds = list(kinds)[:nint.value]

    V = n ** 3
    stages = stage_bytes(n)
    times = np.array([[ev[k][i].elapsed_time(ev[k][i + 1]) for i in range(len(kinds))] for k in range(K)])
    kernels = []
    for kind, (name, bpv) in enumerate(stages):
        cols = [i for i, kk in enumerate(kinds) if kk == kind]
        nl = len(cols)
        if nl == 0:  # a kernel this step does not launch (the fused first sweep replaces two)
            continue
        ms = float(times[:, cols].sum(axis=1).mean())
        kernels.append({"kernel": name, "ms": round(ms, 4), "launches_per_step": nl,
                        "launch_ms": round(ms / nl, 4), "alg_bytes_per_voxel": round(bpv, 4),
                        "alg_bytes_per_launch": int(round(bpv * V / nl)),
                        "gbps": round(bpv * V / (ms * 1e-3) / 1e9, 1)})
    mean_ms = np.array([k["ms"] for k in kernels])
    dom = int(np.argmax(mean_ms))
    dk = kernels[dom]
    roof = {"bound": "hbm", "kernel": dk["kernel"], "achieved": dk["gbps"], "peak": PEAK_GBPS, "unit": "GB/s",
            "frac": round(dk["gbps"] / PEAK_GBPS, 4), "traffic": traffic_from_profiles(dk["kernel"]),
            "alg_bytes_per_launch": dk["alg_bytes_per_launch"], "launch_ms": dk["launch_ms"],
            "launches_per_step": dk["launches_per_step"]}
    b_alg = 4 * (3 + 8 * (1 + 2 / n))
    step_gbps = b_alg * V / (ms_step * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": V / (ms_step * 1e-3),
        "unit": "voxel-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"cahn_hilliard_semi_implicit_{n}^3_fp32", "grid": [n, n, n], "domain": [n, n, n],
                   "slabs": args.slabs,
                   "dt": 0.05, "eps": 3.0, "D": 1.0, "A": 0.25, "bc": "periodic",
                   "init": "0.5+0.05*N(0,1), numpy default_rng(2)", "parallelism": "single GPU"},
        "hbm_gbps": round(step_gbps, 1),
        "roofline": roof,
        "step_roofline": {"bound": "hbm", "achieved": round(step_gbps, 1), "peak": PEAK_GBPS, "unit": "GB/s",
                          "frac": round(step_gbps / PEAK_GBPS, 4), "alg_bytes_per_voxel": b_alg,
                          "actual_sweep_bytes_per_voxel": round(sum(k["alg_bytes_per_voxel"] for k in kernels), 4),
                          "kernels_per_step": len(kinds)},
        "kernels": kernels,
        "sum_kernel_ms": round(float(mean_ms.sum()), 4),
        "wall_ms_per_step": round(1e3 * wall / K, 4),
        "mass_rel_drift": abs(mass1 - mass0) / abs(mass0),
        "sum_f64": sum_f64,
    }
    # free-energy decay over the workload's 500 steps (frames every 50)
    if not args.no_energy:
        energy_decay_leg(out, vg, stepper, c0)
    if not args.no_batched:
        try:
            out["batched"] = batched_leg(args.batch_steps, 10, 0 if args.no_cpu_baseline else 2)
        except Exception as e:  # reported in the JSON line; the headline stands on its own
            out["batched"] = {"error": f"{type(e).__name__}: {e}"}
            log(f"[bench] batched leg failed: {e}")
        torch.cuda.empty_cache()
    if not args.no_configs:
        try:
            out["copy_ceiling"] = copy_ceiling()
            log(f"[bench] copy ceiling: {out['copy_ceiling']}")
            out["configs"] = configs_leg(not args.no_cpu_baseline)
        except Exception as e:  # reported in the JSON line; the headline stands on its own
            out["configs"] = {"error": f"{type(e).__name__}: {e}"}
            log(f"[bench] configs leg failed: {e}")
        torch.cuda.empty_cache()
    if not args.no_frames:
        try:
            out["frames"] = frames_leg(256, 100, 5)
        except Exception as e:  # reported in the JSON line; the headline stands on its own
            out["frames"] = {"error": f"{type(e).__name__}: {e}"}
            log(f"[bench] frames leg failed: {e}")
        torch.cuda.empty_cache()
    if not args.no_inversion:
        try:
            out["inversion"] = inversion_leg(args.inv_steps, args.inv_reps, args.inv_torch_reps)
        except Exception as e:  # reported in the JSON line; the headline stands on its own
            out["inversion"] = {"error": f"{type(e).__name__}: {e}"}
            log(f"[bench] inversion leg failed: {e}")
        torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        log(f"[bench] CPU baseline: oracle on {args.cpu_steps} steps (+1 warm-up)")
        out["cpu_baseline"] = cpu_baseline(c0.astype(np.float32), args.cpu_steps, step1)
        out["rel_l2_vs_fp64_oracle"] = out["cpu_baseline"].pop("rel_l2_vs_fp64_oracle")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
