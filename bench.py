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
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

METRIC = "voxel-steps/s + HBM GB/s (CH 512^3 FP32); fwd+adjoint step vs torch autograd"
PEAK_GBPS = 8000.0
ROOT = Path(__file__).resolve().parent


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def stage_bytes(n, s=4):
    """(kernel, algorithmic HBM bytes per voxel) for the trace intervals of a
    step: [0,1] RHS, [1,2] z R2C, [2,3] y forward, [3,4] x forward*prefac*inverse,
    [4,5] y inverse, [5,6] z C2R + add."""
    nh = n // 2 + 1
    spec = 2 * s * nh / n  # one half-spectrum array, per voxel
    return [
        ("ch_rhs_kernel", 2 * s),                    # read c, write rhs
        ("z_r2c_fast", s + spec),                    # read rhs, write spectrum
        ("strided_fast<y,fwd>", 2 * spec),           # read + write spectrum
        ("strided_fast<x,fwd*prefac*inv>", 2 * spec),
        ("strided_fast<y,inv>", 2 * spec),
        ("z_c2r_fast", spec + 2 * s),                # read spectrum + u, write u
    ]


def traffic_from_profiles(kernel):
    f = ROOT / "profiles" / "traffic_ch_imex_512_fp32.json"
    if not f.exists():
        return None
    d = json.loads(f.read_text())
    k = d.get("kernels", {}).get(kernel)
    return None if k is None else k.get("hbm_bytes_per_launch")


def cpu_baseline(c0, steps):
    """Times the CPU oracle on `steps` steps (after one warm-up step)."""
    import torch
    sys.path.insert(0, str(ROOT))
    from oracle import evoxels_oracle as O
    n = c0.shape[0]
    g = O.OracleGrid((n, n, n), (1.0, 1.0, 1.0), "float32")
    P3 = (("periodic", None),) * 3
    pre = O.imex_prefac(0.05, O.ch_symbol(O.rfft_k_squared((n, n, n), g.spacing), 3.0, 1.0))
    rhs = lambda t, v: O.ch_rhs(v, g, P3, 3.0, 1.0)  # noqa: E731
    u = g.field(c0)
    u = O.imex_step(u, rhs, pre)
    t0 = time.perf_counter()
    for _ in range(steps):
        u = O.imex_step(u, rhs, pre)
    dt = time.perf_counter() - t0
    return {"value": n ** 3 * steps / dt, "unit": "voxel-steps/s", "cores": torch.get_num_threads(),
            "kind": "port", "s_per_step": dt / steps,
            "sample": f"{steps} timed steps (after 1 warm-up) of the same 512^3 float32 workload, "
                      f"torch {torch.__version__} eager on CPU, os.cpu_count()={os.cpu_count()}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--n", type=int, default=512)
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

    mass0 = float(vg.channel_sums(u)[0])
    log(f"[bench] warmup {W} steps")
    for _ in range(W):
        stepper.step_into(u, u)
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(7)] for _ in range(K)]
    for row in ev:  # materialise the underlying hipEvent_t handles
        for e in row:
            e.record(stream)
    torch.cuda.synchronize()
    handles = [(ctypes.c_void_p * 7)(*[e.cuda_event for e in row]) for row in ev]
    log(f"[bench] timing {K} steps")
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    t_start.record(stream)
    for k in range(K):
        _lib.check(lib.evx_fft_plan_trace_next_step(plan.handle, handles[k], 7), "trace")
        stepper.step_into(u, u)
    t_end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms_step = t_start.elapsed_time(t_end) / K
    mass1 = float(vg.channel_sums(u)[0])
    log(f"[bench] {ms_step:.3f} ms/step (events), wall {1e3 * wall / K:.3f} ms/step")

    V = n ** 3
    stages = stage_bytes(n)
    times = np.array([[ev[k][i].elapsed_time(ev[k][i + 1]) for i in range(6)] for k in range(K)])
    mean_ms = times.mean(axis=0)
    kernels = []
    for (name, bpv), ms in zip(stages, mean_ms):
        kernels.append({"kernel": name, "ms": round(float(ms), 4), "alg_bytes_per_voxel": round(bpv, 4),
                        "alg_bytes_per_launch": int(round(bpv * V)),
                        "gbps": round(bpv * V / (ms * 1e-3) / 1e9, 1)})
    dom = int(np.argmax([k["ms"] for k in kernels]))
    dk = kernels[dom]
    roof = {"bound": "hbm", "kernel": dk["kernel"], "achieved": dk["gbps"], "peak": PEAK_GBPS, "unit": "GB/s",
            "frac": round(dk["gbps"] / PEAK_GBPS, 4), "traffic": traffic_from_profiles(dk["kernel"]),
            "alg_bytes_per_launch": dk["alg_bytes_per_launch"], "launch_ms": dk["ms"]}
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
                   "dt": 0.05, "eps": 3.0, "D": 1.0, "A": 0.25, "bc": "periodic",
                   "init": "0.5+0.05*N(0,1), numpy default_rng(2)", "parallelism": "single GPU"},
        "hbm_gbps": round(step_gbps, 1),
        "roofline": roof,
        "step_roofline": {"bound": "hbm", "achieved": round(step_gbps, 1), "peak": PEAK_GBPS, "unit": "GB/s",
                          "frac": round(step_gbps / PEAK_GBPS, 4), "alg_bytes_per_voxel": b_alg,
                          "actual_sweep_bytes_per_voxel": round(sum(b for _, b in stages), 4)},
        "kernels": kernels,
        "sum_kernel_ms": round(float(mean_ms.sum()), 4),
        "wall_ms_per_step": round(1e3 * wall / K, 4),
        "mass_rel_drift": abs(mass1 - mass0) / abs(mass0),
    }
    del u
    torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        log(f"[bench] CPU baseline: oracle on {args.cpu_steps} steps (+1 warm-up)")
        out["cpu_baseline"] = cpu_baseline(c0.astype(np.float32), args.cpu_steps)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
