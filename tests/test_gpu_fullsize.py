"""The semi-implicit Cahn-Hilliard step at the benchmark's own sizes.

* 512^3 float32 (periodic, domain 512, eps 3, D 1, dt 0.05, c0 = 0.5 +
  0.05*N(0,1) from numpy default_rng(2), the benchmark's input): one HIP step
  against one step of the float64 CPU oracle from the same float32 values
  (timesteppers.py:85-89, problem_definition.py:328-371).  Tolerances: field
  relative L2 <= 1e-5, update relative L2 <= 5e-4.
* Free-energy decay: F = int eps|grad c|^2 + 9/eps c^2(1-c)^2 dV (the
  functional of docs/notebooks/01-using-solvers.ipynb:32, whose variation is
  the chemical potential of problem_definition.py:304-305,353) must not grow
  from one frame to the next, over 500 steps of the 512^3 float32 run and
  over the 1000 steps of the 100^3 float64 run (dt 0.1, c0 = 0.5 + 0.1*U[0,1)
  from default_rng(0)); the mass drift is reported beside it.
"""
import numpy as np
import pytest
import torch

import evoxels_amd as evx
from evoxels_amd.problem_definition import CahnHilliard
from evoxels_amd.timesteppers import PseudoSpectralIMEX
from evoxels_amd.voxelgrid import VoxelGridHip
from oracle import evoxels_oracle as O

pytestmark = pytest.mark.gpu
P3 = (("periodic", None),) * 3


def headline_c0(n=512):
    return 0.5 + 0.05 * np.random.default_rng(2).standard_normal((n, n, n))


def setup(n, precision):
    vf = evx.VoxelFields((n, n, n), (n, n, n))
    vf.precision = precision
    vg = VoxelGridHip(vf.grid_info(), precision=precision, device="cuda")
    return vf, vg


def test_headline_512_step_matches_fp64_oracle():
    n, dt = 512, 0.05
    c32 = headline_c0(n).astype(np.float32)
    vf, vg = setup(n, "float32")
    st = PseudoSpectralIMEX(CahnHilliard(vg, eps=3.0, D=1.0), dt)
    u = vg.init_scalar_field(c32).contiguous()
    got = st.step(0.0, u).cpu().double()
    del u
    torch.cuda.empty_cache()
    print("512^3 HIP step done; running the float64 oracle step on the CPU", flush=True)
    og = O.OracleGrid((n, n, n), vf.spacing, "float64")
    pre = O.imex_prefac(dt, O.ch_symbol(O.rfft_k_squared((n, n, n), vf.spacing), 3.0, 1.0))
    u0 = og.field(c32.astype(np.float64))
    ref = O.imex_step(u0, lambda t, v: O.ch_rhs(v, og, P3, 3.0, 1.0), pre)
    ferr = float(torch.linalg.norm(got - ref) / torch.linalg.norm(ref))
    uerr = float(torch.linalg.norm((got - u0) - (ref - u0)) / torch.linalg.norm(ref - u0))
    print(f"512^3 fp32 step vs fp64 oracle: field rel-L2 {ferr:.3e}, update rel-L2 {uerr:.3e}", flush=True)
    assert ferr <= 1e-5
    assert uerr <= 5e-4


def energy_frames(vg, st, u, eps, steps, every):
    e = [float(vg.ch_energy(u, eps)[0])]
    m = [float(vg.channel_sums(u)[0])]
    for k in range(1, steps + 1):
        st.step_into(u, u)
        if k % every == 0:
            e.append(float(vg.ch_energy(u, eps)[0]))
            m.append(float(vg.channel_sums(u)[0]))
    return np.array(e), np.array(m)


@pytest.mark.parametrize("case", ["512_fp32_500_steps", "100_fp64_1000_steps"])
def test_free_energy_decays_frame_to_frame(case):
    if case.startswith("512"):
        n, precision, dt, steps, every = 512, "float32", 0.05, 500, 50
        c0 = headline_c0(n)
    else:
        n, precision, dt, steps, every = 100, "float64", 0.1, 1000, 100
        c0 = 0.5 + 0.1 * np.random.default_rng(0).random((n, n, n))
    vf, vg = setup(n, precision)
    st = PseudoSpectralIMEX(CahnHilliard(vg, eps=3.0, D=1.0), dt)
    u = vg.init_scalar_field(c0).contiguous()
    e, m = energy_frames(vg, st, u, 3.0, steps, every)
    de = np.diff(e)
    drift = float(np.abs(m - m[0]).max() / abs(m[0]))
    print(f"{case}: F per frame {e.tolist()}", flush=True)
    print(f"{case}: largest frame-to-frame change {de.max():.3e}, total {e[-1] - e[0]:.3e}, "
          f"mass drift {drift:.3e}", flush=True)
    assert np.all(np.isfinite(e))
    # non-increasing up to the rounding of the float64 energy sum itself
    assert np.all(de <= 1e-12 * np.abs(e[:-1]))
    assert e[-1] < e[0]
    assert drift <= (1e-6 if precision == "float32" else 1e-12)
