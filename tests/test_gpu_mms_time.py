"""Manufactured-solution runs of the coupled reaction-diffusion problem with
the semi-implicit stepper on the GPU, against the slopes the reference's
notebook prints for the same study (10-appendix-convergence-testing.ipynb
cells 27-28, results5/results6).

The harness restates evoxels/utils.py:135-345 (``mms_convergence_test``) for
the 'hip' backend: grids 2^p per axis on the unit cube, the forcing built
from the sympy test functions ('spatial': du/dt - rhs_analytic; 'temporal':
du/dt evaluated exactly, minus the numerical RHS of the exact field), the
solver run to t = 1 with 8 frames, relative L2 errors of the final fields,
spatial slopes from the finest time step and temporal slopes from the
coarsest grid.  The reference ran the study with its JAX backend in float32;
the slopes agree to the stated tolerance (measured: 1e-4 spatial, 6e-4
temporal), not bit for bit."""
import json
from pathlib import Path

import numpy as np
import pytest
import sympy as sp
import sympy.vector as spv
import torch

import evoxels_amd as evx
from evoxels_amd.problem_definition import CoupledReactionDiffusion, ReactionDiffusion
from evoxels_amd.solvers import TimeDependentSolver
from evoxels_amd.timesteppers import ForwardEuler, PseudoSpectralIMEX
from evoxels_amd.voxelgrid import VoxelGridHip

pytestmark = pytest.mark.gpu
KAT = json.loads((Path(__file__).parent / "golden" / "reference_kats.json").read_text())["coupled_rd_imex_mms"]

CS = spv.CoordSys3D("CS")
t = sp.symbols("t", real=True)
FUNS = (0.5 + 0.4 * sp.cos(2 * sp.pi * CS.x) * sp.cos(2 * sp.pi * CS.y) ** 3 * sp.cos(t * 2 * sp.pi),
        0.4 + 0.1 * sp.sin(4 * sp.pi * CS.x) * sp.exp(-t / 10))


# the diffusion study of the same notebook (cells 21-23, results1-results4)
FUNS_RD = ((0.5 + 0.4 * sp.cos(4 * sp.pi * CS.x) * sp.cos(2 * sp.pi * CS.y) ** 3 * sp.cos(t * 2 * sp.pi))
           * sp.exp(-t),)


def mms_errors(mode, g_powers, t_powers, t_final=1.0, precision="float32", funs=FUNS,
               make_ode=lambda vg: CoupledReactionDiffusion(vg, D_A=1.0, D_B=0.5), stepper_cls=PseudoSpectralIMEX):
    """errors[j, k, i] for function j, time power k, grid power i.  The
    forcing is installed the way the reference's driver does it, by assigning
    ``ode.rhs`` on the instance (utils.py:285), so the steppers must take the
    replaced right-hand side instead of their fused kernels."""
    FUNS = funs
    args = (t, CS.x, CS.y, CS.z)
    u_t = [sp.lambdify(args, sp.N(sp.diff(f, t)), "torch") for f in FUNS]
    u_ex = [sp.lambdify(args, sp.N(f), "torch") for f in FUNS]
    errors = np.zeros((len(FUNS), len(t_powers), len(g_powers)))
    dx = np.zeros(len(g_powers))
    for i, p in enumerate(g_powers):
        n = 2 ** p
        vf = evx.VoxelFields((n, n, n), (1, 1, 1))
        vf.precision = precision
        dx[i] = vf.spacing[0]
        vg = VoxelGridHip(vf.grid_info(), precision=precision, device="cuda")
        ode = make_ode(vg)
        rhs_orig = ode.rhs
        grid = vg.meshgrid()
        if mode == "spatial":
            # one function: the scalar form, as utils.py:266-268 passes it
            an = ode.rhs_analytic(t, FUNS) if len(FUNS) > 1 else (ode.rhs_analytic(t, FUNS[0]),)
            rhs_an = [sp.lambdify(args, sp.N(f), "torch") for f in an]

            def mms_rhs(tt, u):
                rhs = rhs_orig(tt, u)
                t_ = vg.to_backend(tt)
                for j in range(len(FUNS)):
                    rhs[j] = rhs[j] - rhs_an[j](t_, *grid)
                    rhs[j] = rhs[j] + u_t[j](t_, *grid)
                return rhs
        else:
            def mms_rhs(tt, u):
                rhs = rhs_orig(tt, u)
                t_ = vg.to_backend(tt)
                ex = []
                for j in range(len(FUNS)):
                    ex.append(torch.broadcast_to(u_ex[j](t_, *grid), grid[0].shape)[None])
                    rhs[j] = rhs[j] + u_t[j](t_, *grid)
                rhs -= rhs_orig(tt, torch.cat(ex, 0).contiguous())
                return rhs
        ode.rhs = mms_rhs
        mesh = vf.meshgrid()
        for k, q in enumerate(t_powers):
            names = []
            for j, f in enumerate(FUNS):
                fun = sp.lambdify(args, f, "numpy")
                vf.add_field(f"u{j}", np.broadcast_to(fun(0, *mesh), vf.shape).copy())
                vf.add_field(f"u{j}_final", np.broadcast_to(fun(t_final, *mesh), vf.shape).copy())
                names.append(f"u{j}")
            dt = t_final / 2 ** q
            st = stepper_cls(ode, dt)
            assert not ode.fused, "a replaced rhs must not run on the fused kernels"
            solver = TimeDependentSolver(vf, names, backend="hip", step_fn=st.step, device="cuda")
            solver.solve(dt, 8, int(2 ** q), None, verbose=False)
            for j in range(len(FUNS)):
                exact = vf.fields[f"u{j}_final"]
                errors[j, k, i] = np.linalg.norm(vf.fields[f"u{j}"] - exact) / np.linalg.norm(exact)
        print(f"{mode} grid 2^{p}: errors {errors[:, :, i].tolist()}", flush=True)
    return dx, errors


def slope(x, y):
    return np.polyfit(np.log(x), np.log(y), 1)[0]


def test_coupled_rd_imex_spatial_slopes_match_notebook():
    """results5: spatial mode, slopes over grids 8^3..128^3 at dt = 1/1024."""
    dx, err = mms_errors("spatial", [3, 4, 5, 6, 7], [10])
    g = [slope(dx, err[j, -1, :]) for j in range(2)]
    print("spatial slopes", g, "notebook", KAT["spatial_slopes"], flush=True)
    np.testing.assert_allclose(g, KAT["spatial_slopes"], atol=2e-3)


def test_coupled_rd_imex_temporal_slopes_match_notebook():
    """results6: temporal mode, slopes over dt = 1/256..1/1024 on 8^3."""
    _, err = mms_errors("temporal", [3], [8, 9, 10])
    dts = np.array([1 / 256, 1 / 512, 1 / 1024])
    ts = [slope(dts, err[j, :, 0]) for j in range(2)]
    print("temporal slopes", ts, "notebook", KAT["temporal_slopes"], flush=True)
    np.testing.assert_allclose(ts, KAT["temporal_slopes"], atol=3e-3)


def _rd(vg):
    return ReactionDiffusion(vg, D=1.0)


@pytest.mark.parametrize("stepper_cls", [PseudoSpectralIMEX, ForwardEuler])
def test_rd_temporal_slope_with_replaced_rhs(stepper_cls):
    """results3/results4 (diffusion, temporal mode) on 8^3 in float64 (dt
    within forward Euler's stability limit dx^2/6): the forcing enters
    through an instance-level ``rhs``; both first-order steppers must show
    order 1 in time.  Dropping the forcing (the fused kernel's default
    right-hand side) leaves an O(1) error that does not converge."""
    _, err = mms_errors("temporal", [3], [9, 10, 11], precision="float64", funs=FUNS_RD, make_ode=_rd,
                        stepper_cls=stepper_cls)
    dts = np.array([1 / 512, 1 / 1024, 1 / 2048])
    ts = slope(dts, err[0, :, 0])
    print(stepper_cls.__name__, "temporal errors", err[0, :, 0], "slope", ts, flush=True)
    assert err[0, -1, 0] < 1e-2
    assert 0.9 < ts < 1.15


def test_rd_imex_spatial_slope_with_replaced_rhs():
    """results1 (diffusion, semi-implicit, spatial mode) on 8^3..32^3 at
    dt = 1/1024: second order in space."""
    dx, err = mms_errors("spatial", [3, 4, 5], [10], precision="float64", funs=FUNS_RD, make_ode=_rd)
    g = slope(dx, err[0, -1, :])
    print("RD IMEX spatial errors", err[0, -1, :], "slope", g, flush=True)
    assert 1.8 < g < 2.3
