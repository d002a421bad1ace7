"""Solver-level parity on the GPU: compensated reductions, the frame loop
(with and without hipGraph capture), precompiled entry points, NaN abort,
the reference's solver tests and MMS convergence slopes of the HIP kernels."""
import math

import numpy as np
import pytest
import sympy as sp
import torch

import evoxels_amd as evx
from evoxels_amd.problem_definition import CahnHilliard, ReactionDiffusion, TwoPhaseAllenCahn
from evoxels_amd.solvers import TimeDependentSolver
from evoxels_amd.timesteppers import ForwardEuler, PseudoSpectralIMEX, RungeKutta4
from evoxels_amd.voxelgrid import VoxelGridHip
from oracle import evoxels_oracle as O
from tests import mms

pytestmark = pytest.mark.gpu
P3 = (("periodic", None),) * 3
N3 = (("neumann", None),) * 3


# ------------------------------------------------------------ reductions --
@pytest.mark.parametrize("precision", ["float32", "float64"])
def test_compensated_sum_and_flags(precision):
    vf = evx.VoxelFields((33, 40, 50))
    vg = VoxelGridHip(vf.grid_info(), precision=precision, device="cuda")
    rng = np.random.default_rng(5)
    x = rng.standard_normal((3, 33, 40, 50)) * 10.0 ** rng.integers(-8, 8, (3, 33, 40, 50))
    xt = torch.tensor(x, dtype=vg.precision, device="cuda")
    flags = torch.zeros(2, dtype=torch.int32, device="cuda")
    s = vg.channel_sums(xt, flags).cpu().numpy()
    xs = xt.cpu().double().numpy()
    exact = [math.fsum(xs[c].ravel()) for c in range(3)]
    np.testing.assert_allclose(s, exact, rtol=1e-15, atol=1e-15 * np.abs(xs).max())
    assert flags.tolist() == [0, 0]
    xt[1, 3, 4, 5] = float("inf")
    xt[2, 0, 0, 0] = float("nan")
    vg.channel_sums(xt, flags)
    assert flags.tolist() == [1, 1]
    np.testing.assert_allclose(vg.average(torch.tensor(x[:1], dtype=vg.precision, device="cuda")).cpu().numpy(),
                               [exact[0] / x[0].size], rtol=1e-6 if precision == "float32" else 1e-14)


@pytest.mark.parametrize("precision", ["float32", "float64"])
@pytest.mark.parametrize("shape", [(3, 7, 9, 11), (2, 1, 1, 1), (3, 1, 1, 3), (4, 1, 5, 1), (2, 64, 64, 65)])
def test_compensated_sum_ragged_channels(precision, shape):
    """Channel lengths that are not a multiple of the 16-byte vector width:
    every channel but the first starts off the vector boundary, so the sum
    runs through its head, vector and tail parts; NaN/Inf flags from the head
    and the tail elements."""
    vf = evx.VoxelFields(shape[1:])
    vg = VoxelGridHip(vf.grid_info(), precision=precision, device="cuda")
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape) * 10.0 ** rng.integers(-6, 6, shape)
    xt = torch.tensor(x, dtype=vg.precision, device="cuda")
    flags = torch.zeros(2, dtype=torch.int32, device="cuda")
    s = vg.channel_sums(xt, flags).cpu().numpy()
    xs = xt.cpu().double().numpy()
    exact = [math.fsum(xs[c].ravel()) for c in range(shape[0])]
    np.testing.assert_allclose(s, exact, rtol=1e-15, atol=1e-15 * np.abs(xs).max())
    assert flags.tolist() == [0, 0]
    flat = xt.view(shape[0], -1)
    flat[1, 0] = float("nan")
    vg.channel_sums(xt, flags)
    assert flags.tolist() == [1, 0]
    flat[1, 0] = 1.0
    flags.zero_()
    flat[-1, -1] = float("-inf")
    vg.channel_sums(xt, flags)
    assert flags.tolist() == [0, 1]


@pytest.mark.parametrize("precision", ["float32", "float64"])
def test_ch_energy_matches_oracle(precision):
    shape = (20, 17, 66)
    vf = evx.VoxelFields(shape, (20.0, 17.0, 66.0))
    vg = VoxelGridHip(vf.grid_info(), precision=precision, device="cuda")
    arr = 0.5 + 0.1 * np.random.default_rng(6).standard_normal(shape)
    og = O.OracleGrid(shape, vf.spacing, precision)
    for bc in (P3, (("dirichlet", (0.1, 0.9)), ("neumann", None), ("periodic", None))):
        got = float(vg.ch_energy(vg.init_scalar_field(arr), 3.0, bc)[0])
        want = float(O.ch_energy(og.field(arr), og, bc, 3.0)[0])
        assert abs(got - want) < 1e-12 * abs(want)


# ---------------------------------------------------------------- loop --
def test_time_solver_multiple_fields():
    """test_solvers.py:14-27 on the 'hip' backend."""
    vf = evx.VoxelFields((4, 4, 4))
    vf.add_field("a", np.ones(vf.shape))
    vf.add_field("b", np.zeros(vf.shape))
    solver = TimeDependentSolver(vf, ["a", "b"], backend="hip", step_fn=lambda t, u: u + 1, device="cuda")
    solver.solve(frames=1, max_iters=1, verbose=False, jit=False)
    assert np.allclose(vf.fields["a"], 2) and np.allclose(vf.fields["b"], 1)


def _oracle_frames(arr, step, dt, frames, iters):
    out = []
    O.run_loop(arr, step, dt, frames, iters, lambda v, f, t: out.append(v[0].numpy().copy()))
    return out


@pytest.mark.parametrize("jit", [False, True])
def test_cahn_hilliard_solver_frames_match_oracle(jit):
    n, dt, iters, frames = 24, 0.1, 30, 4
    vf = evx.VoxelFields((n, n, n), (n, n, n))
    vf.precision = "float64"
    c0 = 0.5 + 0.1 * np.random.default_rng(0).random((n, n, n))
    vf.add_field("c", c0.copy())
    seen = []
    orig = TimeDependentSolver._export_fields

    def spy(self, u_out):
        orig(self, u_out)
        seen.append(self.vf.fields["c"].copy())

    TimeDependentSolver._export_fields = spy
    try:
        evx.run_cahn_hilliard_solver(vf, "c", "hip", jit=jit, time_increment=dt, frames=frames,
                                     max_iters=iters, verbose=False)
    finally:
        TimeDependentSolver._export_fields = orig
    og = O.OracleGrid((n, n, n), vf.spacing, "float64")
    pre = O.imex_prefac(dt, O.ch_symbol(O.rfft_k_squared((n, n, n), vf.spacing), 3.0, 1.0))
    ref = _oracle_frames(og.field(c0), lambda t, v: O.imex_step(v, lambda tt, w: O.ch_rhs(w, og, P3), pre),
                         dt, frames, iters)
    # n_out = 30 // 4 = 7: outputs at i = 0, 7, 14, 21, 28 and a final one
    assert len(seen) == len(ref) == 6
    for a, b in zip(seen, ref):
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-12


def test_second_solve_recaptures_with_new_dt():
    """Two solve() calls on one solver with different time increments and
    parameters: with jit=True each solve records its own graph (the kernel
    arguments hold dt, eps and the stepper's workspace), so both results equal
    the uncaptured runs."""
    n = 20
    c0 = 0.5 + 0.1 * np.random.default_rng(11).random((n, n, n))
    results = {}
    for jit in (False, True):
        vf = evx.VoxelFields((n, n, n), (n, n, n))
        vf.precision = "float64"
        vf.add_field("c", c0.copy())
        solver = TimeDependentSolver(vf, "c", backend="hip", problem_cls=CahnHilliard,
                                     timestepper_cls=PseudoSpectralIMEX, device="cuda")
        out = []
        for dt, eps in ((0.1, 3.0), (0.05, 2.0), (0.1, 3.0)):
            vf.set_field("c", c0.copy())
            solver.solve(time_increment=dt, frames=2, max_iters=6, problem_kwargs={"eps": eps}, jit=jit,
                         verbose=False)
            out.append(vf.fields["c"].copy())
        results[jit] = out
    assert not np.array_equal(results[False][0], results[False][1])
    for a, b in zip(results[False], results[True]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("jit", [False, True])
def test_allen_cahn_solver_frames_match_oracle(jit):
    shape, dt, iters, frames = (20, 18, 70), 0.2, 12, 3
    vf = evx.VoxelFields(shape, shape)
    phi0 = np.random.default_rng(3).random(shape).astype(np.float32)
    vf.add_field("phi", phi0.copy())
    evx.run_allen_cahn_solver(vf, "phi", "hip", jit=jit, time_increment=dt, frames=frames, max_iters=iters,
                              verbose=False)
    og = O.OracleGrid(shape, vf.spacing, "float32")
    ref = _oracle_frames(og.field(phi0), lambda t, v: O.rk4_step(v, lambda tt, w: O.ac_rhs(w, og, N3), dt, t),
                         dt, frames, iters)
    got = vf.fields["phi"]
    assert np.linalg.norm(got - ref[-1]) / np.linalg.norm(ref[-1]) < 1e-5


def test_nan_aborts_like_reference():
    vf = evx.VoxelFields((4, 4, 4))
    vf.add_field("a", np.ones(vf.shape))
    solver = TimeDependentSolver(vf, "a", backend="hip", step_fn=lambda t, u: u * float("nan"), device="cuda")
    with pytest.raises(SystemExit):
        solver.solve(frames=1, max_iters=2, verbose=False, jit=False)


@pytest.mark.parametrize("async_output", [False, True])
def test_nan_abort_frame_and_fields(capsys, async_output):
    """The abort names the first frame holding a NaN, and vf.fields holds that
    frame, with synchronous and asynchronous export alike."""
    vf = evx.VoxelFields((6, 5, 4))
    vf.add_field("a", np.ones(vf.shape))

    def step(t, u):
        return u * float("nan") if t >= 0.35 else u + 1
    solver = TimeDependentSolver(vf, "a", backend="hip", step_fn=step, device="cuda", async_output=async_output)
    with pytest.raises(SystemExit):
        solver.solve(time_increment=0.1, frames=5, max_iters=10, verbose=False, jit=False)
    # outputs at i = 0, 2, 4, 6, 8: the NaN appears after the step at t = 0.4
    assert "NaN detected in frame 3 at time 0.6000000000000001" in capsys.readouterr().out
    assert np.isnan(vf.fields["a"]).all()


@pytest.mark.parametrize("jit", [False, True])
def test_async_output_equals_synchronous(tmp_path, monkeypatch, jit):
    """Frames, mass history and VTK files of the asynchronous export are those
    of the synchronous loop, byte for byte."""
    shape, dt, iters, frames = (32, 24, 40), 0.05, 20, 5
    c0 = 0.5 + 0.05 * np.random.default_rng(4).standard_normal(shape)
    results = {}
    orig = TimeDependentSolver._export_fields
    for mode in (False, True):
        vf = evx.VoxelFields(shape, shape)
        vf.add_field("c", c0.copy())
        seen = []

        def spy(self, u_out):
            orig(self, u_out)
            seen.append(self.vf.fields["c"].copy())
        monkeypatch.setattr(TimeDependentSolver, "_export_fields", spy)
        d = tmp_path / f"async_{mode}"
        d.mkdir()
        monkeypatch.chdir(d)
        solver = TimeDependentSolver(vf, "c", backend="hip", problem_cls=CahnHilliard,
                                     timestepper_cls=PseudoSpectralIMEX, device="cuda", async_output=mode)
        solver.solve(time_increment=dt, frames=frames, max_iters=iters, problem_kwargs={"eps": 3.0, "D": 1.0},
                     jit=jit, verbose=False, vtk_out=True)
        files = {f.name: f.read_bytes() for f in sorted(d.iterdir())}
        results[mode] = (seen, solver.mass_history, files)
    (s0, m0, f0), (s1, m1, f1) = results[False], results[True]
    assert len(s0) == len(s1) == 6
    assert all(np.array_equal(a, b) for a, b in zip(s0, s1))
    assert [(f, t) for f, t, _ in m0] == [(f, t) for f, t, _ in m1]
    assert all(np.array_equal(a, b) for (_, _, a), (_, _, b) in zip(m0, m1))
    assert sorted(f0) == sorted(f1) == [f"CahnHilliard_c_{k:03d}.vtk" for k in range(6)]
    assert all(f0[k] == f1[k] for k in f0)


def test_inf_does_not_abort():
    """The reference checks isnan only (solvers.py:163)."""
    vf = evx.VoxelFields((4, 4, 4))
    vf.add_field("a", np.ones(vf.shape))
    solver = TimeDependentSolver(vf, "a", backend="hip", step_fn=lambda t, u: u * float("inf"), device="cuda")
    solver.solve(frames=1, max_iters=1, verbose=False, jit=False)
    assert np.isinf(vf.fields["a"]).all()


def test_allen_cahn_tanh_profile():
    """test_solvers.py:30-82, Allen-Cahn leg, on the 'hip' backend."""
    nx, eps = 16, 3.0
    vf = evx.VoxelFields((nx, 1, 1), domain_size=(nx, 1, 1))
    phi = np.zeros((nx, 1, 1), dtype=np.float32)
    phi[: nx // 2] = 1.0
    vf.add_field("phi1", phi.copy())
    evx.run_allen_cahn_solver(vf, "phi1", backend="hip", frames=1, max_iters=10, time_increment=0.5, eps=eps,
                              jit=False, verbose=False)
    x = np.arange(nx) + 0.5
    ana = 0.5 - 0.5 * np.tanh(3 * (x - 0.5 * nx) / 2 / eps)
    assert np.linalg.norm(vf.fields["phi1"].squeeze() - ana) < 0.05


def test_cahn_hilliard_constant_is_fixed_point():
    """test_inversion.py:23-34 idea: a constant 0.5 field stays 0.5 under CH."""
    vf = evx.VoxelFields((4, 4, 4))
    vf.add_field("c", np.full((4, 4, 4), 0.5, dtype=np.float32))
    evx.run_cahn_hilliard_solver(vf, "c", "hip", frames=1, max_iters=2, time_increment=0.1, verbose=False)
    np.testing.assert_allclose(vf.fields["c"], 0.5, atol=1e-6)


# ------------------------------------------------------ MMS on the GPU --
def _hip_rhs(cls, bc, precision, **kw):
    def f(arr, n):
        vf = evx.VoxelFields((n, n, n), (1, 1, 1))
        vf.precision = precision
        vg = VoxelGridHip(vf.grid_info(), precision=precision, device="cuda")
        import warnings
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            prob = cls(vg, bc=bc, **kw)
        return prob.rhs(0.0, vg.init_scalar_field(arr))[0].cpu().numpy()
    return f


def test_mms_slopes_of_hip_kernels():
    CS = mms.CS
    e = sp.sin(2 * sp.pi * CS.x) * sp.sin(4 * sp.pi * CS.y) * sp.sin(6 * sp.pi * CS.z) / (56 * sp.pi ** 2)
    _, _, s = mms.convergence(e, mms.rd_analytic(e), _hip_rhs(ReactionDiffusion, P3, "float32", D=1.0), "float32")
    assert abs(s - 2) < 0.1
    e = sp.cos(sp.pi * CS.x) ** 3
    bcn = (("neumann", None), ("periodic", None), ("periodic", None))
    _, _, s = mms.convergence(e, mms.rd_analytic(e), _hip_rhs(ReactionDiffusion, bcn, "float32", D=1.0), "float32")
    assert abs(s - 1.9533270430146406) < 2e-3
    e = 0.4 + 0.1 * sp.sin(2 * sp.pi * CS.x)
    _, _, s = mms.convergence(e, mms.ch_analytic(e), _hip_rhs(CahnHilliard, P3, "float64", eps=3.0, D=1.0))
    assert abs(s - 2) < 0.1
    e = 0.5 + 0.3 * sp.cos(4 * sp.pi * CS.x) * sp.cos(2 * sp.pi * CS.y) * (CS.z ** 2 / 2 - CS.z ** 3 / 3)
    _, errs, s = mms.convergence(e, mms.ac_analytic(e, eps=3.0, curvature=0.5),
                                 _hip_rhs(TwoPhaseAllenCahn, N3, "float64", eps=3.0, curvature=0.5))
    np.testing.assert_allclose(errs, [0.19028211, 0.05396869, 0.01346211, 0.00334846, 0.00102839], rtol=5e-6)
    assert abs(s - 1.907376693930398) < 1e-5


@pytest.mark.parametrize("precision", ["float32", "float64"])
def test_step_records_give_the_mass_and_nan_flag(precision):
    """The semi-implicit step's final sweep leaves per-block compensated sums
    of the state it writes: folded, they equal the standalone reduction of
    that state (to float64 rounding), also after graph replay, and flag NaN."""
    shape = (24, 32, 512) if precision == "float32" else (20, 16, 64)
    vf = evx.VoxelFields(shape, shape)
    vf.precision = precision
    vg = VoxelGridHip(vf.grid_info(), precision=precision, device="cuda")
    st = PseudoSpectralIMEX(CahnHilliard(vg, eps=3.0, D=1.0), 0.05)
    u = vg.init_scalar_field(0.5 + 0.05 * np.random.default_rng(21).standard_normal(shape)).contiguous()
    v = st.step(0.0, u)
    flags = torch.zeros(2, dtype=torch.int32, device="cuda")
    rec = st.last_step_sums(v, flags)
    assert rec is not None and flags.tolist() == [0, 0]
    ref = vg.channel_sums(v)
    exact = math.fsum(v.cpu().double().numpy().ravel())
    print(f"{precision}: records {float(rec[0])!r}, reduction {float(ref[0])!r}, fsum {exact!r}", flush=True)
    assert abs(float(rec[0]) - exact) <= 1e-14 * abs(exact)
    assert abs(float(ref[0]) - exact) <= 1e-14 * abs(exact)
    g = torch.cuda.CUDAGraph()
    w = u.clone()
    st.step_into(w, w)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.graph(g, stream=side):
        st.step_into(w, w)
    for _ in range(3):
        g.replay()
        r = st.last_step_sums(w)
        assert abs(float(r[0]) - float(vg.channel_sums(w)[0])) <= 1e-14 * abs(float(r[0]))
    bad = u.clone()
    bad.view(-1)[12345] = float("nan")
    out = st.step(0.0, bad)
    flags.zero_()
    st.last_step_sums(out, flags)
    assert flags.tolist()[0] == 1
    # several channels keep no records
    u2 = torch.cat([u, u], 0)
    st2 = PseudoSpectralIMEX(CahnHilliard(vg, eps=3.0, D=1.0), 0.05)
    assert st2.last_step_sums(st2.step(0.0, u2)) is None


def test_solver_mass_history_from_step_records():
    """The solver's per-frame mass comes from the step records after the
    first frame: equal to the reduction of the exported frames."""
    n = 32
    vf = evx.VoxelFields((16, n, 512), (16, n, 512))
    c0 = 0.5 + 0.05 * np.random.default_rng(22).standard_normal((16, n, 512))
    vf.add_field("c", c0.astype(np.float32))
    frames = []
    orig = TimeDependentSolver._export_fields

    def spy(self, u_out):
        orig(self, u_out)
        frames.append(self.vf.fields["c"].astype(np.float64).copy())

    TimeDependentSolver._export_fields = spy
    try:
        solver = TimeDependentSolver(vf, "c", backend="hip", problem_cls=CahnHilliard,
                                     timestepper_cls=PseudoSpectralIMEX, device="cuda")
        solver.solve(time_increment=0.05, frames=4, max_iters=8, jit=True, verbose=False)
    finally:
        TimeDependentSolver._export_fields = orig
    masses = [float(m[0]) for _, _, m in solver.mass_history]
    assert len(masses) == len(frames) == 5
    for m, f in zip(masses, frames):
        assert abs(m - math.fsum(f.ravel())) <= 1e-14 * abs(m)
