"""Parity of the HIP stencil kernels (pad, Laplacians, fused RD/CH/AC right-hand
sides and their Euler / RK4 epilogues) with the CPU oracle, through the C ABI."""
import numpy as np
import pytest
import torch

import evoxels_amd as evx
from evoxels_amd import _lib
from evoxels_amd.problem_definition import CahnHilliard, ReactionDiffusion, TwoPhaseAllenCahn
from evoxels_amd.timesteppers import ForwardEuler, RungeKutta4
from evoxels_amd.voxelgrid import VoxelGridHip
from oracle import evoxels_oracle as O

pytestmark = pytest.mark.gpu

P, N = ("periodic", None), ("neumann", None)
BCS = {
    "ppp": (P, P, P),
    "npp": (N, P, P),
    "dpp": (("dirichlet", (0.25, 0.75)), P, P),
    "n_d_p": (N, ("dirichlet", (0.0, 1.0)), P),  # the explicit-diffusion benchmark's BCs
    "nnn": (N, N, N),
    "d_n_d": (("dirichlet", (0.2, 0.9)), N, ("dirichlet", (0.1, 0.3))),
    "p_n_d": (P, N, ("dirichlet", (1.0, 0.0))),
}
SHAPES = [(5, 7, 9), (20, 19, 70), (9, 33, 130)]
TOL_SAME = {"float32": 2e-6, "float64": 1e-13}


def setup(shape, precision, seed, spacing=(0.9, 1.1, 1.3), lo=0.05, hi=0.95):
    vf = evx.VoxelFields(shape, tuple(s * n for s, n in zip(spacing, shape)))
    vf.precision = precision
    vg = VoxelGridHip(vf.grid_info(), precision=precision, device="cuda")
    og = O.OracleGrid(shape, vf.spacing, precision)
    rng = np.random.default_rng(seed)
    arr = lo + (hi - lo) * rng.random(shape)
    return vg, og, arr


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def quiet_problem(cls, vg, **kw):
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return cls(vg, **kw)


@pytest.mark.parametrize("precision", ["float32", "float64"])
@pytest.mark.parametrize("bcname", list(BCS))
def test_pad_bc_bitwise(precision, bcname):
    vg, og, arr = setup((6, 5, 67), precision, 0)
    bc = BCS[bcname]
    u = vg.init_scalar_field(arr)
    got = vg.pad_bc(u, bc).cpu()
    want = O.pad_bc(og.field(arr), bc)
    assert torch.equal(got, want), f"{bcname}: max diff {float((got - want).abs().max())}"


def test_pad_kat_reference_mixed_bc():
    """test_solvers.py:105-126 on the GPU kernel."""
    vf = evx.VoxelFields((2, 2, 2))
    vg = VoxelGridHip(vf.grid_info(), device="cuda")
    bc = (("dirichlet", (10.0, 20.0)), N, P)
    field = vg.init_scalar_field(np.arange(1, 9, dtype=np.float32).reshape(2, 2, 2))
    padded = vg.to_numpy(vg.pad_bc(field, bc))[0]
    expected = np.pad(np.arange(1, 9, dtype=np.float32).reshape(2, 2, 2), 1, mode="wrap")
    expected[0, :, :] = 2.0 * 10.0 - expected[1, :, :]
    expected[-1, :, :] = 2.0 * 20.0 - expected[-2, :, :]
    expected[:, 0, :] = expected[:, 1, :]
    expected[:, -1, :] = expected[:, -2, :]
    assert np.array_equal(padded, expected)


@pytest.mark.parametrize("precision", ["float32", "float64"])
def test_padded_laplacians(precision):
    vg, og, arr = setup((11, 13, 66), precision, 1)
    bc = BCS["d_n_d"]
    pu = O.pad_bc(og.field(arr), bc)
    pd = pu.to("cuda")
    lap = vg.laplace(pd).cpu()
    nl = vg.normal_laplace(pd).cpu()
    assert rel_l2(lap, O.laplace(pu, og.div_dx2)) < TOL_SAME[precision]
    assert rel_l2(nl, O.normal_laplace(pu, og.div_dx, og.div_dx2)) < 10 * TOL_SAME[precision]


def oracle_rhs(kind, og, bc, u, kw):
    if kind == "rd":
        return O.rd_rhs(u, og, bc, kw["D"])
    if kind == "ch":
        return O.ch_rhs(u, og, bc, kw["eps"], kw["D"])
    return O.ac_rhs(u, og, bc, **kw)


KINDS = {
    "rd": (ReactionDiffusion, {"D": 0.7}),
    "ch": (CahnHilliard, {"eps": 2.5, "D": 1.3}),
    "ac": (TwoPhaseAllenCahn, {"eps": 2.0, "gab": 1.2, "M": 0.8, "force": 0.3, "curvature": 0.2}),
}


@pytest.mark.parametrize("precision", ["float32", "float64"])
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("bcname", list(BCS))
@pytest.mark.parametrize("kind", list(KINDS))
def test_fused_rhs_matches_oracle(kind, bcname, shape, precision):
    cls, kw = KINDS[kind]
    vg, og, arr = setup(shape, precision, hash((kind, bcname)) % 1000, lo=-0.05, hi=1.05)
    bc = BCS[bcname]
    prob = quiet_problem(cls, vg, bc=bc, **kw)
    assert prob.fused
    got = prob.rhs(0.0, vg.init_scalar_field(arr)).cpu()
    want = oracle_rhs(kind, og, prob.bc, og.field(arr), kw)
    err = rel_l2(got, want)
    tol = TOL_SAME[precision] * (20 if kind == "ch" else 5)
    print(f"{kind} {bcname} {shape} {precision}: rel_l2 {err:.3e}")
    assert err < tol


@pytest.mark.parametrize("precision", ["float32", "float64"])
@pytest.mark.parametrize("kind", list(KINDS))
def test_generic_path_matches_fused(kind, precision):
    """A user closure equal to the default runs the generic HIP-operator path."""
    cls, kw = KINDS[kind]
    vg, og, arr = setup((10, 9, 20), precision, 7)
    bc = BCS["n_d_p"]
    fused = quiet_problem(cls, vg, bc=bc, **kw)
    if kind == "rd":
        generic = quiet_problem(cls, vg, bc=bc, f=lambda t, c, lib: 0 * c, **kw)
    elif kind == "ch":
        eps = kw["eps"]
        generic = quiet_problem(cls, vg, bc=bc, mu_hom=lambda c, lib=None: 18 / eps * c * (1 - c) * (1 - 2 * c), **kw)
    else:
        eps = kw["eps"]
        generic = quiet_problem(cls, vg, bc=bc, potential=lambda u, lib=None: 18 / eps * u * (1 - u) * (1 - 2 * u),
                                **kw)
    assert fused.fused and not generic.fused
    u = vg.init_scalar_field(arr)
    want = oracle_rhs(kind, og, fused.bc, og.field(arr), kw)
    assert rel_l2(generic.rhs(0.0, u).cpu(), want) < TOL_SAME[precision] * 20
    assert rel_l2(fused.rhs(0.0, u).cpu(), want) < TOL_SAME[precision] * 20


@pytest.mark.parametrize("precision", ["float32", "float64"])
@pytest.mark.parametrize("kind", list(KINDS))
def test_euler_and_rk4_epilogues(kind, precision):
    cls, kw = KINDS[kind]
    vg, og, arr = setup((12, 18, 66), precision, 11)
    bc = BCS["d_n_d"]
    prob = quiet_problem(cls, vg, bc=bc, **kw)
    dt = 0.01
    u = vg.init_scalar_field(arr)
    rhs = lambda t, v: oracle_rhs(kind, og, prob.bc, v, kw)  # noqa: E731
    u64 = O.OracleGrid(og.shape, og.spacing, "float64").field(arr)
    og64 = O.OracleGrid(og.shape, og.spacing, "float64")
    rhs64 = lambda t, v: oracle_rhs(kind, og64, prob.bc, v, kw)  # noqa: E731
    for stepper, ostep in ((ForwardEuler, O.fe_step), (RungeKutta4, O.rk4_step)):
        got = stepper(prob, dt).step(0.0, u).cpu()
        same = ostep(og.field(arr), rhs, dt)
        ref64 = ostep(u64, rhs64, dt)
        assert rel_l2(got, same) < TOL_SAME[precision]
        # per-step field error against the float64 oracle (1e-5 f32, 1e-6 f64)
        assert rel_l2(got, ref64) < (1e-5 if precision == "float32" else 1e-6)


def test_explicit_diffusion_benchmark_config_steps():
    """256^3 float32, Neumann x / Dirichlet(0,1) y / periodic z, dt = 0.1*dx^2/6,
    phi0 = U[0,1) seed 1: three forward-Euler steps against the float64 oracle."""
    n = 256
    vf = evx.VoxelFields((n, n, n))
    vg = VoxelGridHip(vf.grid_info(), precision="float32", device="cuda")
    bc = (N, ("dirichlet", (0, 1)), P)
    prob = quiet_problem(ReactionDiffusion, vg, D=1.0, bc=bc)
    dt = 0.1 * vf.spacing[0] ** 2 / 6
    arr = np.random.default_rng(1).random((n, n, n))
    og64 = O.OracleGrid((n, n, n), vf.spacing, "float64")
    u_ref = og64.field(arr)
    u = vg.init_scalar_field(arr)
    step = ForwardEuler(prob, dt)
    for i in range(3):
        u = step.step(0.0, u)
        u_ref = O.fe_step(u_ref, lambda t, v: O.rd_rhs(v, og64, prob.bc, 1.0), dt)
        err = rel_l2(u.cpu(), u_ref)
        print(f"step {i + 1}: rel_l2 vs float64 oracle {err:.3e}")
        assert err < 1e-5


@pytest.mark.parametrize("precision", ["float32", "float64"])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_ch_rhs_face_mobility_kat_by_hand(axis, precision):
    """The hand-worked CH profile of the oracle pins (0/1 values with -0.1 and
    1.1 entries, face mobility cbar(1 - cbar), clip first) through the fused
    HIP right-hand side, exactly, and through the semi-implicit step, whose
    contracted periodic right-hand side must reproduce it to rounding."""
    from tests.test_oracle_pins import ch_kat_case
    shape, c, want = ch_kat_case(axis, other=4)
    vf = evx.VoxelFields(shape, tuple(float(n) for n in shape))
    vg = VoxelGridHip(vf.grid_info(), precision=precision, device="cuda")
    prob = CahnHilliard(vg, eps=3.0, D=1.0)
    got = prob.rhs(0.0, vg.init_scalar_field(c)).cpu()[0].double().numpy()
    np.testing.assert_array_equal(got, want)
    from evoxels_amd.timesteppers import PseudoSpectralIMEX
    dt = 0.05
    step = PseudoSpectralIMEX(prob, dt).step(0.0, vg.init_scalar_field(c)).cpu()
    og = O.OracleGrid(shape, vf.spacing, precision)
    pre = O.imex_prefac(dt, O.ch_symbol(O.rfft_k_squared(shape, vf.spacing), 3.0, 1.0))
    hand = torch.tensor(want[None], dtype=og.dtype)
    ref = O.imex_step(og.field(c), lambda t, v: hand, pre)
    upd, upd_ref = (step - og.field(c)).double(), (ref - og.field(c)).double()
    err = float(torch.linalg.norm(upd - upd_ref) / torch.linalg.norm(upd_ref))
    assert err < (5e-6 if precision == "float32" else 1e-13), err
