"""Parity of the multiphase Allen-Cahn, two-species reaction-diffusion and
smoothed-boundary reaction-diffusion problems with the CPU oracle: the fused
two-species kernel (with its Euler / RK4 epilogues), the generic paths (HIP
pad and Laplacian kernels with device tensor algebra), and their explicit,
semi-implicit and exponential steps."""
import warnings

import numpy as np
import pytest
import torch

import evoxels_amd as evx
from evoxels_amd.problem_definition import CoupledReactionDiffusion, MultiPhaseAllenCahn, ReactionDiffusionSBM
from evoxels_amd.timesteppers import ExponentialEuler, ForwardEuler, PseudoSpectralIMEX, RungeKutta4
from evoxels_amd.voxelgrid import VoxelGridHip
from oracle import evoxels_oracle as O

pytestmark = pytest.mark.gpu

P, N = ("periodic", None), ("neumann", None)
BCS = {
    "ppp": (P, P, P),
    "npp": (N, P, P),
    "d_n_p": (("dirichlet", (0.3, 0.8)), N, P),
    "nnn": (N, N, N),
}
# same operations in the same order: float32 agreement is to a few roundings
TOL = {"float32": 4e-6, "float64": 1e-12}


def setup(shape, precision, spacing=(0.9, 1.1, 1.3)):
    vf = evx.VoxelFields(shape, tuple(s * n for s, n in zip(spacing, shape)))
    vf.precision = precision
    vg = VoxelGridHip(vf.grid_info(), precision=precision, device="cuda")
    og = O.OracleGrid(shape, vf.spacing, precision)
    return vf, vg, og


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def problem(cls, vg, **kw):
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return cls(vg, **kw)


def phases(shape, n, seed):
    """Phase fractions near the simplex with some negative entries, so that
    both projections have work to do; phase 0 stays positive (an all-zero
    voxel is 0/0 in the sloppy projection, as in the reference)."""
    rng = np.random.default_rng(seed)
    a = rng.dirichlet([1.5] * n, size=shape).transpose(3, 0, 1, 2) + 0.08 * rng.standard_normal((n,) + shape)
    a[0] = np.abs(a[0]) + 0.02
    return a


# ------------------------------------------------------- multiphase Allen-Cahn
@pytest.mark.parametrize("precision", ["float32", "float64"])
@pytest.mark.parametrize("fast", [True, False])
@pytest.mark.parametrize("bcname,curvature", [("ppp", 1.0), ("npp", 1.0), ("npp", 0.5), ("nnn", 0.0)])
def test_multiphase_rhs_matches_oracle(precision, fast, bcname, curvature):
    shape = (9, 14, 67)
    _, vg, og = setup(shape, precision)
    arr = phases(shape, 3, 4)
    bc = BCS[bcname]
    p = problem(MultiPhaseAllenCahn, vg, eps=2.5, gab=1.2, M=0.8, curvature=curvature, fast=fast, bc=bc)
    got = p.rhs(0.0, torch.tensor(arr, dtype=vg.precision, device="cuda")).cpu().numpy()
    want = O.multiphase_ac_rhs(og.field(arr)[0], og, bc, eps=2.5, gab=1.2, M=0.8, curvature=curvature,
                               fast=fast).numpy()
    err = rel_l2(got, want)
    print(f"multiphase {precision} fast={fast} {bcname} curv={curvature}: rel l2 {err:.2e}", flush=True)
    assert err < TOL[precision]
    assert float(np.abs(got.sum(0)).max()) < (1e-5 if precision == "float32" else 1e-12) * np.abs(got).max()


def test_multiphase_four_phases_imex_step():
    """Four phases, periodic, semi-implicit step with the symbol -M gab k^2
    (problem_definition.py:463)."""
    shape = (12, 16, 20)
    _, vg, og = setup(shape, "float64", (1.0, 1.0, 1.0))
    arr = phases(shape, 4, 6)
    bc = BCS["ppp"]
    p = problem(MultiPhaseAllenCahn, vg, eps=3.0, gab=1.0, M=1.0, bc=bc)
    st = PseudoSpectralIMEX(p, 0.05)
    got = st.step(0.0, torch.tensor(arr, dtype=torch.float64, device="cuda")).cpu().numpy()
    k2 = O.rfft_k_squared(shape, og.spacing)
    pre = O.imex_prefac(0.05, O.ac_symbol(k2, 1.0, 1.0))
    want = O.imex_step(og.field(arr)[0], lambda t, x: O.multiphase_ac_rhs(x, og, bc, eps=3.0), pre).numpy()
    assert rel_l2(got - arr, want - arr) < 1e-11


# -------------------------------------------------- two-species (Gray-Scott)
def coupled(vg, bc, **kw):
    """CoupledReactionDiffusion has no ``bc`` field (problem_definition.py:
    576-595): it is periodic unless the attribute is set before the BCs are
    initialised again, which exercises the kernel's ghost rules."""
    p = problem(CoupledReactionDiffusion, vg, **kw)
    if bc != BCS["ppp"]:
        p.bc = bc
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            p.initialize_boundary_conditions()
    return p


def species(shape, seed):
    rng = np.random.default_rng(seed)
    return np.stack([0.6 + 0.3 * rng.random(shape), 0.1 + 0.4 * rng.random(shape)])


@pytest.mark.parametrize("precision", ["float32", "float64"])
@pytest.mark.parametrize("bcname", list(BCS))
@pytest.mark.parametrize("shape", [(5, 7, 9), (20, 19, 70), (9, 33, 130)])
def test_coupled_rd_fused_rhs_matches_oracle(precision, bcname, shape):
    _, vg, og = setup(shape, precision)
    arr = species(shape, 2)
    bc = BCS[bcname]
    p = coupled(vg, bc, D_A=0.9, D_B=0.45, feed=0.04, kill=0.1)
    assert p.fused
    got = p.rhs(0.0, torch.tensor(arr, dtype=vg.precision, device="cuda")).cpu().numpy()
    want = O.coupled_rd_rhs(og.field(arr)[0], og, bc, 0.9, 0.45, 0.04, 0.1).numpy()
    err = rel_l2(got, want)
    print(f"gray-scott {precision} {bcname} {shape}: rel l2 {err:.2e}", flush=True)
    assert err < TOL[precision]


@pytest.mark.parametrize("precision", ["float32", "float64"])
def test_coupled_rd_generic_equals_fused(precision):
    """A user interaction equal to the default takes the generic path; both
    paths give the same right-hand side."""
    shape = (11, 13, 66)
    _, vg, _ = setup(shape, precision)
    u = torch.tensor(species(shape, 3), dtype=vg.precision, device="cuda")
    fused = coupled(vg, BCS["d_n_p"])
    generic = coupled(vg, BCS["d_n_p"], interaction=lambda v: v[0] * v[1] ** 2)
    assert fused.fused and not generic.fused
    assert rel_l2(generic.rhs(0.0, u).cpu(), fused.rhs(0.0, u).cpu()) < TOL[precision]


@pytest.mark.parametrize("precision", ["float32", "float64"])
@pytest.mark.parametrize("stepper", ["euler", "rk4"])
def test_coupled_rd_fused_steps(precision, stepper):
    shape = (16, 18, 40)
    _, vg, og = setup(shape, precision)
    arr = species(shape, 5)
    bc = BCS["npp"]
    p = coupled(vg, bc)
    dt = 0.2
    st = (ForwardEuler if stepper == "euler" else RungeKutta4)(p, dt)
    u = torch.tensor(arr, dtype=vg.precision, device="cuda")
    for _ in range(3):
        u = st.step(0.0, u)
    rhs = lambda t, x: O.coupled_rd_rhs(x, og, bc)  # noqa: E731
    ur = og.field(arr)[0]
    for _ in range(3):
        ur = (O.fe_step if stepper == "euler" else O.rk4_step)(ur, rhs, dt)
    err = rel_l2(u.cpu().numpy() - arr, ur.numpy() - arr)
    print(f"gray-scott {stepper} {precision}: rel l2 of the update {err:.2e}", flush=True)
    assert err < (5e-5 if precision == "float32" else 1e-11)


@pytest.mark.parametrize("stepper", ["imex", "etd1"])
def test_coupled_rd_spectral_steps(stepper):
    """Semi-implicit and exponential steps with the symbol -max(D_A, D_B) k^2
    (problem_definition.py:593), periodic."""
    shape = (16, 20, 24)
    _, vg, og = setup(shape, "float64", (1.0, 1.0, 1.0))
    arr = species(shape, 7)
    p = problem(CoupledReactionDiffusion, vg, D_A=1.0, D_B=0.5)
    dt = 0.5
    st = (PseudoSpectralIMEX if stepper == "imex" else ExponentialEuler)(p, dt)
    got = st.step(0.0, torch.tensor(arr, dtype=torch.float64, device="cuda")).cpu().numpy()
    sym = -1.0 * O.rfft_k_squared(shape, og.spacing)
    rhs = lambda t, x: O.coupled_rd_rhs(x, og, BCS["ppp"])  # noqa: E731
    if stepper == "imex":
        want = O.imex_step(og.field(arr)[0], rhs, O.imex_prefac(dt, sym)).numpy()
    else:
        want = O.etd1_step(og.field(arr)[0], rhs, O.etd1_multiplier(dt, sym)).numpy()
    # both sides use the same float32 multiplier table (as the reference
    # rounds it); phi_1 evaluates exp on different devices
    assert rel_l2(got - arr, want - arr) < (1e-11 if stepper == "imex" else 1e-6)


# ------------------------------------------------------ smoothed boundary RD
def smooth_mask(shape, seed):
    rng = np.random.default_rng(seed)
    x, y, z = np.meshgrid(*[np.arange(n) / n for n in shape], indexing="ij")
    ph = rng.random(3) * 2 * np.pi
    return 0.5 + 0.3 * np.cos(2 * np.pi * x + ph[0]) * np.cos(4 * np.pi * y + ph[1]) + 0.15 * np.sin(2 * np.pi * z + ph[2])


@pytest.mark.parametrize("precision", ["float32", "float64"])
@pytest.mark.parametrize("case", ["no_mask", "mask_periodic", "mask_dirichlet", "mask_neumann_source"])
def test_sbm_rhs_matches_oracle(precision, case):
    shape = (10, 12, 66)
    _, vg, og = setup(shape, precision)
    rng = np.random.default_rng(11)
    arr = 0.2 + 0.6 * rng.random(shape)
    mask = None if case == "no_mask" else smooth_mask(shape, 12)
    bc = {"no_mask": BCS["npp"], "mask_periodic": BCS["ppp"], "mask_dirichlet": BCS["d_n_p"],
          "mask_neumann_source": BCS["nnn"]}[case]
    f = (lambda t, c, lib=None: c * (1 - c)) if case == "mask_neumann_source" else None
    p = problem(ReactionDiffusionSBM, vg, D=0.7, bc=bc, mask=mask, bc_flux=0.3, f=f)
    got = p.rhs(0.0, torch.tensor(arr, dtype=vg.precision, device="cuda")[None]).cpu().numpy()
    m, norm, bcm = O.rd_sbm_setup(og, bc, mask)
    want = O.rd_sbm_rhs(og.field(arr), og, bcm, 0.7, m, norm, 0.3,
                        f=(lambda c: c * (1 - c)) if f is not None else None).numpy()
    err = rel_l2(got, want)
    print(f"sbm {precision} {case}: rel l2 {err:.2e}", flush=True)
    assert err < TOL[precision] * (4 if precision == "float32" else 1)


def test_sbm_euler_steps():
    shape = (12, 10, 20)
    _, vg, og = setup(shape, "float64")
    rng = np.random.default_rng(13)
    arr = 0.3 + 0.4 * rng.random(shape)
    mask = smooth_mask(shape, 14)
    bc = BCS["d_n_p"]
    p = problem(ReactionDiffusionSBM, vg, D=1.0, bc=bc, mask=mask, bc_flux=0.1)
    st = ForwardEuler(p, 0.05)
    u = torch.tensor(arr, dtype=torch.float64, device="cuda")[None]
    for _ in range(4):
        u = st.step(0.0, u)
    m, norm, bcm = O.rd_sbm_setup(og, bc, mask)
    ur = og.field(arr)
    for _ in range(4):
        ur = O.fe_step(ur, lambda t, x: O.rd_sbm_rhs(x, og, bcm, 1.0, m, norm, 0.1), 0.05)
    assert rel_l2(u.cpu().numpy(), ur.numpy()) < 1e-12


# ------------------------------------------------- kernel path vs tensor path
@pytest.mark.parametrize("precision", ["float32", "float64"])
@pytest.mark.parametrize("fast", [True, False])
@pytest.mark.parametrize("curvature", [1.0, 0.3])
def test_multiphase_kernels_equal_tensor_path(precision, fast, curvature):
    """The projection + padded right-hand-side kernels against the tensor
    path of the same class (taken when the projection is replaced)."""
    shape = (10, 12, 66)
    _, vg, _ = setup(shape, precision)
    ph = torch.tensor(phases(shape, 4, 12), dtype=vg.precision, device="cuda")
    bc = BCS["npp"]
    k = problem(MultiPhaseAllenCahn, vg, eps=2.0, curvature=curvature, fast=fast, bc=bc)
    t = problem(MultiPhaseAllenCahn, vg, eps=2.0, curvature=curvature, fast=fast, bc=bc)
    builtin = t.project_to_simplex
    t.project_to_simplex = lambda x: builtin(x)  # replaced: the tensor path runs
    assert k._kernel_path(4) and not t._kernel_path(4)
    a, b = k.rhs(0.0, ph).cpu().numpy(), t.rhs(0.0, ph).cpu().numpy()
    assert rel_l2(a, b) < TOL[precision]


def test_multiphase_nine_phases_take_the_tensor_path():
    shape = (6, 8, 10)
    _, vg, og = setup(shape, "float64")
    arr = phases(shape, 9, 13)
    p = problem(MultiPhaseAllenCahn, vg, eps=3.0, bc=BCS["ppp"])
    assert not p._kernel_path(9)
    got = p.rhs(0.0, torch.tensor(arr, dtype=torch.float64, device="cuda")).cpu().numpy()
    want = O.multiphase_ac_rhs(og.field(arr)[0], og, BCS["ppp"], eps=3.0).numpy()
    assert rel_l2(got, want) < 1e-12


@pytest.mark.parametrize("precision", ["float32", "float64"])
def test_sbm_kernel_equals_tensor_path(precision):
    """A float bc_flux takes the kernel, a tensor bc_flux the tensor path."""
    shape = (10, 12, 66)
    _, vg, _ = setup(shape, precision)
    mask = smooth_mask(shape, 14)
    u = torch.tensor(0.2 + 0.6 * np.random.default_rng(15).random((1,) + shape), dtype=vg.precision, device="cuda")
    k = problem(ReactionDiffusionSBM, vg, D=0.7, bc=BCS["d_n_p"], mask=mask, bc_flux=0.3)
    t = problem(ReactionDiffusionSBM, vg, D=0.7, bc=BCS["d_n_p"], mask=mask,
                bc_flux=torch.tensor(0.3, dtype=vg.precision, device="cuda"))
    assert rel_l2(k.rhs(0.0, u).cpu(), t.rhs(0.0, u).cpu()) < TOL[precision]
