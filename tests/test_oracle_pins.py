"""Pins of the CPU oracle (oracle/evoxels_oracle.py) against the reference's own
known answers and against what the mathematics fixes.  CPU only."""
import json
import math
from pathlib import Path

import numpy as np
import pytest
import sympy as sp
import torch

from oracle import evoxels_oracle as O
from tests import mms

CS = mms.CS
P3 = (("periodic", None),) * 3
KAT = json.loads((Path(__file__).parent / "golden" / "reference_kats.json").read_text())


def grid(shape, domain=None, precision="float64"):
    domain = domain or tuple(float(n) for n in shape)
    return O.OracleGrid.from_domain(shape, domain, precision)


# ----------------------------------------------------------------- padding --
def test_pad_kat_reference_mixed_bc():
    """test_solvers.py:105-126, the reference's exact padded array."""
    with pytest.warns(UserWarning, match="order of convergence to 0.5"):
        bc = O.normalize_bc((("dirichlet", (10.0, 20.0)), "neumann", "periodic"))
    k = KAT["pad_mixed_bc"]
    f = torch.tensor(np.array(k["input"], dtype=np.float32).reshape([1] + k["input_shape"]))
    padded = O.pad_bc(f, bc)[0].numpy()
    assert np.array_equal(padded, np.array(k["expected_padded"], dtype=np.float32))


def test_bc_normalisation_and_errors():
    """problem_definition.py:60-103 and test_solvers.py:85-102."""
    with pytest.warns(UserWarning, match="reduces the spatial order of convergence to 0.5!"):
        bc = O.normalize_bc((("dirichlet", (1, -1)), "periodic", "periodic"))
    assert bc == (("dirichlet", (1, -1)), ("periodic", None), ("periodic", None))
    assert O.normalize_bc(None) == P3
    for bad, msg in [(("periodic",) * 2, "exactly three"), (("dirichlet", "periodic", "periodic"), "explicit values"),
                     (("robin", "periodic", "periodic"), "Unsupported BC type"),
                     ((("neumann", (1, 2)), "periodic", "periodic"), "do not accept"),
                     ((("dirichlet", (1,)), "periodic", "periodic"), "two boundary values")]:
        with pytest.raises(ValueError, match=msg):
            O.normalize_bc(bad)


def test_pad_ghost_values_by_hand():
    """Hand-computed ghost values, including edges/corners where rules compose."""
    u = torch.arange(1, 28, dtype=torch.float64).reshape(1, 3, 3, 3)  # u[i,j,k] = 1 + 9i + 3j + k
    bc = (("dirichlet", (1.0, 2.0)), ("neumann", None), ("dirichlet", (3.0, 4.0)))
    p = O.pad_bc(u, bc)[0]
    U = lambda i, j, k: 1 + 9 * i + 3 * j + k  # noqa: E731
    # face ghosts
    assert p[0, 2, 2] == 2 * 1.0 - U(0, 1, 1)
    assert p[4, 2, 2] == 2 * 2.0 - U(2, 1, 1)
    assert p[2, 0, 2] == U(1, 0, 1)
    assert p[2, 2, 0] == 2 * 3.0 - U(1, 1, 0)
    # x-y edge: x Dirichlet then y Neumann copy of the x ghost
    assert p[0, 0, 2] == 2 * 1.0 - U(0, 0, 1)
    # corner: z rule applied to the x-ruled, y-copied value
    assert p[0, 0, 0] == 2 * 3.0 - (2 * 1.0 - U(0, 0, 0))
    assert p[4, 4, 4] == 2 * 4.0 - (2 * 2.0 - U(2, 2, 2))
    # periodic y with Neumann x: edge = x copy of the y-wrapped value
    bc2 = (("neumann", None), ("periodic", None), ("periodic", None))
    p2 = O.pad_bc(u, bc2)[0]
    assert p2[0, 0, 2] == U(0, 2, 1)
    assert p2[4, 4, 0] == U(2, 0, 2)


# --------------------------------------------------------------- Laplacian --
@pytest.mark.parametrize("shape,spacing", [((8, 6, 10), (0.7, 1.3, 0.9)), ((5, 7, 9), (1.0, 1.0, 1.0))])
def test_laplace_periodic_fourier_eigenvalue(shape, spacing):
    """Discrete Fourier modes are exact eigenfunctions of the periodic 7-point Laplacian."""
    g = O.OracleGrid(shape, spacing, "float64")
    a = (2, 1, 3)
    idx = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    phase = sum(2 * np.pi * a[i] * idx[i] / shape[i] for i in range(3))
    u = torch.tensor(np.cos(phase))[None]
    lam = -sum(4 / spacing[i] ** 2 * np.sin(np.pi * a[i] / shape[i]) ** 2 for i in range(3))
    lap = O.laplace(O.pad_bc(u, P3), g.div_dx2)
    np.testing.assert_allclose(lap.numpy(), lam * u.numpy(), rtol=0, atol=1e-12 * abs(lam))


def _oracle_rhs(kind, bc, precision, **kw):
    def f(arr, n):
        g = O.OracleGrid.from_domain((n, n, n), (1, 1, 1), precision)
        u = g.field(arr)
        if kind == "rd":
            r = O.rd_rhs(u, g, bc, kw.get("D", 1.0))
        elif kind == "ch":
            r = O.ch_rhs(u, g, bc, **kw)
        else:
            r = O.ac_rhs(u, g, bc, **kw)
        return r[0].numpy()
    return f


def test_mms_periodic_laplace_slope():
    """test_laplace.py:21-33 (float32, periodic)."""
    e = sp.sin(2 * sp.pi * CS.x) * sp.sin(4 * sp.pi * CS.y) * sp.sin(6 * sp.pi * CS.z) / (56 * sp.pi ** 2)
    _, _, slope = mms.convergence(e, mms.rd_analytic(e), _oracle_rhs("rd", P3, "float32"), "float32")
    assert abs(slope - 2) < 0.1


def test_mms_neumann_laplace_matches_notebook_slope():
    """Zero-flux 7-point Laplacian, cell_center, float32: the reference's notebook
    reports slope 1.9533270430146406 (10-appendix-convergence-testing.ipynb, cell 9)."""
    e = sp.cos(sp.pi * CS.x) ** 3
    bc = (("neumann", None), ("periodic", None), ("periodic", None))
    _, _, slope = mms.convergence(e, mms.rd_analytic(e), _oracle_rhs("rd", bc, "float32"), "float32")
    assert abs(slope - KAT["neumann_laplace_slope"]["value"]) < 2e-3


def test_mms_dirichlet_laplace_matches_notebook_slope():
    """Zero Dirichlet, cell_center, float32: notebook slope 0.4954937374912848 (cell 11)."""
    e = (CS.x * (1 - CS.x)) ** 2
    bc = (("dirichlet", (0, 0)), ("periodic", None), ("periodic", None))
    _, _, slope = mms.convergence(e, mms.rd_analytic(e), _oracle_rhs("rd", bc, "float32"), "float32")
    assert abs(slope - KAT["dirichlet_laplace_slope"]["value"]) < 2e-3


def test_mms_cahn_hilliard_slope():
    """test_rhs.py:13-21 (float64)."""
    e = 0.4 + 0.1 * sp.sin(2 * sp.pi * CS.x)
    _, _, slope = mms.convergence(e, mms.ch_analytic(e, 3.0, 1.0), _oracle_rhs("ch", P3, "float64", eps=3.0, D=1.0))
    assert abs(slope - 2) < 0.1


def test_mms_allen_cahn_matches_notebook_errors():
    """test_rhs.py:29-37 slope, and the error array of
    10-appendix-convergence-testing.ipynb cell 16 (curvature 0.5, force 0)."""
    e = 0.5 + 0.3 * sp.cos(4 * sp.pi * CS.x) * sp.cos(2 * sp.pi * CS.y) * (CS.z ** 2 / 2 - CS.z ** 3 / 3)
    bc = (("neumann", None),) * 3
    _, errs, slope = mms.convergence(e, mms.ac_analytic(e, eps=3.0, curvature=0.5),
                                     _oracle_rhs("ac", bc, "float64", eps=3.0, curvature=0.5))
    np.testing.assert_allclose(errs, KAT["allen_cahn_errors"]["errors"], rtol=5e-6)
    assert abs(slope - KAT["allen_cahn_errors"]["slope"]) < 1e-5
    _, _, slope_f = mms.convergence(e, mms.ac_analytic(e, eps=3.0, curvature=0.5, force=1.0),
                                    _oracle_rhs("ac", bc, "float64", eps=3.0, curvature=0.5, force=1.0))
    assert abs(slope_f - 2) < 0.1


# ------------------------------------------------------------ CH invariants --
def _rand_field(shape, seed, lo=0.2, hi=0.8):
    rng = np.random.default_rng(seed)
    return torch.tensor(lo + (hi - lo) * rng.random(shape))[None]


@pytest.mark.parametrize("bc", [P3, (("neumann", None), ("periodic", None), ("neumann", None))])
def test_ch_rhs_conserves_mass(bc):
    g = grid((9, 8, 7))
    r = O.ch_rhs(_rand_field((9, 8, 7), 1), g, bc, eps=2.5, D=1.3)
    assert abs(float(r.sum())) < 1e-12 * float(r.abs().sum())


def test_ch_rhs_constant_field_and_translation():
    g = grid((8, 8, 8))
    assert float(O.ch_rhs(torch.full((1, 8, 8, 8), 0.37, dtype=torch.float64), g, P3).abs().max()) == 0.0
    u = _rand_field((8, 8, 8), 2)
    r = O.ch_rhs(u, g, P3)
    rr = O.ch_rhs(torch.roll(u, (2, -3, 1), (1, 2, 3)), g, P3)
    assert torch.equal(torch.roll(r, (2, -3, 1), (1, 2, 3)), rr)


def test_ch_rhs_single_mode_linearisation():
    """rhs(cbar + d*phi) = D*cbar(1-cbar)*lam*(f''(cbar) - 2*eps*lam)*d*phi + O(d^2)."""
    shape, sp_ = (16, 12, 10), (0.8, 1.1, 0.9)
    g = O.OracleGrid(shape, sp_, "float64")
    a = (1, 2, 1)
    idx = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    phi = np.cos(sum(2 * np.pi * a[i] * idx[i] / shape[i] for i in range(3)))
    lam = -sum(4 / sp_[i] ** 2 * np.sin(np.pi * a[i] / shape[i]) ** 2 for i in range(3))
    cbar, eps, D = 0.4, 2.0, 1.5
    fpp = 18 / eps * (1 - 6 * cbar + 6 * cbar ** 2)
    for d in (1e-4, 1e-5):
        r = O.ch_rhs(torch.tensor(cbar + d * phi)[None], g, P3, eps=eps, D=D)[0].numpy()
        lin = D * cbar * (1 - cbar) * lam * (fpp - 2 * eps * lam) * d * phi
        assert np.abs(r - lin).max() < 50 * d * np.abs(lin).max()


# ---------------------------------------------------------- normal Laplacian --
# A 1-D periodic profile of 0/1 values with two out-of-range entries, constant
# in the other two axes.  Worked by hand (dx = 1, eps = 3, D = 1):
#   clip:            c  = [0, 0, 1, 1, 0, 1, 0, 1]       (-0.1 -> 0, 1.1 -> 1)
#   c(1-c) = 0 everywhere, so mu = -2*eps*lap(c) = -6*(c[i+1] + c[i-1] - 2 c[i])
#                    mu = [-6, -6, 6, 6, -12, 12, -12, 12]
#   face i+1/2 mobility cbar(1-cbar), cbar = (c[i]+c[i+1])/2:
#                    m  = [0, 1/4, 0, 1/4, 1/4, 1/4, 1/4, 1/4]
#   flux m*(mu[i+1]-mu[i]) = [0, 3, 0, -4.5, 6, -6, 6, -4.5]
#   rhs = flux[i+1/2] - flux[i-1/2]
# A mobility averaged from the cells, avg(c(1-c)), is 0 on every face and
# gives rhs = 0; a clip after the mobility gives 0.2475 on the 1|-0.1 face.
CH_KAT_RAW = [0.0, 0.0, 1.0, 1.0, -0.1, 1.1, 0.0, 1.0]
CH_KAT_RHS = [4.5, 3.0, -3.0, -4.5, 10.5, -12.0, 12.0, -10.5]


def ch_kat_case(axis, other=3):
    shape = [other, other, other]
    shape[axis] = len(CH_KAT_RAW)
    prof = np.asarray(CH_KAT_RAW).reshape([-1 if a == axis else 1 for a in range(3)])
    want = np.asarray(CH_KAT_RHS).reshape([-1 if a == axis else 1 for a in range(3)])
    return tuple(shape), np.broadcast_to(prof, shape).copy(), np.broadcast_to(want, shape).copy()


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_ch_rhs_face_mobility_and_clip_by_hand(axis):
    """CahnHilliard.rhs (problem_definition.py:350-371): mobility
    cbar_f(1 - cbar_f) on faces of the clipped field, against the worked
    profile above, along each axis; D scales it, eps scales the gradient
    term, and a spacing of 2 divides it by 2^4."""
    shape, c, want = ch_kat_case(axis)
    for precision in ("float64", "float32"):
        g = grid(shape, precision=precision)
        got = O.ch_rhs(g.field(c), g, P3, eps=3.0, D=1.0)[0].double().numpy()
        np.testing.assert_array_equal(got, want)
        got = O.ch_rhs(g.field(c), g, P3, eps=3.0, D=2.0)[0].double().numpy()
        np.testing.assert_array_equal(got, 2 * want)
    g = grid(shape)
    np.testing.assert_allclose(O.ch_rhs(g.field(c), g, P3, eps=1.5)[0].numpy(), want / 2, rtol=0, atol=1e-13)
    dom = [float(n) for n in shape]
    dom[axis] *= 2
    g2 = grid(shape, tuple(dom))
    np.testing.assert_allclose(O.ch_rhs(g2.field(c), g2, P3)[0].numpy(), want / 16, rtol=0, atol=1e-14)


def test_normal_laplace_planar_profile_and_bulk():
    g = grid((12, 5, 6), (12.0, 5.0, 6.0))
    x = np.arange(12) + 0.5
    prof = np.tanh((x - 6.0) / 2.0)
    u = torch.tensor(np.broadcast_to(prof[:, None, None], (12, 5, 6)).copy())[None]
    bc = (("neumann", None), ("periodic", None), ("periodic", None))
    pu = O.pad_bc(u, bc)
    nl = O.normal_laplace(pu, g.div_dx, g.div_dx2)
    lx = (pu[:, 2:, 1:-1, 1:-1] - 2 * pu[:, 1:-1, 1:-1, 1:-1] + pu[:, :-2, 1:-1, 1:-1])
    grad = 0.5 * (pu[:, 2:, 1:-1, 1:-1] - pu[:, :-2, 1:-1, 1:-1])
    mask = (grad ** 2 > 1e-7)
    np.testing.assert_allclose(nl[mask].numpy(), lx[mask].numpy(), rtol=1e-12, atol=1e-14)
    const = O.pad_bc(torch.full((1, 4, 4, 4), 0.3, dtype=torch.float64), P3)
    assert float(O.normal_laplace(const, g.div_dx, g.div_dx2).abs().max()) == 0.0


def test_allen_cahn_tanh_equilibrium():
    """test_solvers.py:30-82 (Allen-Cahn leg): RK4, 10 steps of 0.5 from a step profile."""
    nx, eps = 16, 3.0
    g = grid((nx, 1, 1), (nx, 1, 1))
    phi = np.zeros((nx, 1, 1), dtype=np.float32)
    phi[: nx // 2] = 1.0
    u = torch.tensor(phi)[None]
    g32 = O.OracleGrid.from_domain((nx, 1, 1), (nx, 1, 1), "float32")
    bc = (("neumann", None),) * 3
    rhs = lambda t, v: O.ac_rhs(v, g32, bc, eps=eps)  # noqa: E731
    for i in range(10):
        u = O.rk4_step(u, rhs, 0.5, t=i * 0.5)
    x = np.arange(nx) + 0.5
    ana = 0.5 - 0.5 * np.tanh(3 * (x - 0.5 * nx) / 2 / eps)
    assert np.linalg.norm(u[0, :, 0, 0].numpy() - ana) < 0.05


# ------------------------------------------------------- k tables, steppers --
def test_k_table_float32_quirk():
    for n, d in [(7, 0.3), (100, 1.0), (512, 1.0), (256, 1 / 256)]:
        f = torch.fft.fftfreq(n, d)
        assert f.dtype == torch.float32
        j = np.arange(n)
        js = np.where(j < (n + 1) // 2, j, j - n).astype(np.float32)
        assert np.array_equal(f.numpy(), js * np.float32(1.0 / (n * d)))
    k2 = O.rfft_k_squared((6, 5, 8), (1.0, 0.5, 2.0))
    assert k2.dtype == torch.float32 and k2.shape == (6, 5, 5)
    x = torch.tensor([3.0, 0.7, 11.0], dtype=torch.float32)
    assert torch.equal(O.imex_prefac(0.1, x), torch.reciprocal(1 - 0.1 * x) * 0.1)


def _mode(shape, a):
    idx = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    return np.cos(sum(2 * np.pi * a[i] * idx[i] / shape[i] for i in range(3)))


@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_rd_imex_single_mode_closed_form(precision):
    """RD + IMEX on one Fourier mode: u1 = u0*(1 + prefac(k)*D*lam_h(k)) with
    prefac from float32 k^2 (computed here independently with numpy)."""
    shape, spc, D, dt = (8, 6, 10), (0.5, 1.0, 2.0), 0.7, 0.3
    a = (1, 2, 3)
    g = O.OracleGrid(shape, spc, precision)
    u0 = torch.tensor(_mode(shape, a), dtype=g.dtype)[None]
    lam = -sum(4 / spc[i] ** 2 * np.sin(np.pi * a[i] / shape[i]) ** 2 for i in range(3))
    k = [np.float32(2 * np.pi) * (np.float32(a[i]) * np.float32(1.0 / (shape[i] * spc[i]))) for i in range(3)]
    k2 = (k[0] * k[0] + k[1] * k[1]) + k[2] * k[2]
    sym = np.float32(-D * 0.25) * k2
    prefac = (np.float32(1) / (np.float32(1) - np.float32(dt) * sym)) * np.float32(dt)
    k2t = O.rfft_k_squared(shape, spc)
    pre = O.imex_prefac(dt, O.rd_symbol(k2t, D))
    u1 = O.imex_step(u0, lambda t, v: O.rd_rhs(v, g, P3, D), pre)
    expect = u0.numpy() * (1 + float(prefac) * D * lam)
    tol = 1e-12 if precision == "float64" else 2e-6
    np.testing.assert_allclose(u1.numpy(), expect, atol=tol)


def _dft3(x, inverse=False):
    """Brute-force multidimensional DFT (O(N^2) per axis)."""
    out = x.astype(complex)
    for ax in range(x.ndim):
        n = x.shape[ax]
        s = 1 if inverse else -1
        W = np.exp(s * 2j * np.pi * np.outer(np.arange(n), np.arange(n)) / n)
        out = np.moveaxis(np.tensordot(W, np.moveaxis(out, ax, 0), axes=(1, 0)), 0, ax)
    return out / (np.prod(x.shape) if inverse else 1)


def test_imex_step_matches_brute_force_dft():
    shape = (3, 4, 5)
    g = O.OracleGrid(shape, (1.0, 1.0, 1.0), "float64")
    u = _rand_field(shape, 4)
    k2 = O.rfft_k_squared(shape, g.spacing)
    pre = O.imex_prefac(0.2, O.ch_symbol(k2, 3.0, 1.0))
    rhs = lambda t, v: O.ch_rhs(v, g, P3, 3.0, 1.0)  # noqa: E731
    u1 = O.imex_step(u, rhs, pre)[0].numpy()
    # full-spectrum prefac by Hermitian symmetry of the float32 table
    kk = [2 * np.pi * np.fft.fftfreq(n, 1.0) for n in shape]
    KX, KY, KZ = np.meshgrid(*kk, indexing="ij")
    full = np.empty(shape)
    nh = shape[2] // 2 + 1
    for i in range(shape[0]):
        for j in range(shape[1]):
            for l in range(shape[2]):
                if l < nh:
                    full[i, j, l] = pre[i, j, l]
                else:
                    full[i, j, l] = pre[(-i) % shape[0], (-j) % shape[1], shape[2] - l]
    r = rhs(0, u)[0].numpy()
    upd = _dft3(full * _dft3(r), inverse=True).real
    np.testing.assert_allclose(u1, u[0].numpy() + upd, atol=1e-13)


def test_fe_and_rk4_closed_forms_on_a_mode():
    shape, D, dt = (8, 8, 8), 0.9, 0.05
    g = grid(shape)
    a = (1, 0, 2)
    u0 = torch.tensor(_mode(shape, a))[None]
    lam = -sum(4 * np.sin(np.pi * a[i] / shape[i]) ** 2 for i in range(3))
    z = dt * D * lam
    rhs = lambda t, v: O.rd_rhs(v, g, P3, D)  # noqa: E731
    np.testing.assert_allclose(O.fe_step(u0, rhs, dt).numpy(), (1 + z) * u0.numpy(), atol=1e-13)
    np.testing.assert_allclose(O.rk4_step(u0, rhs, dt).numpy(),
                               (1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24) * u0.numpy(), atol=1e-13)


def test_run_loop_semantics():
    """test_solvers.py:14-27 (u+1 with two fields) and the frame schedule of solvers.py:192-208."""
    u = torch.stack([torch.ones(4, 4, 4), torch.zeros(4, 4, 4)])
    seen = []
    out = O.run_loop(u, lambda t, v: v + 1, 0.1, 1, 1, lambda v, f, t: seen.append((f, round(t, 12))))
    assert torch.all(out[0] == 2) and torch.all(out[1] == 1)
    assert seen == [(0, 0.0), (1, 0.1)]
    seen.clear()
    O.run_loop(u, lambda t, v: v, 0.5, 3, 10, lambda v, f, t: seen.append((f, t)))
    assert [f for f, _ in seen] == [0, 1, 2, 3, 4] and seen[-1][1] == 5.0
    with pytest.raises(ZeroDivisionError):
        O.run_loop(u, lambda t, v: v, 0.5, 20, 10)


def test_ch_energy_gradient_term_closed_form():
    shape, L = (16, 4, 4), 16.0
    g = grid(shape, (L, 4.0, 4.0))
    x = np.arange(16) + 0.5
    delta, eps = 0.05, 2.0
    c = 0.5 + delta * np.cos(2 * np.pi * x / L)
    u = torch.tensor(np.broadcast_to(c[:, None, None], shape).copy())[None]
    e = float(O.ch_energy(u, g, P3, eps)[0])
    grad_part = eps * delta ** 2 * np.sin(2 * np.pi / L) ** 2 * (16 * 16 / 2)
    bulk = 9 / eps * np.sum(np.broadcast_to((c ** 2 * (1 - c) ** 2)[:, None, None], shape))
    assert abs(e - (grad_part + bulk)) < 1e-12 * e


def test_phi1_pade_branch_and_formula_branch():
    """phi_1(z) = expm1(z)/z: the (6,6) Pade branch (|z| < 0.5) agrees with the
    exact value to float64 roundoff (a dropped or mis-signed coefficient
    fails), the formula branch is the direct expression, phi_1(0) = 1."""
    z = torch.tensor([0.0, 1e-8, -1e-3, 0.1, -0.3, 0.49, -0.4999, 0.5, -0.75, -3.0, -40.0, 2.0],
                     dtype=torch.float64)
    got = O.phi1(z.clone()).numpy()
    want = np.array([1.0 if v == 0 else math.expm1(v) / v for v in z.tolist()])
    assert np.abs(got - want).max() <= 2e-15 * np.abs(want).max()
    small = np.abs(z.numpy()) < 0.5
    # Pade (6,6) error is O(z^13): ~1e-17 at |z| = 0.5, far below float64 eps
    assert np.abs(got[small] - want[small]).max() < 1e-15


@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_etd1_single_mode_closed_form(precision):
    """ReactionDiffusion + ExponentialEuler on one Fourier mode:
    u1 = u0 * (1 + m(k) * D * lambda_h(k)), m = dt*phi_1(dt*sigma_k) from the
    float32 symbol sigma = -D*A*k^2."""
    n, D, dt = 16, 0.8, 2.5
    g = grid((n, n, n), precision=precision)
    a = (1, 2, 3)
    x = np.arange(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    u0 = np.cos(2 * np.pi * (a[0] * X + a[1] * Y + a[2] * Z) / n)
    k2 = O.rfft_k_squared((n, n, n), g.spacing)
    mult = O.etd1_multiplier(dt, O.rd_symbol(k2, D))
    u1 = O.etd1_step(g.field(u0), lambda t, v: O.rd_rhs(v, g, P3, D), mult)[0].numpy()
    lam = -sum(4 * np.sin(np.pi * ai / n) ** 2 for ai in a)
    kk = [np.float32(2 * np.pi) * (np.float32(ai) * np.float32(1.0 / n)) for ai in a]
    kq = (kk[0] * kk[0] + kk[1] * kk[1]) + kk[2] * kk[2]
    zf = np.float32(dt) * (np.float32(-D * 0.25) * kq)
    m = float(np.float32(dt) * np.float32(math.expm1(float(zf)) / float(zf)))
    expect = u0 * (1 + m * D * lam)
    # m is a float32 value: one ulp of it (float32 exp vs expm1) is ~1e-6 here
    assert np.abs(u1 - expect).max() < 5e-6 * np.abs(expect).max()


# ------------------------------------------- multiphase AC / coupled RD / SBM --
MPAC_FUNS = (
    sp.Rational(1, 3) + sp.sin(2 * sp.pi * CS.z) / 20 + sp.cos(2 * sp.pi * CS.x) / 10,
    sp.Rational(1, 3) + sp.sin(2 * sp.pi * CS.y) / 10 - sp.cos(2 * sp.pi * CS.x) / 10,
    sp.Rational(1, 3) - sp.sin(2 * sp.pi * CS.z) / 20 - sp.sin(2 * sp.pi * CS.y) / 10,
)
N3 = (("neumann", None), ("periodic", None), ("periodic", None))


def _oracle_channels(fn):
    def f(arr, n, mask):
        g = O.OracleGrid.from_domain((n, n, n), (1, 1, 1), "float64")
        return fn(torch.tensor(arr), g, mask).numpy()
    return f


@pytest.mark.parametrize("fast", [True, False])
def test_mms_multiphase_allen_cahn_slopes(fast):
    """test_rhs.py:55-79: eps 3, gab 1, M 1, curvature 1, Neumann x, float64;
    both simplex projections (the test phases already lie on the simplex)."""
    _, _, slopes = mms.convergence_channels(
        MPAC_FUNS, mms.multiphase_ac_analytic(MPAC_FUNS),
        _oracle_channels(lambda u, g, m: O.multiphase_ac_rhs(u, g, N3, fast=fast)))
    assert all(abs(s - 2) < 0.1 for s in slopes), slopes


@pytest.mark.parametrize("curvature", [1.0, 0.4])
def test_multiphase_two_phases_reduce_to_two_phase_allen_cahn(curvature):
    """With phases (phi, 1 - phi) the multiphase well potential difference is
    4 phi (1-phi)(1-2phi) and the normal Laplacian is odd in phi, so channel 0
    equals the pinned two-phase Allen-Cahn RHS (force 0) for any curvature."""
    g = grid((9, 12, 7), (4.0, 5.0, 3.0))
    phi = _rand_field((9, 12, 7), 21, 0.05, 0.95)
    phis = torch.cat([phi, 1 - phi])
    bc = (("neumann", None), ("periodic", None), ("neumann", None))
    for eps in (2.0, 3.5):
        got = O.multiphase_ac_rhs(phis, g, bc, eps=eps, gab=1.3, M=0.7, curvature=curvature)
        want = O.ac_rhs(phi, g, bc, eps=eps, gab=1.3, M=0.7, force=0.0, curvature=curvature)
        np.testing.assert_allclose(got[0:1].numpy(), want.numpy(), rtol=0, atol=1e-12)
        np.testing.assert_allclose(got[1:2].numpy(), -want.numpy(), rtol=0, atol=1e-12)


def _simplex_by_bisection(v):
    """Euclidean projection of one point onto the simplex by bisection on the
    threshold theta of sum(max(v - theta, 0)) = 1."""
    lo, hi = v.min() - 1.0, v.max()
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if np.maximum(v - mid, 0).sum() > 1:
            lo = mid
        else:
            hi = mid
    return np.maximum(v - 0.5 * (lo + hi), 0)


def test_euclidean_simplex_projection_against_bisection():
    """The sort/cumsum projection inside multiphase_ac_rhs (problem_definition.py:
    514-530) equals an independent bisection; recovered here from the rhs of a
    zero-gradient state, where the RHS is a function of the projected point."""
    rng = np.random.default_rng(5)
    pts = rng.normal(0.3, 0.6, size=(4, 40))
    g = grid((1, 1, 1), (1.0, 1.0, 1.0))
    for v in pts.T:
        want = _simplex_by_bisection(v)
        u = torch.tensor(v).reshape(4, 1, 1, 1)
        # constant field: the gradient term vanishes, the well term of the
        # projected point remains, and it is reproduced from ``want``
        got = O.multiphase_ac_rhs(u, g, P3, eps=3.0, fast=False)
        w = torch.tensor(want).reshape(4, 1, 1, 1)
        pf = 9 / (2 * 9.0)
        df = pf * (3 * w * (torch.sum(w ** 2, 0, keepdim=True) - w ** 2) + w ** 3 - w)
        np.testing.assert_allclose(got.numpy(), -(df - df.mean(0, keepdim=True)).numpy(), rtol=0, atol=1e-12)
        assert abs(want.sum() - 1) < 1e-12 and want.min() >= 0


def test_sloppy_projection_and_phase_sum():
    """Clip-and-normalise projection: the projected phases are nonnegative
    and sum to one; the RHS of every voxel sums to zero over the phases."""
    g = grid((8, 6, 10))
    rng = np.random.default_rng(9)
    u = torch.tensor(rng.uniform(-0.1, 1.0, (3, 8, 6, 10)))
    u[0] = u[0].abs() + 0.01  # a voxel whose phases all clip to zero is 0/0, as in the reference
    r = O.multiphase_ac_rhs(u, g, N3, fast=True)
    assert float(r.sum(0).abs().max()) < 1e-12
    # a point already on the simplex is left alone by both projections
    phis = torch.tensor(rng.dirichlet([2, 2, 2], size=(8, 6, 10)).transpose(3, 0, 1, 2).copy())
    np.testing.assert_allclose(O.multiphase_ac_rhs(phis, g, N3, fast=True).numpy(),
                               O.multiphase_ac_rhs(phis, g, N3, fast=False).numpy(), rtol=0, atol=1e-12)


RD_FUNS = (0.5 + 0.3 * sp.cos(4 * sp.pi * CS.x) * sp.cos(4 * sp.pi * CS.y) ** 3,
           0.4 + 0.1 * sp.sin(2 * sp.pi * CS.x) * sp.cos(4 * sp.pi * CS.z))


def test_mms_coupled_reaction_diffusion_slopes():
    """test_rhs.py:40-52: D_A 1, D_B 0.5, default feed/kill, periodic, float64."""
    _, _, slopes = mms.convergence_channels(
        RD_FUNS, mms.coupled_rd_analytic(RD_FUNS, 1.0, 0.5),
        _oracle_channels(lambda u, g, m: O.coupled_rd_rhs(u, g, P3, 1.0, 0.5)))
    assert all(abs(s - 2) < 0.1 for s in slopes), slopes


def test_coupled_rd_uniform_state_is_the_kinetics_ode():
    """A uniform state has zero Laplacian: the RHS is the Gray-Scott kinetics
    -a b^2 + F (1 - a), a b^2 - k b at every voxel."""
    g = grid((5, 4, 6))
    a, b, F, k = 0.3, 0.6, 0.04, 0.09
    u = torch.stack([torch.full((5, 4, 6), a, dtype=torch.float64), torch.full((5, 4, 6), b, dtype=torch.float64)])
    r = O.coupled_rd_rhs(u, g, N3, D_A=0.8, D_B=0.3, feed=F, kill=k)
    assert float((r[0] - (-0.108 + 0.028)).abs().max()) < 1e-15
    assert float((r[1] - (0.108 - 0.054)).abs().max()) < 1e-15


def _sbm_oracle(bc, D=1.0, bc_flux=0.0):
    def fn(u, g, mask):
        m, norm, bcm = O.rd_sbm_setup(g, bc, mask)
        return O.rd_sbm_rhs(u, g, bcm, D, m, norm, bc_flux)
    return fn


def test_mms_smoothed_boundary_periodic_slope():
    """Cell-centred form of test_rhs.py:82-97 (the reference runs it on a
    staggered_x grid): periodic test function times the reference's mask,
    bc_flux 1, float64, second order."""
    mask_fun = 0.5 + 0.3 * sp.cos(4 * sp.pi * CS.x) * sp.cos(2 * sp.pi * CS.y)
    u = (0.6 + 0.2 * sp.sin(2 * sp.pi * CS.x) * sp.cos(2 * sp.pi * CS.z)) * mask_fun
    _, _, slopes = mms.convergence_channels(
        (u,), (mms.sbm_analytic(u, mask_fun, 1.0, 1.0),),
        _oracle_channels(_sbm_oracle(P3, 1.0, 1.0)), mask=mask_fun)
    assert abs(slopes[0] - 2) < 0.1, slopes


@pytest.mark.parametrize("bc", [P3, N3, (("dirichlet", (0.2, 0.7)), ("periodic", None), ("neumann", None))])
def test_smoothed_boundary_reductions(bc):
    """Without a mask the SBM RHS is the plain diffusion RHS plus bc_flux; with
    a constant mask the mask gradients vanish and it is the diffusion RHS."""
    g = grid((7, 9, 6), (3.0, 4.0, 2.5))
    u = _rand_field((7, 9, 6), 31)
    m, norm, bcm = O.rd_sbm_setup(g, bc, None)
    np.testing.assert_allclose(O.rd_sbm_rhs(u, g, bcm, 0.7, m, norm, 0.25).numpy(),
                               (O.rd_rhs(u, g, bc, 0.7) + 0.25).numpy(), rtol=0, atol=1e-12)
    if bc[0][0] != "dirichlet":
        m, norm, bcm = O.rd_sbm_setup(g, bc, np.full((7, 9, 6), 0.6))
        assert float(norm.abs().max()) == 0.0
        np.testing.assert_allclose(O.rd_sbm_rhs(u, g, bcm, 0.7, m, norm, 0.25).numpy(),
                                   O.rd_rhs(u, g, bc, 0.7).numpy(), rtol=0, atol=1e-12)


# ---------------------------------------------------- staggered_x and spectral BCs --
def _staggered_rd(bc, precision="float32", D=1.0):
    def f(arr, n):
        g = O.OracleGrid((n + 1, n, n), (1.0 / n, 1.0 / n, 1.0 / n), precision, "staggered_x")
        u = O.trim_boundary_nodes(g.field(arr), "staggered_x")
        return O.rd_rhs(u, g, bc, D)[0].numpy()
    return f


def test_mms_staggered_neumann_laplace_matches_notebook_slope():
    """Zero-flux 4/3, -1/3 ghost on staggered_x, float32: notebook slope
    2.0915744090903576 (10-appendix-convergence-testing.ipynb cell 9)."""
    e = sp.cos(sp.pi * CS.x) ** 3
    bc = (("neumann", None), ("periodic", None), ("periodic", None))
    _, _, slope = mms.convergence(e, mms.rd_analytic(e), _staggered_rd(bc), "float32", convention="staggered_x")
    assert abs(slope - KAT["staggered_neumann_laplace_slope"]["value"]) < 2e-3


def test_mms_staggered_dirichlet_laplace_matches_notebook_slope():
    """Boundary-node Dirichlet on staggered_x, float32: notebook slope
    2.0314069655338005 (cell 11)."""
    e = (CS.x * (1 - CS.x)) ** 2
    bc = (("dirichlet", (0, 0)), ("periodic", None), ("periodic", None))
    _, _, slope = mms.convergence(e, mms.rd_analytic(e), _staggered_rd(bc), "float32", convention="staggered_x")
    assert abs(slope - KAT["staggered_dirichlet_laplace_slope"]["value"]) < 2e-3


def test_mms_staggered_nonzero_dirichlet_slope():
    """test_laplace.py:63-76: cos(pi x)^3 with node values (1, -1), order 2."""
    e = sp.cos(sp.pi * CS.x) ** 3
    bc = (("dirichlet", (1.0, -1.0)), ("periodic", None), ("periodic", None))
    _, _, slope = mms.convergence(e, mms.rd_analytic(e), _staggered_rd(bc), "float32", convention="staggered_x")
    assert abs(slope - 2) < 0.1


def test_mms_staggered_smoothed_boundary_slope():
    """test_rhs.py:82-97: smoothed-boundary diffusion on staggered_x, node
    values (0, 1) scaled by the mask, bc_flux 1, float32, order 2."""
    mask_fun = 0.5 + 0.3 * sp.cos(4 * sp.pi * CS.x) * sp.cos(2 * sp.pi * CS.y)
    u = (CS.x + CS.x * (1 - CS.x)) * mask_fun
    bc = (("dirichlet", (0.0, 1.0)), ("periodic", None), ("periodic", None))

    def fn(arr, n, mask):
        g = O.OracleGrid((n + 1, n, n), (1.0 / n, 1.0 / n, 1.0 / n), "float32", "staggered_x")
        m, norm, bcm = O.rd_sbm_setup(g, bc, mask)
        v = O.trim_boundary_nodes(torch.tensor(arr, dtype=torch.float32), "staggered_x")
        return O.rd_sbm_rhs(v, g, bcm, 1.0, m, norm, 1.0).numpy()
    _, _, slopes = mms.convergence_channels((u,), (mms.sbm_analytic(u, mask_fun, 1.0, 1.0),), fn, "float32",
                                            mask=mask_fun, convention="staggered_x")
    assert abs(slopes[0] - 2) < 0.1, slopes


def test_staggered_pad_by_hand():
    """Ghost planes of the two staggered_x paddings on a 3x2x2 interior,
    written out element by element."""
    u = torch.arange(1.0, 13.0, dtype=torch.float64).reshape(1, 3, 2, 2)
    p = O.pad_staggered_x(u, "dirichlet", (5.0, -2.0))
    assert torch.all(p[:, 0] == 5.0) and torch.all(p[:, -1] == -2.0)
    assert torch.equal(p[:, 1:-1], O.pad_periodic(u)[:, 1:-1])
    q = O.pad_staggered_x(u, "neumann")
    per = O.pad_periodic(u)
    for j in range(4):
        for k in range(4):
            assert float(q[0, 0, j, k]) == 4 / 3 * float(per[0, 1, j, k]) - 1 / 3 * float(per[0, 2, j, k])
            assert float(q[0, -1, j, k]) == 4 / 3 * float(per[0, -2, j, k]) - 1 / 3 * float(per[0, -3, j, k])
    # a quadratic with zero slope at the boundary node is reproduced exactly
    xs = torch.arange(1, 4, dtype=torch.float64)  # interior nodes 1..3, boundary node 0
    quad = (0.7 + 0.2 * xs ** 2).reshape(1, 3, 1, 1).expand(1, 3, 2, 2).contiguous()
    assert abs(float(O.pad_staggered_x(quad, "neumann")[0, 0, 1, 1]) - 0.7) < 1e-14


@pytest.mark.parametrize("case", ["cc_neumann", "cc_dirichlet", "stag_dirichlet"])
@pytest.mark.parametrize("m", [1, 3])
def test_nonperiodic_imex_step_on_eigenmodes(case, m):
    """The mirror extension turns the boundary eigenmodes of the
    finite-difference Laplacian into single Fourier modes, so one
    semi-implicit RD step is u1 = u0 (1 + prefac(k) D lambda_h) with
    k_x = pi m / L and lambda_h the discrete eigenvalue."""
    n, ny, nz, D, dt = 12, 8, 6, 0.8, 0.3
    conv = "cell_center" if case.startswith("cc") else "staggered_x"
    nx_nodes = n if conv == "cell_center" else n + 2
    L = (3.0, 2.0, 1.5)
    sp_, _ = O.spacing_for((nx_nodes, ny, nz), L, conv)
    g = O.OracleGrid((nx_nodes, ny, nz), sp_, "float64", conv)
    i = np.arange(n)
    if case == "cc_neumann":
        mode_x, lam_x = np.cos(np.pi * m * (i + 0.5) / n), np.sin(np.pi * m / (2 * n)) ** 2
    elif case == "cc_dirichlet":
        mode_x, lam_x = np.sin(np.pi * m * (i + 0.5) / n), np.sin(np.pi * m / (2 * n)) ** 2
    else:  # interior nodes 1..N-1 of N = n + 1 intervals
        N = n + 1
        mode_x, lam_x = np.sin(np.pi * m * (i + 1) / N), np.sin(np.pi * m / (2 * N)) ** 2
    j = np.arange(ny)
    u0 = (mode_x[:, None, None] * np.cos(2 * np.pi * j / ny)[None, :, None] * np.ones(nz)[None, None, :])
    lam = -4 / sp_[0] ** 2 * lam_x - 4 / sp_[1] ** 2 * np.sin(np.pi / ny) ** 2
    kx = np.pi * m / (n * sp_[0] if conv == "cell_center" else (n + 1) * sp_[0])
    k2 = kx ** 2 + (2 * np.pi / (ny * sp_[1])) ** 2
    prefac = dt / (1 + dt * D * 0.25 * k2)
    kind = "neumann" if case == "cc_neumann" else "dirichlet"
    bc = ((kind, None if kind == "neumann" else (0.0, 0.0)), ("periodic", None), ("periodic", None))
    mult = O.imex_prefac(dt, O.rd_symbol(O.rfft_k_squared_nonperiodic(g.shape, g.spacing, conv), D))
    u1 = O.spectral_step_bc(g.field(u0), lambda t, x: O.rd_rhs(x, g, bc, D), mult, kind, conv)
    want = u0 * (1 + prefac * D * lam)
    # the multiplier table is float32, as the reference's
    np.testing.assert_allclose(u1[0].numpy(), want, rtol=0, atol=2e-7 * np.abs(u0).max())
