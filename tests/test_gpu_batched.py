"""Batched uneven grids: every grid of the packed batched step equals the
single-grid fused semi-implicit step bit for bit, under both schedules, any
tile shape, graph replay and the autotuner's choice; one grid is also checked
against the float64 oracle."""
import numpy as np
import pytest
import torch

import evoxels_amd as evx
from evoxels_amd.batched import BatchedCahnHilliard
from evoxels_amd.problem_definition import CahnHilliard
from evoxels_amd.timesteppers import PseudoSpectralIMEX
from evoxels_amd.voxelgrid import VoxelGridHip
from oracle import evoxels_oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [(16, 32, 48), (32, 32, 32), (48, 16, 64), (64, 48, 16), (96, 32, 16), (16, 128, 32), (128, 16, 96)]
DT = 0.05


def initial(shapes, seed=3):
    rng = np.random.default_rng(seed)
    return [rng.uniform(0.3, 0.7) + 0.05 * rng.standard_normal(s) for s in shapes]


def single_grid_steps(arr, eps, D, steps):
    shape = arr.shape
    vf = evx.VoxelFields(shape, shape)
    vg = VoxelGridHip(vf.grid_info(), precision="float32", device="cuda")
    st = PseudoSpectralIMEX(CahnHilliard(vg, eps=float(eps), D=float(D)), DT)
    u = vg.init_scalar_field(arr).contiguous()
    for _ in range(steps):
        st.step_into(u, u)
    return u[0]


@pytest.fixture(scope="module")
def reference():
    arrs = initial(SHAPES)
    eps = np.linspace(2.0, 3.5, len(SHAPES))
    D = np.linspace(0.6, 1.4, len(SHAPES))
    want = [single_grid_steps(a, e, d, 3) for a, e, d in zip(arrs, eps, D)]
    return arrs, eps, D, want


@pytest.mark.parametrize("mode,xchunk,blocks", [("persistent", 8, 0), ("persistent", 4, 300), ("staged", 16, 0),
                                                ("persistent", 32, 256)])
def test_batched_grids_equal_single_grid_steps(reference, mode, xchunk, blocks):
    arrs, eps, D, want = reference
    b = BatchedCahnHilliard(SHAPES, DT, eps=eps, D=D)
    b.set_schedule(mode, blocks=blocks, xchunk=xchunk)
    state = b.pack(arrs)
    b.step(state, 3)
    torch.cuda.synchronize()
    for i, (v, w) in enumerate(zip(b.views(state), want)):
        assert torch.equal(v, w), f"grid {i} {SHAPES[i]} differs ({mode}, xchunk {xchunk})"
    assert b.error_flag() == 0


def test_batched_graph_replay_and_autotune(reference):
    arrs, eps, D, want = reference
    b = BatchedCahnHilliard(SHAPES, DT, eps=eps, D=D)
    state = b.pack(arrs)
    sched = b.autotune(state, steps=2)
    assert sched["mode"] in ("persistent", "staged") and sched["ms_per_step"] > 0
    # the autotuner ran on a copy: state is untouched
    assert all(torch.equal(v, torch.tensor(a, dtype=torch.float32, device="cuda"))
               for v, a in zip(b.views(state), arrs))
    g = b.capture(state, 1)
    torch.cuda.synchronize()
    # capture only records; three replays with eager work in between
    for r in range(3):
        g.replay()
        torch.cuda.synchronize()
        _ = (state * 2).sum().item()
    for i, (v, w) in enumerate(zip(b.views(state), want)):
        assert torch.equal(v, w), f"grid {i} differs after graph replays"
    assert b.error_flag() == 0


def test_batched_grid_against_float64_oracle():
    shapes = [(32, 48, 16), (16, 16, 64)]
    arrs = initial(shapes, seed=8)
    b = BatchedCahnHilliard(shapes, 0.1, eps=3.0, D=1.0)
    state = b.pack(arrs)
    b.step(state)
    P3 = (("periodic", None),) * 3
    for v, a, s in zip(b.views(state), arrs, shapes):
        og = O.OracleGrid(s, (1.0, 1.0, 1.0), "float64")
        pre = O.imex_prefac(0.1, O.ch_symbol(O.rfft_k_squared(s, og.spacing), 3.0, 1.0))
        ref = O.imex_step(og.field(a), lambda t, x: O.ch_rhs(x, og, P3, 3.0, 1.0), pre)[0].numpy()
        got = v.cpu().numpy().astype(np.float64)
        a32 = a.astype(np.float32).astype(np.float64)
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-5
        assert np.linalg.norm((got - a32) - (ref - a)) / np.linalg.norm(ref - a) < 5e-4
