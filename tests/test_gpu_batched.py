"""Batched uneven grids: every grid of the packed batched step equals the
single-grid fused semi-implicit step (to float32 rounding) for several tile
shapes, under graph replay and with the autotuner's choice; two grids are
also checked against the float64 oracle."""
import time

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


def compare(views, want, what):
    """Largest difference to the single-grid steps, relative to the field."""
    worst = 0.0
    for i, (v, w) in enumerate(zip(views, want)):
        d = float((v - w).abs().max() / w.abs().max())
        print(f"{what}: grid {i} {SHAPES[i]} max rel diff {d:.3e} bitwise {torch.equal(v, w)}", flush=True)
        worst = max(worst, d)
    # same tile bodies, compiled into different kernels: the compiler may
    # contract multiply-adds differently, so agreement is to float32 rounding
    assert worst < 2e-6, f"{what}: batched grids differ from single-grid steps by {worst:.3e}"


PERSISTENT_SKIP = ("a persistent single-launch schedule (workgroups pulling items from a device queue with "
                   "per-grid sweep counters) stalled in its dependency waits on the GPU twice, with no cause shown; "
                   "it is not part of this build")


@pytest.mark.parametrize("mode,xchunk", [("staged", 4), ("staged", 16), ("staged", 64),
                                         pytest.param("persistent", 8, marks=pytest.mark.skip(reason=PERSISTENT_SKIP)),
                                         pytest.param("persistent", 32, marks=pytest.mark.skip(reason=PERSISTENT_SKIP))])
def test_batched_grids_equal_single_grid_steps(reference, mode, xchunk):
    arrs, eps, D, want = reference
    b = BatchedCahnHilliard(SHAPES, DT, eps=eps, D=D)
    b.set_tiles(xchunk)
    state = b.pack(arrs)
    t0 = time.perf_counter()
    b.step(state, 3)
    torch.cuda.synchronize()
    print(f"{mode} xchunk {xchunk}: 3 steps in {time.perf_counter() - t0:.3f} s", flush=True)
    compare(b.views(state), want, f"{mode}/{xchunk}")


def test_batched_fft_group_is_bitwise_neutral(reference):
    """FFT items covering 3 (partial last groups) or 8 consecutive tiles give
    the same bits as one tile per item: a workgroup runs the same tile bodies
    back to back."""
    arrs, eps, D, want = reference
    b = BatchedCahnHilliard(SHAPES, DT, eps=eps, D=D)
    outs = {}
    for g in (1, 3, 8):
        b.set_tiles(8, fft_group=g)
        sched = b.schedule
        assert sched["fft_group"] == g
        state = b.pack(arrs)
        b.step(state, 3)
        torch.cuda.synchronize()
        print(f"fft_group {g}: {sched['items']} items", flush=True)
        outs[g] = [v.clone() for v in b.views(state)]
    assert all(torch.equal(a, c) for a, c in zip(outs[1], outs[3]))
    assert all(torch.equal(a, c) for a, c in zip(outs[1], outs[8]))
    compare(outs[3], want, "fft_group 3")


@pytest.mark.parametrize("fft_group", [1, 3])
def test_batched_one_launch_equals_six_launches(reference, fft_group):
    """The step as one cooperative launch (blocks walking each sweep's items,
    grid barriers between sweeps) gives the same bits as one launch per
    sweep: the same tile bodies on the same items in the same sweep order."""
    arrs, eps, D, want = reference
    b = BatchedCahnHilliard(SHAPES, DT, eps=eps, D=D)
    outs = {}
    for one in (False, True):
        b.set_tiles(8, fft_group=fft_group, one_launch=one)
        assert b.schedule["one_launch"] is one and b.schedule["launches_per_step"] == (1 if one else 6)
        state = b.pack(arrs)
        t0 = time.perf_counter()
        b.step(state, 3)
        torch.cuda.synchronize()
        print(f"one_launch {one} fft_group {fft_group}: 3 steps in {time.perf_counter() - t0:.3f} s", flush=True)
        outs[one] = [v.clone() for v in b.views(state)]
    assert all(torch.equal(x, y) for x, y in zip(outs[False], outs[True]))
    compare(outs[True], want, f"one launch, fft_group {fft_group}")


def test_batched_one_launch_graph_replay(reference):
    """A captured graph of one-launch steps, replayed three times with eager
    work in between, each replay checked."""
    arrs, eps, D, want = reference
    b = BatchedCahnHilliard(SHAPES, DT, eps=eps, D=D)
    b.set_tiles(8, one_launch=True)
    state = b.pack(arrs)
    g = b.capture(state, 1)
    torch.cuda.synchronize()
    assert all(torch.equal(v, torch.tensor(a, dtype=torch.float32, device="cuda"))
               for v, a in zip(b.views(state), arrs))
    ref = BatchedCahnHilliard(SHAPES, DT, eps=eps, D=D)
    rstate = ref.pack(arrs)
    for r in range(3):
        g.replay()
        ref.step(rstate, 1)
        torch.cuda.synchronize()
        _ = (state * 2).sum().item()
        assert all(torch.equal(x, y) for x, y in zip(b.views(state), ref.views(rstate))), f"replay {r}"
    compare(b.views(state), want, "one-launch graph")


def test_batched_graph_replay(reference):
    arrs, eps, D, want = reference
    b = BatchedCahnHilliard(SHAPES, DT, eps=eps, D=D)
    state = b.pack(arrs)
    g = b.capture(state, 1)
    torch.cuda.synchronize()
    # capture only records: the state is untouched
    assert all(torch.equal(v, torch.tensor(a, dtype=torch.float32, device="cuda"))
               for v, a in zip(b.views(state), arrs))
    for r in range(3):  # three replays with eager work in between
        t0 = time.perf_counter()
        g.replay()
        torch.cuda.synchronize()
        print(f"replay {r}: {time.perf_counter() - t0:.3f} s", flush=True)
        _ = (state * 2).sum().item()
    compare(b.views(state), want, "graph")


def test_batched_autotune(reference):
    arrs, eps, D, want = reference
    b = BatchedCahnHilliard(SHAPES, DT, eps=eps, D=D)
    state = b.pack(arrs)
    t0 = time.perf_counter()
    sched = b.autotune(state, steps=1)
    print(f"autotune {time.perf_counter() - t0:.3f} s -> {sched}", flush=True)
    assert sched["xchunk"] in (4, 6, 8, 12, 16, 24, 32, 48, 64) and sched["fft_group"] in (1, 2, 4, 8)
    assert sched["launches_per_step"] == (1 if sched["one_launch"] else 6)
    assert sched["ms_per_step"] > 0
    # the autotuner ran on a copy: state is untouched
    assert all(torch.equal(v, torch.tensor(a, dtype=torch.float32, device="cuda"))
               for v, a in zip(b.views(state), arrs))
    b.step(state, 3)
    torch.cuda.synchronize()
    compare(b.views(state), want, "autotuned")


def test_batched_autotune_starting_from_one_launch(reference):
    """The tile shapes are timed with one launch per sweep whatever the
    schedule before the call; the last pass picks the launch schedule."""
    arrs, eps, D, want = reference
    b = BatchedCahnHilliard(SHAPES, DT, eps=eps, D=D)
    b.set_tiles(8, fft_group=1, one_launch=True)
    state = b.pack(arrs)
    sched = b.autotune(state, steps=1)
    print(f"autotune from one launch -> {sched}", flush=True)
    assert sched["launches_per_step"] == (1 if sched["one_launch"] else 6)
    b.step(state, 3)
    torch.cuda.synchronize()
    compare(b.views(state), want, "autotuned from one launch")


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


def config4_draw():
    """The 48-grid workload: edges from {32, 48, 64, 96, 128} per axis and
    c0 = U[0.3, 0.7] + 0.05*N(0,1) per grid, numpy default_rng(3), in this
    order; dt 0.1, eps 3, D 1, unit spacing."""
    rng = np.random.default_rng(3)
    shapes = [tuple(int(v) for v in s) for s in rng.choice([32, 48, 64, 96, 128], (48, 3))]
    arrs = [rng.uniform(0.3, 0.7) + 0.05 * rng.standard_normal(s) for s in shapes]
    return shapes, arrs


def test_config4_48_uneven_grids_three_steps():
    """The 48-grid draw for three autotuned, graph-replayed steps: every grid
    against the single-grid fused step, and three grids (one with a 128
    edge) against three steps of the float64 oracle (field relative L2
    <= 1e-5, update <= 5e-4 per step)."""
    shapes, arrs = config4_draw()
    assert len(shapes) == 48 and any(128 in s for s in shapes)
    b = BatchedCahnHilliard(shapes, 0.1, eps=3.0, D=1.0)
    state = b.pack(arrs)
    sched = b.autotune(state, steps=2)
    print(f"config 4: {len(shapes)} grids, schedule {sched}", flush=True)
    g = b.capture(state, 3)
    g.replay()
    torch.cuda.synchronize()
    worst = 0.0
    for i, (v, a, s) in enumerate(zip(b.views(state), arrs, shapes)):
        vf = evx.VoxelFields(s, s)
        vg = VoxelGridHip(vf.grid_info(), precision="float32", device="cuda")
        st = PseudoSpectralIMEX(CahnHilliard(vg, eps=3.0, D=1.0), 0.1)
        u = vg.init_scalar_field(a).contiguous()
        for _ in range(3):
            st.step_into(u, u)
        d = float((v - u[0]).abs().max() / u[0].abs().max())
        worst = max(worst, d)
    print(f"config 4: largest batched vs single-grid difference {worst:.3e}", flush=True)
    assert worst < 2e-6
    P3 = (("periodic", None),) * 3
    pick = [next(i for i, s in enumerate(shapes) if 128 in s)]
    pick += [i for i in (0, 23, 47) if i not in pick][:2]
    assert len(set(pick)) == 3
    views = b.views(state)
    for i in pick:
        s, a = shapes[i], arrs[i]
        og = O.OracleGrid(s, (1.0, 1.0, 1.0), "float64")
        pre = O.imex_prefac(0.1, O.ch_symbol(O.rfft_k_squared(s, og.spacing), 3.0, 1.0))
        ref = og.field(a.astype(np.float32).astype(np.float64))
        bb = BatchedCahnHilliard([s], 0.1, eps=3.0, D=1.0)
        st1 = bb.pack([a])
        for k in range(3):
            prev = ref
            ref = O.imex_step(ref, lambda t, x: O.ch_rhs(x, og, P3, 3.0, 1.0), pre)
            prev_got = bb.views(st1)[0].cpu().double().numpy()
            bb.step(st1)
            got = bb.views(st1)[0].cpu().double().numpy()
            r = ref[0].numpy()
            ferr = np.linalg.norm(got - r) / np.linalg.norm(r)
            uerr = np.linalg.norm((got - prev_got) - (r - prev[0].numpy())) / np.linalg.norm(r - prev[0].numpy())
            print(f"config 4 grid {i} {s} step {k + 1}: field {ferr:.3e} update {uerr:.3e}", flush=True)
            assert ferr <= 1e-5 and uerr <= 5e-4
        np.testing.assert_allclose(got, views[i].cpu().double().numpy(), rtol=0, atol=2e-6)
