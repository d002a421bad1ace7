"""Host-side logic of the 'hip' backend that needs no GPU: the VoxelFields
mirror (tests/test_fields.py of the reference), boundary-condition
validation, FFT-BC checks and backend selection."""
import types
import warnings

import numpy as np
import pytest
import torch

import evoxels_amd as evx
from evoxels_amd.problem_definition import CahnHilliard, ReactionDiffusion, TwoPhaseAllenCahn
from evoxels_amd.solvers import TimeDependentSolver
from evoxels_amd.timesteppers import PseudoSpectralIMEX
from evoxels_amd.voxelgrid import CellCenteredBCs, VoxelGridHip


def test_voxelfields_geometry():
    """test_fields.py:8-33."""
    vf = evx.VoxelFields((10, 10, 10))
    assert vf.shape == (10, 10, 10) and (vf.Nx, vf.Ny, vf.Nz) == (10, 10, 10)
    vf = evx.VoxelFields((10, 10, 10), (10, 10, 10))
    assert (vf.domain_size, vf.spacing) == ((10, 10, 10), (1, 1, 1))
    vf = evx.VoxelFields((10, 5, 7), (10, 5, 7))
    x, y, z = vf.meshgrid()
    assert x[-1, 0, 0] == 10 - vf.spacing[0] / 2
    vf = evx.VoxelFields((11, 5, 7), (10, 5, 7), convention="staggered_x")
    x, y, z = vf.meshgrid()
    assert x[-1, 0, 0] == 10 and y[0, -1, 0] == 5 - vf.spacing[1] / 2


def test_voxelfields_fields_and_errors():
    vf = evx.VoxelFields((10, 10, 10), (10, 10, 10))
    vf.add_field("c", 0.123 * np.ones((10, 10, 10)))
    assert (vf.fields["c"][1, 2, 3], *vf.fields["c"].shape) == (0.123, 10, 10, 10)
    with pytest.raises(ValueError):
        vf.set_field("c", np.ones((3, 3, 3)))
    with pytest.raises(TypeError):
        vf.set_field("c", [1, 2])
    with pytest.raises(ValueError):
        evx.VoxelFields((10, 10))
    with pytest.raises(ValueError):
        evx.VoxelFields((4, 4, 4), convention="bad")
    with pytest.warns(UserWarning, match="questionable"):
        evx.VoxelFields((4, 4, 4), (1, 1, 100))


def test_sphere_volume_fraction():
    """test_fields.py:60-65."""
    vf = evx.VoxelFields((6, 5, 5), convention="staggered_x")
    vf.add_field("sphere")
    vf.set_voxel_sphere("sphere", center=(0.5, 0.5, 0.5), radius=0.31, label=1)
    assert np.count_nonzero(vf.fields["sphere"] == 1) == 20
    assert vf.average("sphere") == 0.16


def _stub_vg(convention="cell_center"):
    vg = types.SimpleNamespace(convention=convention, shape=(4, 4, 4))
    vg.bc = CellCenteredBCs(vg)
    return vg


def test_bc_normalisation_matches_reference():
    """test_solvers.py:85-102 and problem_definition.py:60-103."""
    with pytest.warns(UserWarning, match="reduces the spatial order of convergence to 0.5!"):
        p = ReactionDiffusion(_stub_vg(), D=1.0, bc=(("dirichlet", (1, -1)), "periodic", "periodic"))
    assert p.bc == (("dirichlet", (1, -1)), ("periodic", None), ("periodic", None))
    assert p.bc_type == ("dirichlet", "periodic", "periodic")
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        p = CahnHilliard(_stub_vg())
    assert p.bc_type == ("periodic",) * 3 and p.fused
    for bad, msg in [(("periodic",) * 2, "exactly three"), (("dirichlet", "periodic", "periodic"), "explicit values"),
                     (("robin", "periodic", "periodic"), "Unsupported BC type"),
                     ((("neumann", (1, 2)), "periodic", "periodic"), "do not accept")]:
        with pytest.raises(ValueError, match=msg):
            ReactionDiffusion(_stub_vg(), D=1.0, bc=bad)


def test_exponential_euler_rejects_full_neumann_semilinear_problem():
    """test_solvers.py:165-171."""
    from evoxels_amd.timesteppers import ExponentialEuler
    with pytest.raises(ValueError, match="support at most one non-periodic axis"):
        ExponentialEuler(TwoPhaseAllenCahn(_stub_vg()), 0.1)


def test_fft_bc_validation():
    """test_solvers.py:165-171 and problem_definition.py:138-167."""
    with pytest.raises(ValueError, match="support at most one non-periodic axis"):
        PseudoSpectralIMEX(TwoPhaseAllenCahn(_stub_vg()), 0.1)
    p = ReactionDiffusion(_stub_vg(), D=1.0, bc=("periodic", "neumann", "periodic"))
    with pytest.raises(NotImplementedError, match="only implement the single non-periodic axis"):
        p.verify_fft_bc_config()


def test_custom_closures_disable_fusion():
    assert not CahnHilliard(_stub_vg(), mu_hom=lambda c: c).fused
    assert not ReactionDiffusion(_stub_vg(), D=1.0, f=lambda t, c, lib: c).fused
    assert TwoPhaseAllenCahn(_stub_vg()).fused


def test_backend_selection():
    vf = evx.VoxelFields((4, 4, 4))
    vf.add_field("c")
    with pytest.raises(ValueError, match="Unsupported backend: torch"):
        TimeDependentSolver(vf, "c", backend="torch")


def test_hip_grid_refuses_cpu_device():
    vf = evx.VoxelFields((4, 4, 4))
    with pytest.raises(RuntimeError, match="ROCm"):
        VoxelGridHip(vf.grid_info(), device="cpu")
    if not torch.cuda.is_available():
        with pytest.raises(RuntimeError, match="ROCm GPU"):
            VoxelGridHip(vf.grid_info(), device="cuda")


def test_inversion_checkpoint_schedule_and_argument_checks():
    from evoxels_amd.inversion import _obs_map, _params_tensor, checkpoint_schedule
    assert checkpoint_schedule(12, 5) == [(10, 12), (5, 10), (0, 5)]
    assert checkpoint_schedule(40, 20) == [(20, 40), (0, 20)]
    assert checkpoint_schedule(0, 20) == []
    with pytest.raises(ValueError):
        checkpoint_schedule(10, 0)
    assert _obs_map([4, 8, 12], 12) == {4: 0, 8: 1, 12: 2}
    with pytest.raises(ValueError):
        _obs_map([4, 13], 12)
    with pytest.raises(ValueError):
        _obs_map([4, 4], 12)
    p = _params_tensor({"eps": [3.0, 2.0], "D": 1.5}, 2, "cpu")
    assert p.dtype == torch.float64 and p.tolist() == [[3.0, 1.5], [2.0, 1.5]]
    with pytest.raises(ValueError):
        _params_tensor([[1.0, 2.0, 3.0]], 1, "cpu")


def test_batched_plan_rejects_unsupported_shapes_and_dtype_before_any_gpu_work():
    import ctypes as C
    from evoxels_amd import _lib
    lib = _lib.load()
    h = C.c_void_p()
    shp = np.array([[32, 32, 100]], dtype=np.int64)
    sp = np.ones((1, 3))
    assert lib.evx_batch_plan_create(C.byref(h), _lib.EVX_F32, 1, shp.ctypes.data, sp.ctypes.data) \
        == _lib.EVX_ERR_UNSUPPORTED
    assert "edges" in _lib.last_error()
    shp = np.array([[32, 32, 32]], dtype=np.int64)
    assert lib.evx_batch_plan_create(C.byref(h), _lib.EVX_F64, 1, shp.ctypes.data, sp.ctypes.data) \
        == _lib.EVX_ERR_UNSUPPORTED
    assert lib.evx_batch_plan_create(C.byref(h), _lib.EVX_F32, 0, shp.ctypes.data, sp.ctypes.data) \
        == _lib.EVX_ERR_INVALID


def test_gc_paused_restores_the_collector_state():
    """Graph captures run with automatic garbage collection paused; the
    previous state comes back afterwards, also after an exception."""
    import gc
    from evoxels_amd import _lib
    assert gc.isenabled()
    with _lib.gc_paused():
        assert not gc.isenabled()
    assert gc.isenabled()
    try:
        with _lib.gc_paused():
            raise KeyError("inside")
    except KeyError:
        pass
    assert gc.isenabled()
    gc.disable()
    try:
        with _lib.gc_paused():
            assert not gc.isenabled()
        assert not gc.isenabled()
    finally:
        gc.enable()


def test_batch_graph_refuses_replay_after_schedule_change():
    """A batched graph records the tile schedule it was captured with; after
    set_tiles/autotune bump the plan's generation, replay raises instead of
    launching the old schedule (no GPU: stand-ins for the plan and graph)."""
    from evoxels_amd.batched import BatchGraph
    runner = types.SimpleNamespace(_generation=3)
    calls = []
    g = BatchGraph(runner, types.SimpleNamespace(replay=lambda: calls.append(1)), runner._generation)
    g.replay()
    assert calls == [1] and not g.stale
    runner._generation += 1
    assert g.stale
    with pytest.raises(RuntimeError, match="capture again"):
        g.replay()
    assert calls == [1]


def test_replaced_rhs_disables_fused_steps():
    """The manufactured-solution driver assigns ``problem.rhs`` on the
    instance (utils.py:285); the fused kernels evaluate the built-in
    right-hand side, so such a problem must not take them (no GPU: a
    stand-in grid)."""
    vg = _stub_vg()
    for cls, kw in ((ReactionDiffusion, {"D": 1.0}), (CahnHilliard, {}), (TwoPhaseAllenCahn, {})):
        p = cls(vg, **kw)
        assert p.fused and p._kernel_ok
        st = PseudoSpectralIMEX(p, 0.1) if cls is not TwoPhaseAllenCahn else None
        if st is not None:
            assert st._kind in ("rd", "ch") and st.capturable
        p.rhs = lambda t, u: u
        assert not p.fused and p._kernel_ok
        if st is not None:  # the stepper was built before the replacement
            assert st._kind is None and not st.capturable

    class MyCH(CahnHilliard):
        def rhs(self, t, c):
            return super().rhs(t, c)

    assert not MyCH(vg).fused


def test_inversion_sequence_checks_and_step_counts():
    """InversionModel.train's sequence validation (inversion.py:170-192;
    test_inversion.py:12-22) and the fixed-step counts of output times."""
    from evoxels_amd.inversion import SaveAt, steps_of, validate_inds
    with pytest.raises(ValueError, match="different length"):
        validate_inds([[0, 1, 2], [0, 1]])
    with pytest.raises(ValueError, match="different spacing"):
        validate_inds([[0, 1, 2], [0, 2, 4]])
    with pytest.raises(ValueError, match="at least 2"):
        validate_inds([[3]])
    validate_inds([[1, 2, 3], [4, 5, 6]])
    assert steps_of(np.array([0.0, 0.1, 0.2, 0.30000000000000004]), 0.1) == [0, 1, 2, 3]
    assert steps_of(SaveAt.times(SaveAt(ts=[1.0, 2.0, 4.0])), 0.5) == [0, 2, 6]
    with pytest.raises(ValueError):
        steps_of(np.array([0.0, 0.15]), 0.1)
    s = SaveAt(ts=np.linspace(0, 1, 5))
    assert s.subs.ts is s.ts and len(SaveAt.times(s)) == 5
