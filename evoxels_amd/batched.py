"""Many independent periodic Cahn-Hilliard grids of different shapes, advanced
together on one GPU.

Each grid follows ``CahnHilliard`` + ``PseudoSpectralIMEX``
(problem_definition.py:290-371, timesteppers.py:64-89) with the tile bodies
of the single-grid fused step.  All grids live in one packed float32 state;
one call advances every grid by one step with six launches (one per sweep),
each covering the tiles of every grid.  ``autotune`` times the tile shapes
and keeps the fastest; ``capture`` records a number of steps in a HIP graph,
so a replay advances all grids by many steps with one host call.

The reference has no batched solver: it advances one grid per solver
(solvers.py:80-114).
"""
from __future__ import annotations

import ctypes as C
import weakref
from typing import Sequence

import numpy as np
import torch

from . import _lib


class BatchedCahnHilliard:
    """Semi-implicit CH steps for grids of shapes ``shapes`` (edges in
    {16, 32, 48, 64, 96, 128}); ``eps``/``D`` are scalars or per-grid
    sequences; ``spacing`` is None (unit spacing) or one 3-tuple per grid."""

    def __init__(self, shapes: Sequence[Sequence[int]], dt: float, eps=3.0, D=1.0, A: float = 0.25,
                 spacing=None, device="cuda"):
        self.shapes = [tuple(int(v) for v in s) for s in shapes]
        n = len(self.shapes)
        if n == 0:
            raise ValueError("at least one grid is required")
        if any(len(s) != 3 for s in self.shapes):
            raise ValueError("every shape must be (nx, ny, nz)")
        dev = torch.device(device)
        if dev.type != "cuda" or not torch.cuda.is_available():
            raise RuntimeError("the 'hip' backend needs a ROCm GPU")
        if dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        sp = np.ones((n, 3)) if spacing is None else np.asarray(spacing, dtype=np.float64).reshape(n, 3)
        shp = np.ascontiguousarray(np.asarray(self.shapes, dtype=np.int64))
        sp = np.ascontiguousarray(sp)
        lib = _lib.load()
        h = C.c_void_p()
        _lib.check(lib.evx_batch_plan_create(C.byref(h), _lib.EVX_F32, n, shp.ctypes.data, sp.ctypes.data),
                   "evx_batch_plan_create")
        self.handle = h
        self._fin = weakref.finalize(self, lib.evx_batch_plan_destroy, h)
        offs = np.zeros(n, dtype=np.int64)
        total = C.c_int64()
        _lib.check(lib.evx_batch_state_layout(h, offs.ctypes.data, C.byref(total)), "evx_batch_state_layout")
        self.offsets = [int(v) for v in offs]
        self.total = int(total.value)
        nb = C.c_size_t()
        _lib.check(lib.evx_batch_workspace_bytes(h, C.byref(nb)), "evx_batch_workspace_bytes")
        self.ws = torch.empty(nb.value, dtype=torch.uint8, device=dev)
        self.dt = float(dt)
        self.A = float(A)
        self.set_params(eps, D)

    def set_params(self, eps, D):
        n = len(self.shapes)
        e = np.broadcast_to(np.asarray(eps, dtype=np.float64), (n,))
        d = np.broadcast_to(np.asarray(D, dtype=np.float64), (n,))
        ed = np.ascontiguousarray(np.stack([e, d], 1))
        self.eps, self.D = e.copy(), d.copy()
        _lib.check(_lib.load().evx_batch_set_params(self.handle, ed.ctypes.data, self.dt, self.A),
                   "evx_batch_set_params")

    # ------------------------------------------------------------ state
    def new_state(self):
        return torch.zeros(self.total, dtype=torch.float32, device=self.device)

    def views(self, state):
        """Per-grid (nx, ny, nz) views into the packed state."""
        return [state[o:o + int(np.prod(s))].view(s) for o, s in zip(self.offsets, self.shapes)]

    def pack(self, arrays):
        state = self.new_state()
        for v, a in zip(self.views(state), arrays):
            v.copy_(torch.as_tensor(np.asarray(a), dtype=torch.float32))
        return state

    # ------------------------------------------------------------ steps
    def step(self, state, n: int = 1):
        if state.dtype != torch.float32 or state.numel() < self.total or not state.is_contiguous() \
                or state.device != self.device:
            raise ValueError("state must be a contiguous float32 tensor from new_state()/pack()")
        lib = _lib.load()
        s = _lib.stream_handle(state.device)
        for _ in range(n):
            _lib.check(lib.evx_batch_ch_imex_step(self.handle, _lib.ptr(state), _lib.ptr(self.ws), s),
                       "evx_batch_ch_imex_step")
        return state

    def capture(self, state, n_steps: int):
        """A graph whose replay advances ``state`` by ``n_steps`` steps.  It
        holds the tile schedule of the moment: after ``set_tiles`` or
        ``autotune`` change it, ``replay`` raises instead of running the old
        launches against the new item list."""
        side = torch.cuda.Stream(device=self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        g = torch.cuda.CUDAGraph()
        with _lib.gc_paused(), torch.cuda.graph(g, stream=side):
            self.step(state, n_steps)
        return BatchGraph(self, g, self._generation)

    # ------------------------------------------------------------ tiles
    @property
    def schedule(self):
        x, ni, g, one = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        lib = _lib.load()
        _lib.check(lib.evx_batch_get_tiles(self.handle, C.byref(x), C.byref(ni)), "evx_batch_get_tiles")
        _lib.check(lib.evx_batch_get_fft_group(self.handle, C.byref(g)), "evx_batch_get_fft_group")
        _lib.check(lib.evx_batch_get_one_launch(self.handle, C.byref(one)), "evx_batch_get_one_launch")
        return {"launches_per_step": 1 if one.value else 6, "one_launch": bool(one.value), "xchunk": x.value,
                "fft_group": g.value, "items": ni.value}

    _generation = 0

    def set_tiles(self, xchunk=8, fft_group=None, one_launch=None):
        """Schedule: x planes per CH tile; (if given) the number of
        consecutive FFT tiles each workgroup of the FFT sweeps covers; (if
        given) one cooperative launch per step instead of one per sweep."""
        torch.cuda.synchronize(self.device)
        now = self.schedule
        if int(xchunk) != now["xchunk"] or (fft_group is not None and int(fft_group) != now["fft_group"]) \
                or (one_launch is not None and bool(one_launch) != now["one_launch"]):
            self._generation += 1
        lib = _lib.load()
        _lib.check(lib.evx_batch_set_tiles(self.handle, int(xchunk)), "evx_batch_set_tiles")
        if fft_group is not None:
            _lib.check(lib.evx_batch_set_fft_group(self.handle, int(fft_group)), "evx_batch_set_fft_group")
        if one_launch is not None:
            _lib.check(lib.evx_batch_set_one_launch(self.handle, int(bool(one_launch))), "evx_batch_set_one_launch")

    def autotune(self, state, steps: int = 5):
        """Times the candidate tile shapes on a copy of ``state`` and keeps the
        fastest; returns the choice with its time per step."""
        scratch = state.clone()
        self._generation += 1
        best = C.c_double()
        _lib.check(_lib.load().evx_batch_autotune(self.handle, _lib.ptr(scratch), _lib.ptr(self.ws), int(steps),
                                                  _lib.stream_handle(state.device), C.byref(best)),
                   "evx_batch_autotune")
        out = self.schedule
        out["ms_per_step"] = best.value
        return out


class BatchGraph:
    """A captured batch of steps (``BatchedCahnHilliard.capture``).  ``replay``
    checks that the plan's tile schedule is still the one recorded."""

    def __init__(self, runner: BatchedCahnHilliard, graph, generation: int):
        self.runner = runner
        self.graph = graph
        self.generation = generation

    @property
    def stale(self) -> bool:
        return self.generation != self.runner._generation

    def replay(self):
        if self.stale:
            raise RuntimeError("the tile schedule changed (set_tiles/autotune) after this graph was captured; "
                               "capture again")
        self.graph.replay()
