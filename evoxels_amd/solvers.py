"""Solver driver for the 'hip' backend (mirror of evoxels.solvers).

``TimeDependentSolver(vf, fieldnames, backend='hip', problem_cls,
timestepper_cls, step_fn, device)`` and ``solve(...)`` keep the signatures
and loop semantics of evoxels/solvers.py:12-208: channel stacking, output
before the step whenever ``i % (max_iters // frames) == 0`` plus a final
output, per-frame export into ``vf.fields``, NaN abort with ``sys.exit(1)``.

``jit=True`` means hipGraph capture here (the reference uses torch.compile):
for a stepper whose launches do not depend on ``t`` the steps between two
outputs are captured once and replayed.  The per-frame NaN check is a flag
fused into the compensated float64 mass reduction (one small read-back
instead of a full-field ``isnan().any()``).

Frames are exported asynchronously by default (``async_output``): the frame
is copied on the device, then to pinned host memory on a side stream while
the next steps run, and handed to ``vf.fields`` (with the NaN abort and the
VTK file, written by a background thread) when the following frame is
produced or the run ends.  ``vf.fields``, the files and the abort message are
those of the synchronous loop; only their timing differs.
"""
from __future__ import annotations

import sys
import warnings
from abc import ABC, abstractmethod
from collections import deque
from dataclasses import dataclass
from timeit import default_timer as timer
from typing import Any, Callable, Type

import numpy as np
import torch

from . import _lib
from .problem_definition import ODE
from .timesteppers import TimeStepper
from .vtk_io import AsyncWriter


class _Slot:
    """Device snapshot, pinned host copy and read-back of one output frame."""

    def __init__(self, shape, dtype, channels, device):
        self.key = (tuple(shape), dtype)
        self.snap = torch.empty(shape, dtype=dtype, device=device)
        self.host = torch.empty(shape, dtype=dtype, pin_memory=True)
        self.flags = torch.zeros(2, dtype=torch.int32, device=device)
        self.flags_host = torch.empty(2, dtype=torch.int32, pin_memory=True)
        self.mass_host = torch.empty(channels, dtype=torch.float64, pin_memory=True)
        self.event = torch.cuda.Event()
        self.mass = None


class _FrameExporter:
    """Double-buffered asynchronous frame output.  ``submit`` runs on the
    compute stream only the device-side work (mass + NaN flags, snapshot
    copy) and queues the device-to-host copies on a side stream of default
    priority; ``complete`` waits for the oldest frame and performs the host
    side of solvers.py:152-187 for it."""

    DEPTH = 2

    def __init__(self, solver, device):
        self.solver = solver
        self.device = device
        self.stream = torch.cuda.Stream(device=device)
        self.free: list[_Slot] = []
        self.pending: deque = deque()
        self.writer = None

    def submit(self, u, u_out, frame, time, vtk_out):
        while len(self.pending) >= self.DEPTH:
            self.complete()
        key = (tuple(u_out.shape), u_out.dtype)
        slot = next((s for s in self.free if s.key == key), None)
        if slot is None:
            slot = _Slot(u_out.shape, u_out.dtype, u.shape[0], self.device)
        else:
            self.free.remove(slot)
        cur = torch.cuda.current_stream(self.device)
        slot.flags.zero_()
        slot.mass = self.solver._frame_sums(u, slot.flags)
        slot.snap.copy_(u_out)
        self.stream.wait_stream(cur)
        with torch.cuda.stream(self.stream):
            slot.host.copy_(slot.snap, non_blocking=True)
            slot.mass_host.copy_(slot.mass, non_blocking=True)
            slot.flags_host.copy_(slot.flags, non_blocking=True)
            slot.event.record(self.stream)
        self.pending.append((slot, frame, time, vtk_out))

    def complete(self):
        slot, frame, time, vtk_out = self.pending.popleft()
        slot.event.synchronize()
        self.solver._export_fields(slot.host.clone())
        self.solver.mass_history.append((frame, time, slot.mass_host.numpy().copy()))
        slot.mass = None
        self.free.append(slot)
        if int(slot.flags_host[0]):
            self.drain_files()
            print(f"NaN detected in frame {frame} at time {time}. Aborting simulation.")
            sys.exit(1)
        if vtk_out:
            vf = self.solver.vf
            if not hasattr(vf, "export_to_vtk"):
                raise NotImplementedError("vtk_out needs a VoxelFields with export_to_vtk")
            if self.writer is None:
                self.writer = AsyncWriter()
            names = self.solver.fieldnames
            self.writer.submit(self.solver._vtk_name(frame), {n: vf.fields[n] for n in names},
                               vf.spacing, vf.origin)

    def drain_files(self):
        if self.writer is not None:
            self.writer.close()
            self.writer = None

    def flush(self):
        while self.pending:
            self.complete()
        self.drain_files()


@dataclass
class BaseSolver(ABC):
    """solvers.py:12-187."""
    vf: Any
    fieldnames: str | list[str]
    backend: str
    problem_cls: Type[ODE] | None = None
    timestepper_cls: Type[TimeStepper] | None = None
    step_fn: Callable | None = None
    device: str = "cuda"
    async_output: bool = True

    def __post_init__(self):
        if self.backend == "hip":
            from .voxelgrid import VoxelGridHip
            from .profiler import HipMemoryProfiler
            self.vg = VoxelGridHip(self.vf.grid_info(), precision=self.vf.precision, device=self.device)
            self.profiler = HipMemoryProfiler(self.vg.device)
        else:
            raise ValueError(f"Unsupported backend: {self.backend}")
        if isinstance(self.fieldnames, str):
            self.fieldnames = [self.fieldnames]
        else:
            self.fieldnames = list(self.fieldnames)
        self.timestepper = None
        self.mass_history = []

    def _init_fields(self):
        u_list = [self.vg.init_scalar_field(self.vf.fields[name]) for name in self.fieldnames]
        u = self.vg.concatenate(u_list, 0)
        return self.vg.bc.trim_boundary_nodes(u).contiguous()

    def _init_stepper(self, time_increment, problem_kwargs, jit):
        problem_kwargs = problem_kwargs or {}
        if self.step_fn is not None:
            self.problem = None
            step = self.step_fn
        else:
            if self.problem_cls is None or self.timestepper_cls is None:
                raise ValueError("Either provide step_fn or both problem_cls and timestepper_cls")
            self.problem = self.problem_cls(self.vg, **problem_kwargs)
            self.timestepper = self.timestepper_cls(self.problem, time_increment)
            step = self.timestepper.step
        self._capture = bool(jit) and self.timestepper is not None and \
            getattr(self.timestepper, "capturable", False)
        # A captured graph holds the kernel arguments (dt, parameters) and the
        # workspace of the stepper it was recorded from: a new stepper needs a
        # new capture.
        self._drop_graph()
        return step

    def _drop_graph(self):
        self._graph = self._graph_in = self._graph_out = self._graph_step = None
        self._graph_key = None

    @abstractmethod
    def _run_loop(self, u, step, time_increment, frames, max_iters, vtk_out, verbose, plot_bounds, colormap):
        raise NotImplementedError

    def solve(self, time_increment=0.1, frames=10, max_iters=100, problem_kwargs=None, jit=True,
              verbose=True, vtk_out=False, plot_bounds=None, colormap="viridis"):
        """solvers.py:80-114."""
        u = self._init_fields()
        step = self._init_stepper(time_increment, problem_kwargs, jit)
        self._u_from_step = False
        self._exporter = _FrameExporter(self, self.vg.device) if self.async_output else None
        torch.cuda.synchronize(self.vg.device)
        start = timer()
        try:
            u = self._run_loop(u, step, time_increment, frames, max_iters, vtk_out, verbose, plot_bounds,
                               colormap)
            if self._exporter is not None:
                self._exporter.flush()
        finally:
            if self._exporter is not None:
                self._exporter.drain_files()
        torch.cuda.synchronize(self.vg.device)
        end = timer()
        self.computation_time = end - start
        if verbose:
            self.profiler.print_memory_stats(start, end, max_iters)

    def _frame_sums(self, u, flags):
        """Mass (compensated float64 per channel) and NaN/Inf flags of the
        frame state u: from the records the last step's final sweep left when
        u is that step's output, otherwise by a reduction over u."""
        ts = self.timestepper
        if getattr(self, "_u_from_step", False) and hasattr(ts, "last_step_sums"):
            sums = ts.last_step_sums(u, flags)
            if sums is not None:
                return sums
        return self.vg.channel_sums(u, flags)

    def _export_fields(self, u_out):
        host = u_out.cpu().numpy()
        for i, name in enumerate(self.fieldnames):
            self.vf.set_field(name, host[i])

    def _vtk_name(self, frame):
        """solvers.py:167-169."""
        return (self.problem_cls.__name__ if self.problem_cls else "step") + "_" + self.fieldnames[0] + \
            f"_{frame:03d}.vtk"

    def _output_field(self, u):
        """trim_ghost_nodes(pad_bc(u)) (solvers.py:154-157): u itself on
        cell_center grids; on staggered_x grids the boundary nodes come back
        from the ghost rule."""
        if self.vg.convention == "cell_center":
            return u
        if getattr(self, "problem", None) is not None:
            return self.vg.bc.trim_ghost_nodes(self.problem.pad_bc(u))
        return self.vg.bc.trim_ghost_nodes(self.vg.pad_zeros(u))

    def _handle_outputs(self, u, frame, time, vtk_out, verbose, plot_bounds, colormap):
        """solvers.py:152-187 (synchronous form, or queued on the exporter)."""
        if getattr(self, "_exporter", None) is not None:
            self._exporter.submit(u, self._output_field(u), frame, time, vtk_out)
            if verbose:
                self.profiler.update_memory_stats()
            if verbose == "plot":
                warnings.warn("live plotting is not part of the 'hip' backend; frames are exported to vf.fields")
            return
        flags = torch.zeros(2, dtype=torch.int32, device=u.device)
        mass = self._frame_sums(u, flags)
        self._export_fields(self._output_field(u))
        self.mass_history.append((frame, time, mass.cpu().numpy()))
        if verbose:
            self.profiler.update_memory_stats()
        if int(flags[0].item()):
            print(f"NaN detected in frame {frame} at time {time}. Aborting simulation.")
            sys.exit(1)
        if vtk_out:
            if not hasattr(self.vf, "export_to_vtk"):
                raise NotImplementedError("vtk_out needs a VoxelFields with export_to_vtk")
            self.vf.export_to_vtk(filename=self._vtk_name(frame), field_names=self.fieldnames)
        if verbose == "plot":
            warnings.warn("live plotting is not part of the 'hip' backend; frames are exported to vf.fields")


@dataclass
class TimeDependentSolver(BaseSolver):
    """solvers.py:189-208, with optional hipGraph replay of each output interval."""

    def _captured(self, u, step, n):
        """Replays n steps captured in a hipGraph; returns the new state."""
        key = (n, tuple(u.shape), u.dtype)
        if getattr(self, "_graph_key", None) != key or getattr(self, "_graph_step", None) is not step:
            self._graph_in = torch.empty_like(u)
            self._graph_in.copy_(u)
            # Warm-up on a copy: creates FFT plans and workspaces outside capture.
            s = torch.cuda.Stream(device=u.device)
            s.wait_stream(torch.cuda.current_stream(u.device))
            with torch.cuda.stream(s):
                step(0.0, self._graph_in.clone())
            torch.cuda.current_stream(u.device).wait_stream(s)
            self._graph = torch.cuda.CUDAGraph()
            with _lib.gc_paused(), torch.cuda.graph(self._graph):
                v = self._graph_in
                for _ in range(n):
                    v = step(0.0, v)
                self._graph_out = v
            self._graph_key = key
            # keeps the stepper, and with it the workspace the graph writes, alive
            self._graph_step = step
        self._graph_in.copy_(u)
        self._graph.replay()
        return self._graph_out

    def _run_loop(self, u, step, time_increment, frames, max_iters, vtk_out, verbose, plot_bounds, colormap):
        n_out = max_iters // frames
        frame = 0
        i = 0
        while i < max_iters:
            time = i * time_increment
            if i % n_out == 0:
                self._handle_outputs(u, frame, time, vtk_out, verbose, plot_bounds, colormap)
                frame += 1
            seg_end = min((i // n_out + 1) * n_out, max_iters)
            if self._capture and seg_end - i == n_out:
                u = self._captured(u, step, n_out)
            else:
                for k in range(i, seg_end):
                    u = step(k * time_increment, u)
            self._u_from_step = seg_end > i
            i = seg_end
        time = max_iters * time_increment
        self._handle_outputs(u, frame, time, vtk_out, verbose, plot_bounds, colormap)
        return u
