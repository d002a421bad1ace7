"""HIP backend adapter: the 'hip' counterpart of evoxels.voxelgrid.VoxelGridTorch.

``VoxelGridHip`` keeps the method names and array shapes of the reference's
``VoxelGrid`` / ``VoxelGridTorch`` (voxelgrid.py:19-246) so that problem
classes written against that interface run unchanged.  Fields are torch
tensors on a ROCm device; boundary padding, Laplacians, FFTs and reductions
are HIP kernels of libevoxels_hip.so.  Slicing, concatenation and the face
operators of fd_stencils.py are plain tensor views/elementwise ops exactly as
in the reference.  There is no CPU path: a missing GPU or library raises.
"""
from __future__ import annotations

import ctypes as C
import weakref
from dataclasses import dataclass
from typing import Tuple

import numpy as np
import torch

from . import _lib

__ = slice(None)
_i_ = slice(1, -1)
CENTER = (__, _i_, _i_, _i_)
LEFT = (__, slice(None, -2), _i_, _i_)
RIGHT = (__, slice(2, None), _i_, _i_)
BOTTOM = (__, _i_, slice(None, -2), _i_)
TOP = (__, _i_, slice(2, None), _i_)
BACK = (__, _i_, _i_, slice(None, -2))
FRONT = (__, _i_, _i_, slice(2, None))

PERIODIC3 = (("periodic", None),) * 3


@dataclass
class Grid:
    """voxelgrid.py:9-15."""
    shape: Tuple[int, int, int]
    origin: Tuple[float, float, float]
    spacing: Tuple[float, float, float]
    convention: str


class FDStencils:
    """Face/centre operators of fd_stencils.py:20-60 (elementwise tensor ops);
    laplace and normal_laplace dispatch to HIP kernels (fd_stencils.py:62-103)."""

    def to_x_face(self, field):
        return 0.5 * (field[:, 1:, :, :] + field[:, :-1, :, :])

    def to_y_face(self, field):
        return 0.5 * (field[:, :, 1:, :] + field[:, :, :-1, :])

    def to_z_face(self, field):
        return 0.5 * (field[:, :, :, 1:] + field[:, :, :, :-1])

    def grad_x_face(self, field):
        return (field[:, 1:, :, :] - field[:, :-1, :, :]) * self.div_dx[0]

    def grad_y_face(self, field):
        return (field[:, :, 1:, :] - field[:, :, :-1, :]) * self.div_dx[1]

    def grad_z_face(self, field):
        return (field[:, :, :, 1:] - field[:, :, :, :-1]) * self.div_dx[2]

    def grad_x_center(self, field):
        return 0.5 * (field[RIGHT] - field[LEFT]) * self.div_dx[0]

    def grad_y_center(self, field):
        return 0.5 * (field[TOP] - field[BOTTOM]) * self.div_dx[1]

    def grad_z_center(self, field):
        return 0.5 * (field[FRONT] - field[BACK]) * self.div_dx[2]

    def gradient_norm_squared(self, field):
        return self.grad_x_center(field) ** 2 + self.grad_y_center(field) ** 2 + \
            self.grad_z_center(field) ** 2

    def laplace(self, field):
        """7-point Laplacian of a padded field (HIP kernel)."""
        out = torch.empty(self._interior_shape(field), dtype=field.dtype, device=field.device)
        g = self._grid_desc(out.shape)
        _lib.call("evx_laplace_padded", _lib.dtype_code(field.dtype), _lib.ptr(field.contiguous()),
                  _lib.ptr(out), C.byref(g), _lib.stream_handle(field.device))
        return out

    def normal_laplace(self, field):
        """19-point normal Laplacian of a padded field (HIP kernel)."""
        out = torch.empty(self._interior_shape(field), dtype=field.dtype, device=field.device)
        g = self._grid_desc(out.shape)
        _lib.call("evx_normal_laplace_padded", _lib.dtype_code(field.dtype), _lib.ptr(field.contiguous()),
                  _lib.ptr(out), C.byref(g), _lib.stream_handle(field.device))
        return out

    @staticmethod
    def _interior_shape(padded):
        if padded.dim() != 4 or min(padded.shape[1:]) < 3:
            raise ValueError("expected a padded (C, Nx+2, Ny+2, Nz+2) field")
        c, a, b, d = padded.shape
        return (c, a - 2, b - 2, d - 2)


class CellCenteredBCs:
    """Ghost padding of boundary_conditions.py:5-83, evaluated in one HIP kernel."""

    def __init__(self, vg):
        self.vg = vg

    def pad_periodic(self, field):
        return self.vg.pad_bc(field, PERIODIC3)

    def pad_dirichlet_periodic(self, field, bc0=0, bc1=0):
        return self.vg.pad_bc(field, (("dirichlet", (bc0, bc1)), ("periodic", None), ("periodic", None)))

    def pad_zero_flux_periodic(self, field):
        return self.vg.pad_bc(field, (("neumann", None), ("periodic", None), ("periodic", None)))

    def pad_bc(self, field, bc):
        return self.vg.pad_bc(field, bc)

    def pad_fft_periodic(self, field):
        return field

    def pad_fft_dirichlet_periodic(self, field):
        return torch.cat((field, -torch.flip(field, [1])), 1)

    def pad_fft_zero_flux_periodic(self, field):
        return torch.cat((field, torch.flip(field, [1])), 1)

    def trim_boundary_nodes(self, field):
        return field

    def trim_ghost_nodes(self, field):
        if field[0, _i_, _i_, _i_].shape == self.vg.shape:
            return field[:, _i_, _i_, _i_]
        raise ValueError(f"The provided field has the wrong shape {self.vg.shape}.")


class StaggeredXBCs:
    """Ghost rules of the staggered_x convention (boundary_conditions.py:
    85-156): x carries the boundary nodes, so the x ghost plane of the
    trimmed (interior) field is the node itself.  The two paddings the
    reference implements run in one HIP kernel; the others raise as there."""

    def __init__(self, vg):
        self.vg = vg

    def pad_periodic_BC_staggered_x(self, field):
        raise NotImplementedError

    def pad_dirichlet_periodic(self, field, bc0=0, bc1=0):
        """Periodic pad, then the x ghost planes set to bc0 / bc1 (boundary
        node values; scalars or fields over the padded boundary plane)."""
        return self.vg.pad_staggered_x(field, "dirichlet", (bc0, bc1))

    def pad_zero_flux_periodic(self, field):
        """Periodic pad, then x ghost = 4/3 u1 - 1/3 u2 (zero slope of the
        quadratic through the ghost node and the first two interior nodes)."""
        return self.vg.pad_staggered_x(field, "neumann", None)

    def pad_bc(self, field, bc):
        raise NotImplementedError

    def pad_fft_periodic(self, field):
        raise NotImplementedError

    def pad_fft_dirichlet_periodic(self, field):
        """Odd extension with the zero boundary nodes: length 2*(Nx-1) for a
        grid of Nx nodes (boundary_conditions.py:137-140)."""
        bc = torch.zeros_like(field[:, 0:1])
        return torch.cat((field, bc, -torch.flip(field, [1]), bc), 1)

    def pad_fft_zero_flux_periodic(self, field):
        raise NotImplementedError

    def trim_boundary_nodes(self, field):
        if field.shape[1] == self.vg.shape[0]:
            return field[:, _i_, :, :]
        raise ValueError(f"The provided field must have the shape {self.vg.shape}.")

    def trim_ghost_nodes(self, field):
        if field[0, :, _i_, _i_].shape == self.vg.shape:
            return field[:, :, _i_, _i_]
        raise ValueError(f"The provided field has the wrong shape {self.vg.shape}.")


class _FFTPlan:
    """Owns one evx_fft_plan (twiddle tables only)."""

    def __init__(self, dtype_code, c, nx, ny, nz):
        lib = _lib.load()
        h = C.c_void_p()
        _lib.check(lib.evx_fft_plan_create(C.byref(h), dtype_code, c, nx, ny, nz), "evx_fft_plan_create")
        self.handle = h
        self.shape = (c, nx, ny, nz)
        ws = C.c_size_t()
        _lib.check(lib.evx_fft_plan_workspace_bytes(h, C.byref(ws)), "evx_fft_plan_workspace_bytes")
        self.workspace_bytes = ws.value
        self._fin = weakref.finalize(self, lib.evx_fft_plan_destroy, h)


def _hip_device(device) -> torch.device:
    dev = torch.device(device)
    if dev.type != "cuda":
        raise RuntimeError(f"the 'hip' backend runs on a ROCm GPU; got device {device!r}")
    if not torch.cuda.is_available():
        raise RuntimeError("the 'hip' backend needs a ROCm GPU, and none is visible to torch")
    if dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    return dev


class VoxelGridHip(FDStencils):
    """Grid adapter on a ROCm device (voxelgrid.py:19-152,155-246)."""

    def __init__(self, grid: Grid, precision="float32", device="cuda"):
        _lib.load()
        self.torch = torch
        self.lib = torch
        self.device = _hip_device(device)
        if precision == "float32":
            self.precision = torch.float32
        elif precision == "float64":
            self.precision = torch.float64
        else:
            raise ValueError(f"Unsupported precision: {precision}")
        self.shape = tuple(int(n) for n in grid.shape)
        self.origin = grid.origin
        self.spacing = grid.spacing
        self.convention = grid.convention
        if self.convention not in ("cell_center", "staggered_x"):
            raise ValueError("Chosen convention must be cell_center or staggered_x.")
        # voxelgrid.py:29-30: spacing cast to the working precision, then 1/x and 1/x**2.
        sp = torch.tensor(np.array(self.spacing), dtype=self.precision)
        dd, dd2 = 1 / sp, 1 / sp ** 2
        self.div_dx_host = tuple(float(v) for v in dd)
        self.div_dx2_host = tuple(float(v) for v in dd2)
        self.div_dx = dd.to(self.device)
        self.div_dx2 = dd2.to(self.device)
        self.bc = CellCenteredBCs(self) if self.convention == "cell_center" else StaggeredXBCs(self)
        self._plans = {}
        self._reduce_ws = None

    # ---------------------------------------------------------- descriptors
    def _grid_desc(self, shape4):
        return _lib.make_grid(shape4, self.div_dx_host, self.div_dx2_host)

    def plan(self, shape4, dtype=None):
        dtype = dtype or self.precision
        key = (dtype, tuple(int(v) for v in shape4))
        if key not in self._plans:
            self._plans[key] = _FFTPlan(_lib.dtype_code(dtype), *key[1])
        return self._plans[key]

    def reduce_workspace(self, channels):
        lib = _lib.load()
        n = C.c_size_t()
        _lib.check(lib.evx_reduce_workspace_bytes(channels, C.byref(n)), "evx_reduce_workspace_bytes")
        if self._reduce_ws is None or self._reduce_ws.numel() < n.value:
            self._reduce_ws = torch.empty(n.value, dtype=torch.uint8, device=self.device)
        return self._reduce_ws

    # -------------------------------------------------------------- arrays
    def to_backend(self, np_arr):
        return torch.tensor(np_arr, dtype=self.precision, device=self.device)

    def to_numpy(self, field):
        return field.cpu().numpy()

    def pad_bc(self, field, bc):
        if any(kind == "dirichlet" and any(isinstance(v, torch.Tensor) for v in values) for kind, values in bc):
            return self._pad_bc_field_values(field, bc)
        field = field.contiguous()
        c, nx, ny, nz = field.shape
        out = torch.empty((c, nx + 2, ny + 2, nz + 2), dtype=field.dtype, device=field.device)
        g = self._grid_desc(field.shape)
        b = _lib.make_bc(bc)
        _lib.call("evx_pad_bc", _lib.dtype_code(field.dtype), _lib.ptr(field), _lib.ptr(out), C.byref(g),
                  C.byref(b), _lib.stream_handle(field.device))
        return out

    def _pad_bc_field_values(self, field, bc):
        """pad_bc with Dirichlet values given as fields on the boundary planes
        (the smoothed-boundary problem scales them by its mask): the periodic
        pad in HIP, then the per-axis ghost rules in the order x, y, z as
        boundary_conditions.py:33-59 applies them."""
        padded = self.pad_periodic(field)
        for a, (kind, values) in enumerate(bc):
            if kind == "periodic":
                continue
            idx = lambda i: (slice(None),) * (a + 1) + (i,)  # noqa: E731
            if kind == "neumann":
                padded[idx(0)] = padded[idx(1)]
                padded[idx(-1)] = padded[idx(-2)]
            else:
                padded[idx(0)] = 2.0 * values[0] - padded[idx(1)]
                padded[idx(-1)] = 2.0 * values[1] - padded[idx(-2)]
        return padded

    def pad_staggered_x(self, field, kind, values):
        """staggered_x x padding with y, z periodic: one HIP kernel for scalar
        node values and the zero-flux rule; node values given as fields (the
        smoothed-boundary problem scales them by its mask) are written over
        the periodic pad."""
        field = field.contiguous()
        if kind == "dirichlet" and any(isinstance(v, torch.Tensor) for v in values):
            padded = self.pad_periodic(field)
            padded[:, 0, :, :] = values[0]
            padded[:, -1, :, :] = values[1]
            return padded
        bc = ((kind, tuple(values) if values is not None else None), ("periodic", None), ("periodic", None))
        c, nx, ny, nz = field.shape
        out = torch.empty((c, nx + 2, ny + 2, nz + 2), dtype=field.dtype, device=field.device)
        g = self._grid_desc(field.shape)
        b = _lib.make_bc(bc, "staggered_x")
        _lib.call("evx_pad_bc", _lib.dtype_code(field.dtype), _lib.ptr(field), _lib.ptr(out), C.byref(g),
                  C.byref(b), _lib.stream_handle(field.device))
        return out

    def pad_periodic(self, field):
        return self.pad_bc(field, PERIODIC3)

    def pad_zeros(self, field):
        return torch.nn.functional.pad(field, (1, 1, 1, 1, 1, 1), mode="constant", value=0)

    def rfftn(self, field, shape):
        """Per-channel real 3-D FFT (voxelgrid.py:203-204)."""
        field = field.contiguous()
        if tuple(shape) != tuple(field.shape):
            raise NotImplementedError("rfftn with a shape different from the field's")
        p = self.plan(field.shape, field.dtype)
        c, nx, ny, nz = field.shape
        cdt = torch.complex64 if field.dtype == torch.float32 else torch.complex128
        spec = torch.empty((c, nx, ny, nz // 2 + 1), dtype=cdt, device=field.device)
        _lib.call("evx_rfftn", p.handle, _lib.ptr(field), _lib.ptr(spec), _lib.stream_handle(field.device))
        return spec

    @staticmethod
    def _negate_freq(a, dims):
        """a at the negated frequency index (-k mod n) along ``dims``."""
        return torch.roll(torch.flip(a, dims), shifts=[1] * len(dims), dims=dims)

    def fftn(self, field, shape):
        """Full complex 3-D FFT per channel of a real field (voxelgrid.py:
        200-201): the half spectrum of the HIP R2C, completed by Hermitian
        symmetry X[k] = conj(X[-k])."""
        if field.is_complex():
            raise NotImplementedError("fftn of a complex field: the 'hip' transforms are real-to-complex")
        c, nx, ny, nz = (int(v) for v in shape)
        half = self.rfftn(field, shape)
        nh = nz // 2 + 1
        full = torch.empty((c, nx, ny, nz), dtype=half.dtype, device=half.device)
        full[..., :nh] = half
        if nz > nh:
            mirrored = torch.conj(self._negate_freq(half[..., 1:nz - nh + 1], [1, 2]))
            full[..., nh:] = torch.flip(mirrored, [3])
        return full

    def real_of_ifftn(self, field, shape):
        """real(ifftn(F)) per channel (voxelgrid.py:209-210) as the C2R of the
        Hermitian part (F + conj(F(-k)))/2, which has the same real inverse."""
        c, nx, ny, nz = (int(v) for v in shape)
        if tuple(field.shape) != (c, nx, ny, nz):
            raise NotImplementedError("real_of_ifftn with a spectrum shape that does not match `shape`")
        herm = 0.5 * (field + torch.conj(self._negate_freq(field, [1, 2, 3])))
        return self.irfftn(herm[..., :nz // 2 + 1].contiguous(), shape, overwrite=True)

    def irfftn(self, spec, shape, overwrite=False):
        """Per-channel inverse real 3-D FFT (voxelgrid.py:206-207).  The
        transform runs in place on its input, so the spectrum is copied first
        unless the caller passes ``overwrite=True`` for a temporary."""
        c, nx, ny, nz = (int(v) for v in shape)
        if spec.shape != (c, nx, ny, nz // 2 + 1):
            raise NotImplementedError("irfftn with a spectrum shape that does not match `shape`")
        rdt = torch.float32 if spec.dtype == torch.complex64 else torch.float64
        work = spec if overwrite and spec.is_contiguous() else spec.contiguous().clone()
        p = self.plan((c, nx, ny, nz), rdt)
        out = torch.empty((c, nx, ny, nz), dtype=rdt, device=spec.device)
        _lib.call("evx_irfftn", p.handle, _lib.ptr(work), _lib.ptr(out), _lib.stream_handle(spec.device))
        return out

    def expand_dim(self, field, dim):
        return field.unsqueeze(dim)

    def squeeze(self, field, dim):
        return torch.squeeze(field, dim)

    def concatenate(self, fieldlist, dim):
        return torch.cat(fieldlist, dim=dim)

    def set(self, field, index, value):
        field[index] = value
        return field

    def mean(self, field, dim=None, keepdim=False):
        return torch.mean(field, dim=dim, keepdim=keepdim)

    def sum(self, field, dim=None, keepdim=False):
        return torch.sum(field, dim=dim, keepdim=keepdim)

    def arange(self, start, stop):
        return torch.arange(start, stop, dtype=self.precision, device=self.device)

    def argmax(self, field, dim=None, keepdim=False):
        return torch.argmax(field, dim=dim, keepdim=keepdim)

    def cumsum(self, field, dim=None):
        return torch.cumsum(field, dim=dim)

    def sort(self, field, dim=0, descending=False):
        vals, _ = torch.sort(field, dim=dim, descending=descending)
        return vals

    def take_along_dim(self, field, idx, dim=0):
        return torch.take_along_dim(field, idx, dim=dim)

    def meshgrid(self):
        """Full 3-D mesh of the cell coordinates (voxelgrid.py:92-95)."""
        return tuple(torch.meshgrid(*self.axes(), indexing="ij"))

    def fft_mesh(self):
        return tuple(torch.meshgrid(*self.fft_axes(), indexing="ij"))

    def fft_k_squared(self):
        """Full-spectrum k^2 (voxelgrid.py:105-108), float32 as fftfreq returns."""
        kx, ky, kz = torch.meshgrid(*self.fft_axes(), indexing="ij")
        return kx ** 2 + ky ** 2 + kz ** 2

    # ---------------------------------------------------------- k-space
    def axes(self):
        return tuple(torch.arange(0, n, dtype=self.precision, device=self.device) * self.spacing[i]
                     + self.origin[i] for i, n in enumerate(self.shape))

    def fft_axes(self):
        return tuple(2 * np.pi * torch.fft.fftfreq(n, d, device=self.device)
                     for n, d in zip(self.shape, self.spacing))

    def rfft_axes(self):
        return tuple(2 * np.pi * torch.fft.rfftfreq(n, d, device=self.device)
                     for n, d in zip(self.shape, self.spacing))

    def rfft_k_squared(self):
        """float32 table, as the reference builds it (voxelgrid.py:110-114)."""
        a_x, a_y, _ = self.fft_axes()
        _, _, a_z = self.rfft_axes()
        kx, ky, kz = torch.meshgrid(a_x, a_y, a_z, indexing="ij")
        return kx ** 2 + ky ** 2 + kz ** 2

    def rfft_k_squared_nonperiodic(self):
        """voxelgrid.py:116-124: x frequencies of the mirror extension, 2*Nx
        points (cell_center) or 2*Nx-2 (staggered_x, Nx nodes)."""
        n = 2 * self.shape[0] if self.convention == "cell_center" else 2 * self.shape[0] - 2
        a_x = 2 * np.pi * torch.fft.fftfreq(n, d=self.spacing[0], device=self.device)
        _, a_y, _ = self.fft_axes()
        _, _, a_z = self.rfft_axes()
        kx, ky, kz = torch.meshgrid(a_x, a_y, a_z, indexing="ij")
        return kx ** 2 + ky ** 2 + kz ** 2

    # ---------------------------------------------------------- fields
    def init_scalar_field(self, array):
        return self.expand_dim(self.to_backend(array), 0)

    def export_scalar_field_to_numpy(self, field):
        return self.to_numpy(self.squeeze(field, 0))

    def channel_sums(self, field, nonfinite=None):
        """Compensated float64 sum per channel (HIP); optional NaN/Inf flags."""
        field = field.contiguous()
        c = field.shape[0]
        out = torch.empty(c, dtype=torch.float64, device=field.device)
        ws = self.reduce_workspace(c)
        _lib.call("evx_sum_f64", _lib.dtype_code(field.dtype), _lib.ptr(field), c, field[0].numel(),
                  _lib.ptr(out), _lib.ptr(nonfinite), _lib.ptr(ws), _lib.stream_handle(field.device))
        return out

    def average(self, field):
        """Spatial mean per channel (voxelgrid.py:137-152); compensated float64
        accumulation, returned in the working precision."""
        if tuple(field.shape[1:]) != self.shape:
            raise ValueError(f"The provided field must have the shape {self.shape}.")
        if self.convention == "cell_center":
            s = self.channel_sums(field)
            return (s / float(np.prod(self.shape))).to(field.dtype)
        # staggered_x: the first and last x slices are half cells
        s = self.channel_sums(field[:, 1:-1]) + 0.5 * self.channel_sums(field[:, 0:1]) \
            + 0.5 * self.channel_sums(field[:, -1:])
        return (s / float((self.shape[0] - 1) * self.shape[1] * self.shape[2])).to(field.dtype)

    def ch_energy(self, field, eps, bc=PERIODIC3):
        """Discrete CH free energy per channel (float64)."""
        field = field.contiguous()
        c = field.shape[0]
        out = torch.empty(c, dtype=torch.float64, device=field.device)
        ws = self.reduce_workspace(c)
        g = self._grid_desc(field.shape)
        b = _lib.make_bc(bc)
        dv = float(np.prod(np.asarray(self.spacing, dtype=np.float64)))
        _lib.call("evx_ch_energy", _lib.dtype_code(field.dtype), _lib.ptr(field), C.byref(g), C.byref(b),
                  float(eps), dv, _lib.ptr(out), _lib.ptr(ws), _lib.stream_handle(field.device))
        return out
