"""Legacy-VTK export of voxel fields without pyvista.

``VoxelFields.export_to_vtk`` of the reference (voxelfields.py:196-214)
builds a ``pyvista.ImageData`` whose points are the cell corners
(dimensions N+1 per axis, origin at ``origin - spacing/2``) and stores every
field as cell data flattened in Fortran order (x fastest), then saves it.
This module writes the same image as a legacy VTK ``STRUCTURED_POINTS`` file
(format 3.0, read by ParaView, VisIt and VTK itself) directly from numpy:
binary big-endian by default, ASCII on request.  ``export_async`` hands the
write to one background thread so that frame files are written while the
solver keeps the GPU busy.
"""
from __future__ import annotations

import threading
from concurrent.futures import Future, ThreadPoolExecutor
from typing import Mapping, Sequence

import numpy as np

# numpy dtype -> (legacy VTK type name, big-endian dtype written to disk)
_VTK_TYPES = {
    np.dtype(np.float32): ("float", ">f4"),
    np.dtype(np.float64): ("double", ">f8"),
    np.dtype(np.int8): ("char", "i1"),
    np.dtype(np.uint8): ("unsigned_char", "u1"),
    np.dtype(np.int16): ("short", ">i2"),
    np.dtype(np.uint16): ("unsigned_short", ">u2"),
    np.dtype(np.int32): ("int", ">i4"),
    np.dtype(np.uint32): ("unsigned_int", ">u4"),
    np.dtype(np.int64): ("long", ">i8"),
    np.dtype(np.uint64): ("unsigned_long", ">u8"),
    np.dtype(np.bool_): ("unsigned_char", "u1"),
}


def image_geometry(shape: Sequence[int], spacing: Sequence[float], origin: Sequence[float]):
    """Point dimensions, spacing and corner origin of the cell image
    (voxelfields.py:205-209)."""
    dims = tuple(int(n) + 1 for n in shape)
    corner = tuple(float(o) - float(s) / 2 for o, s in zip(origin, spacing))
    return dims, tuple(float(s) for s in spacing), corner


def write_structured_points(filename: str, fields: Mapping[str, np.ndarray], spacing, origin,
                            binary: bool = True, title: str = "evoxels fields") -> None:
    """Writes ``fields`` (name -> 3-D array, all of one shape) as cell data of
    a STRUCTURED_POINTS image.  ``origin`` is the first cell centre, as in
    VoxelFields."""
    if not fields:
        raise ValueError("no fields to export")
    shapes = {np.shape(a) for a in fields.values()}
    if len(shapes) != 1 or len(next(iter(shapes))) != 3:
        raise ValueError("all exported fields must be 3-D arrays of one shape")
    shape = next(iter(shapes))
    dims, sp, corner = image_geometry(shape, spacing, origin)
    n_cells = int(np.prod(shape))
    fmt = lambda v: repr(float(v))  # noqa: E731  (round-trips exactly)
    head = [
        "# vtk DataFile Version 3.0",
        title.replace("\n", " ")[:255],
        "BINARY" if binary else "ASCII",
        "DATASET STRUCTURED_POINTS",
        "DIMENSIONS {} {} {}".format(*dims),
        "ORIGIN " + " ".join(fmt(v) for v in corner),
        "SPACING " + " ".join(fmt(v) for v in sp),
        f"CELL_DATA {n_cells}",
    ]
    with open(filename, "wb") as fh:
        fh.write(("\n".join(head) + "\n").encode("ascii"))
        for name, arr in fields.items():
            a = np.asarray(arr)
            if a.dtype not in _VTK_TYPES:
                raise TypeError(f"field {name!r}: dtype {a.dtype} has no legacy VTK type")
            vtk_name, disk = _VTK_TYPES[a.dtype]
            if any(c.isspace() for c in name) or not name:
                raise ValueError(f"field name {name!r} must be non-empty without whitespace")
            fh.write(f"SCALARS {name} {vtk_name} 1\nLOOKUP_TABLE default\n".encode("ascii"))
            flat = a.reshape(-1, order="F")  # x fastest, as VTK orders cells
            if binary:
                fh.write(flat.astype(disk, copy=False).tobytes())
                fh.write(b"\n")
            else:
                if a.dtype.kind == "f":
                    text = "\n".join(repr(float(v)) for v in flat)
                else:
                    text = "\n".join(str(int(v)) for v in flat)
                fh.write(text.encode("ascii") + b"\n")


class AsyncWriter:
    """One background thread writing VTK files in submission order."""

    def __init__(self):
        self._pool = ThreadPoolExecutor(max_workers=1, thread_name_prefix="evoxels-vtk")
        self._futures: list[Future] = []
        self._lock = threading.Lock()

    def submit(self, filename, fields, spacing, origin, binary=True) -> Future:
        fut = self._pool.submit(write_structured_points, filename, dict(fields), spacing, origin, binary)
        with self._lock:
            self._futures.append(fut)
        return fut

    def wait(self) -> None:
        """Blocks until every submitted file is written; re-raises the first error."""
        with self._lock:
            futs, self._futures = self._futures, []
        for f in futs:
            f.result()

    def close(self) -> None:
        self.wait()
        self._pool.shutdown(wait=True)
