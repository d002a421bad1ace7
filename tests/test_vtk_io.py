"""Legacy-VTK export (the pyvista-free counterpart of voxelfields.py:196-214):
geometry of the cell image, x-fastest cell order, binary big-endian and ASCII
encodings, dtype mapping and argument errors.  CPU only; the files are read
back by a parser written from the legacy format's specification."""
import numpy as np
import pytest

import evoxels_amd as evx
from evoxels_amd.vtk_io import AsyncWriter, image_geometry, write_structured_points


def parse_legacy(path):
    """Minimal reader of a legacy STRUCTURED_POINTS file with CELL_DATA
    SCALARS sections: returns (header dict, {name: flat array})."""
    raw = open(path, "rb").read()
    pos = 0

    def line():
        nonlocal pos
        end = raw.index(b"\n", pos)
        out = raw[pos:end].decode("ascii")
        pos = end + 1
        return out

    head = {"version": line(), "title": line(), "encoding": line(), "dataset": line()}
    for _ in range(4):
        parts = line().split()
        head[parts[0]] = parts[1:]
    types = {"float": ">f4", "double": ">f8", "int": ">i4", "unsigned_char": "u1", "short": ">i2",
             "long": ">i8", "char": "i1", "unsigned_short": ">u2", "unsigned_int": ">u4", "unsigned_long": ">u8"}
    n = int(head["CELL_DATA"][0])
    arrays = {}
    while pos < len(raw):
        kw, name, vtype, ncomp = line().split()
        assert kw == "SCALARS" and ncomp == "1"
        assert line() == "LOOKUP_TABLE default"
        if head["encoding"] == "BINARY":
            dt = np.dtype(types[vtype])
            arrays[name] = (np.frombuffer(raw, dtype=dt, count=n, offset=pos), vtype)
            pos += n * dt.itemsize
            assert raw[pos:pos + 1] == b"\n"
            pos += 1
        else:
            vals = [line() for _ in range(n)]
            arrays[name] = (np.array([float(v) for v in vals]), vtype)
    return head, arrays


@pytest.mark.parametrize("convention", ["cell_center", "staggered_x"])
def test_export_geometry_and_cell_order(tmp_path, convention):
    shape = (5, 4, 3)
    vf = evx.VoxelFields(shape, (2.5, 1.0, 0.6), convention=convention)
    i, j, k = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    vf.add_field("code", (10000 * i + 100 * j + k).astype(np.float64))
    vf.add_field("c", np.random.default_rng(0).random(shape).astype(np.float32))
    path = tmp_path / "f.vtk"
    vf.export_to_vtk(str(path))
    head, arrays = parse_legacy(path)
    assert head["version"] == "# vtk DataFile Version 3.0" and head["encoding"] == "BINARY"
    assert head["dataset"] == "DATASET STRUCTURED_POINTS"
    # voxelfields.py:205-209: N+1 points per axis, corner at origin - spacing/2
    assert [int(v) for v in head["DIMENSIONS"]] == [6, 5, 4]
    assert [float(v) for v in head["SPACING"]] == list(vf.spacing)
    assert [float(v) for v in head["ORIGIN"]] == [o - s / 2 for o, s in zip(vf.origin, vf.spacing)]
    assert int(head["CELL_DATA"][0]) == 60
    code, vtype = arrays["code"]
    assert vtype == "double"
    # VTK cell index = i + nx*(j + ny*k): x varies fastest (Fortran order)
    for (a, b, c) in [(0, 0, 0), (4, 0, 0), (0, 3, 0), (1, 2, 2), (4, 3, 2)]:
        assert code[a + 5 * (b + 4 * c)] == 10000 * a + 100 * b + c
    cvals, ctype = arrays["c"]
    assert ctype == "float"
    np.testing.assert_array_equal(cvals, vf.fields["c"].reshape(-1, order="F"))


def test_ascii_encoding_round_trips(tmp_path):
    a = np.random.default_rng(1).random((3, 2, 4))
    path = tmp_path / "a.vtk"
    write_structured_points(str(path), {"phi": a}, (0.5, 0.25, 1.0), (0.25, 0.125, 0.5), binary=False)
    head, arrays = parse_legacy(path)
    assert head["encoding"] == "ASCII"
    np.testing.assert_array_equal(arrays["phi"][0], a.reshape(-1, order="F"))


def test_dtypes_and_errors(tmp_path):
    p = str(tmp_path / "x.vtk")
    write_structured_points(p, {"lab": np.arange(24, dtype=np.int32).reshape(2, 3, 4),
                                "m": np.ones((2, 3, 4), dtype=np.uint8)}, (1, 1, 1), (0.5, 0.5, 0.5))
    _, arrays = parse_legacy(p)
    assert arrays["lab"][1] == "int" and arrays["m"][1] == "unsigned_char"
    np.testing.assert_array_equal(arrays["lab"][0], np.arange(24).reshape(2, 3, 4).reshape(-1, order="F"))
    with pytest.raises(ValueError):
        write_structured_points(p, {}, (1, 1, 1), (0, 0, 0))
    with pytest.raises(ValueError):
        write_structured_points(p, {"a": np.ones((2, 2, 2)), "b": np.ones((2, 2, 3))}, (1, 1, 1), (0, 0, 0))
    with pytest.raises(ValueError):
        write_structured_points(p, {"a b": np.ones((2, 2, 2))}, (1, 1, 1), (0, 0, 0))
    with pytest.raises(TypeError):
        write_structured_points(p, {"z": np.ones((2, 2, 2), dtype=np.complex64)}, (1, 1, 1), (0, 0, 0))


def test_async_writer_writes_in_order_and_reports_errors(tmp_path):
    w = AsyncWriter()
    arrs = [np.full((2, 2, 2), float(i)) for i in range(5)]
    for i, a in enumerate(arrs):
        w.submit(str(tmp_path / f"f_{i:03d}.vtk"), {"c": a}, (1, 1, 1), (0.5, 0.5, 0.5))
    w.wait()
    for i in range(5):
        _, arrays = parse_legacy(tmp_path / f"f_{i:03d}.vtk")
        assert np.all(arrays["c"][0] == i)
    w.submit(str(tmp_path / "bad.vtk"), {"bad name": arrs[0]}, (1, 1, 1), (0, 0, 0))
    with pytest.raises(ValueError):
        w.wait()
    w.close()


def test_image_geometry():
    dims, sp, corner = image_geometry((4, 5, 6), (0.1, 0.2, 0.3), (0.05, 0.1, 0.15))
    assert dims == (5, 6, 7) and sp == (0.1, 0.2, 0.3)
    assert np.allclose(corner, (0.0, 0.0, 0.0))
