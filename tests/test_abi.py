"""The C-ABI library loads and exports every entry point include/evoxels_hip.h
declares.  CPU only: only calls that return before any HIP call are made."""
import ctypes as C
import re
from pathlib import Path

import pytest

from evoxels_amd import _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "evoxels_hip.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(evx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_functions():
    names = declared_functions()
    assert "evx_ch_imex_step" in names and "evx_rd_rhs" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_binding_table_matches_header():
    assert sorted(_lib.SIGNATURES) == declared_functions()


def test_version_and_error_paths_without_gpu():
    lib = _lib.load()
    assert lib.evx_version() == 2
    g = _lib.make_grid((1, 4, 4, 4), (1, 1, 1), (1, 1, 1))
    b = _lib.make_bc((("periodic", None),) * 3)
    e = _lib.make_epilogue()
    st = lib.evx_rd_rhs(7, C.c_void_p(16), C.c_void_p(32), C.byref(g), C.byref(b), 1.0, C.byref(e), None)
    assert st == _lib.EVX_ERR_INVALID and "dtype" in _lib.last_error()
    st = lib.evx_ch_rhs(0, C.c_void_p(16), C.c_void_p(16), C.byref(g), C.byref(b), 3.0, 1.0, C.byref(e), None)
    assert st == _lib.EVX_ERR_INVALID and "alias" in _lib.last_error()
    bad = _lib.make_grid((1, 0, 4, 4), (1, 1, 1), (1, 1, 1))
    st = lib.evx_pad_bc(0, C.c_void_p(16), C.c_void_p(32), C.byref(bad), C.byref(b), None)
    assert st == _lib.EVX_ERR_INVALID and "positive" in _lib.last_error()
    b.kind[1] = 9
    st = lib.evx_pad_bc(0, C.c_void_p(16), C.c_void_p(32), C.byref(g), C.byref(b), None)
    assert st == _lib.EVX_ERR_INVALID and "Unsupported BC" in _lib.last_error()
    assert lib.evx_axpy(0, None, None, 1.0, None, 0, None) == _lib.EVX_OK
    h = C.c_void_p()
    assert lib.evx_fft_plan_create(C.byref(h), 0, 1, 0, 4, 4) == _lib.EVX_ERR_INVALID
    assert lib.evx_fft_plan_create(C.byref(h), 0, 1, 37, 4, 4) == _lib.EVX_ERR_UNSUPPORTED
    assert "prime factor 37" in _lib.last_error()


def test_batch_argument_checks_without_gpu():
    """The batched entry points reject bad arguments before touching the
    device: null plans and outputs, float64, grid counts out of range."""
    lib = _lib.load()
    p = C.c_void_p()
    shapes = (C.c_int64 * 3)(32, 32, 32)
    spacing = (C.c_double * 3)(1.0, 1.0, 1.0)
    assert lib.evx_batch_plan_create(None, 0, 1, shapes, spacing) == _lib.EVX_ERR_INVALID
    assert lib.evx_batch_plan_create(C.byref(p), 1, 1, shapes, spacing) == _lib.EVX_ERR_UNSUPPORTED
    assert "float32" in _lib.last_error()
    assert lib.evx_batch_plan_create(C.byref(p), 0, 0, shapes, spacing) == _lib.EVX_ERR_INVALID
    assert "n_inst" in _lib.last_error()
    i = C.c_int()
    ms = C.c_double()
    for st in (lib.evx_batch_set_tiles(None, 8), lib.evx_batch_set_fft_group(None, 1),
               lib.evx_batch_get_fft_group(None, C.byref(i)), lib.evx_batch_set_one_launch(None, 1),
               lib.evx_batch_get_one_launch(None, C.byref(i)),
               lib.evx_batch_autotune(None, None, None, 1, None, C.byref(ms))):
        assert st == _lib.EVX_ERR_INVALID


def test_check_maps_status_to_reference_exceptions():
    with pytest.raises(ValueError):
        _lib.check(_lib.EVX_ERR_INVALID)
    with pytest.raises(NotImplementedError):
        _lib.check(_lib.EVX_ERR_UNSUPPORTED)
    with pytest.raises(_lib.HipBackendError):
        _lib.check(_lib.EVX_ERR_HIP)
