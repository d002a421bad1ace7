"""ctypes binding of libevoxels_hip.so (the C ABI declared in include/evoxels_hip.h).

torch is imported before the library is opened: both torch's bundled HIP
runtime and the system one carry the soname ``libamdhip64.so.7``, so the
dynamic loader binds the library to the runtime torch already loaded and the
kernels run on torch's streams and memory.  There is no fallback: if the
library is missing or cannot be loaded, every entry point raises.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import gc
import os
from pathlib import Path

import torch  # noqa: F401  (must be loaded first, see module docstring)

LIB_PATH = Path(__file__).resolve().parent / "libevoxels_hip.so"

EVX_OK = 0
EVX_ERR_INVALID = -1
EVX_ERR_HIP = -2
EVX_ERR_UNSUPPORTED = -3
EVX_ERR_NOMEM = -4

EVX_F32 = 0
EVX_F64 = 1

BC_KIND = {"periodic": 0, "neumann": 1, "dirichlet": 2}

OUT_RHS, OUT_EULER, OUT_RK4_FIRST, OUT_RK4_MID, OUT_RK4_LAST = range(5)

OPT_SPECIALIZED = 1
OPT_EXACT_SYMBOL = 2
OPT_SLABS = 3
OPT_FUSED_RHS = 4
OPT_STEP_SUMS = 5


class EvxBC(C.Structure):
    _fields_ = [("kind", C.c_int32 * 3), ("lo", C.c_double * 3), ("hi", C.c_double * 3)]


class EvxGrid(C.Structure):
    _fields_ = [("channels", C.c_int64), ("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
                ("div_dx", C.c_double * 3), ("div_dx2", C.c_double * 3)]


class EvxEpilogue(C.Structure):
    _fields_ = [("mode", C.c_int32), ("coef", C.c_double), ("base", C.c_void_p),
                ("acc_in", C.c_void_p), ("acc_out", C.c_void_p)]


class EvxImexParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("spacing", C.c_double * 3), ("eps", C.c_double),
                ("D", C.c_double), ("A", C.c_double), ("chan_params", C.c_void_p)]


_P = C.c_void_p
_I32 = C.c_int32
_I64 = C.c_int64
_D = C.c_double

# name -> argtypes (all return int32 status unless listed in _RESTYPE)
SIGNATURES = {
    "evx_version": [],
    "evx_last_error": [],
    "evx_pad_bc": [_I32, _P, _P, C.POINTER(EvxGrid), C.POINTER(EvxBC), _P],
    "evx_laplace_padded": [_I32, _P, _P, C.POINTER(EvxGrid), _P],
    "evx_normal_laplace_padded": [_I32, _P, _P, C.POINTER(EvxGrid), _P],
    "evx_rd_rhs": [_I32, _P, _P, C.POINTER(EvxGrid), C.POINTER(EvxBC), _D, C.POINTER(EvxEpilogue), _P],
    "evx_ch_rhs": [_I32, _P, _P, C.POINTER(EvxGrid), C.POINTER(EvxBC), _D, _D, C.POINTER(EvxEpilogue), _P],
    "evx_coupled_rd_rhs": [_I32, _P, _P, C.POINTER(EvxGrid), C.POINTER(EvxBC), _D, _D, _D, _D,
                           C.POINTER(EvxEpilogue), _P],
    "evx_ac_rhs": [_I32, _P, _P, C.POINTER(EvxGrid), C.POINTER(EvxBC), _D, _D, _D, _D, _D,
                   C.POINTER(EvxEpilogue), _P],
    "evx_axpy": [_I32, _P, _P, _D, _P, _I64, _P],
    "evx_simplex_project": [_I32, _P, _P, _I32, _I64, _I32, _P],
    "evx_multiphase_ac_rhs_padded": [_I32, _P, _P, _P, C.POINTER(EvxGrid), _D, _D, _D, _D, _P],
    "evx_sbm_rhs_padded": [_I32, _P, _P, _P, _D, _P, C.POINTER(EvxGrid), _D, _D, _P],
    "evx_fft_plan_create": [C.POINTER(_P), _I32, _I64, _I64, _I64, _I64],
    "evx_fft_plan_destroy": [_P],
    "evx_fft_plan_workspace_bytes": [_P, C.POINTER(C.c_size_t)],
    "evx_fft_plan_spectrum_elems": [_P, C.POINTER(_I64)],
    "evx_rfftn": [_P, _P, _P, _P],
    "evx_irfftn": [_P, _P, _P, _P],
    "evx_ch_imex_step": [_P, _P, _P, _P, C.POINTER(EvxImexParams), _P],
    "evx_ch_imex_step_mirror": [_P, _P, _P, _P, C.POINTER(EvxImexParams), C.POINTER(EvxBC), _I32, _I32, _P],
    "evx_rd_imex_step_mirror": [_P, _P, _P, _P, C.POINTER(EvxImexParams), C.POINTER(EvxBC), _I32, _I32, _P],
    "evx_rd_imex_step": [_P, _P, _P, _P, C.POINTER(EvxImexParams), _P],
    "evx_ch_etd1_step": [_P, _P, _P, _P, C.POINTER(EvxImexParams), _P],
    "evx_rd_etd1_step": [_P, _P, _P, _P, C.POINTER(EvxImexParams), _P],
    "evx_ch_rhs_vjp_workspace_bytes": [_I32, C.POINTER(EvxGrid), C.POINTER(C.c_size_t)],
    "evx_ch_rhs_vjp": [_I32, _P, _P, _P, _P, _P, C.POINTER(EvxGrid), _D, _D, _P, _P, _P],
    "evx_ch_imex_vjp_workspace_bytes": [_P, C.POINTER(C.c_size_t)],
    "evx_ch_imex_step_vjp": [_P, _P, _P, _P, _P, _P, _P, C.POINTER(EvxImexParams), _P],
    "evx_batch_plan_create": [C.POINTER(_P), _I32, _I32, _P, _P],
    "evx_batch_plan_destroy": [_P],
    "evx_batch_state_layout": [_P, _P, C.POINTER(_I64)],
    "evx_batch_workspace_bytes": [_P, C.POINTER(C.c_size_t)],
    "evx_batch_set_params": [_P, _P, _D, _D],
    "evx_batch_set_tiles": [_P, C.c_int],
    "evx_batch_get_tiles": [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "evx_batch_set_fft_group": [_P, C.c_int],
    "evx_batch_get_fft_group": [_P, C.POINTER(C.c_int)],
    "evx_batch_set_one_launch": [_P, C.c_int],
    "evx_batch_get_one_launch": [_P, C.POINTER(C.c_int)],
    "evx_batch_autotune": [_P, _P, _P, C.c_int, _P, C.POINTER(_D)],
    "evx_batch_ch_imex_step": [_P, _P, _P, _P],
    "evx_fft_plan_trace_next_step": [_P, C.POINTER(_P), C.c_int],
    "evx_fft_plan_trace_kinds": [_P, C.POINTER(_I32), C.POINTER(_I32)],
    "evx_fft_plan_set_option": [_P, C.c_int, C.c_int],
    "evx_fft_plan_step_sums": [_P, _P, _P, _P, _P],
    "evx_reduce_workspace_bytes": [_I64, C.POINTER(C.c_size_t)],
    "evx_sum_f64": [_I32, _P, _I64, _I64, _P, _P, _P, _P],
    "evx_sum_f64_noalloc": [_I32, _P, _I64, _I64, _P, _P, _P],
    "evx_ch_energy": [_I32, _P, C.POINTER(EvxGrid), C.POINTER(EvxBC), _D, _D, _P, _P, _P],
}
_RESTYPE = {"evx_last_error": C.c_char_p}

_lib = None


class HipBackendError(RuntimeError):
    """A HIP runtime or kernel-launch failure reported by the library."""


def load():
    """Open the library once and attach prototypes; raise if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH.name} is not built. Run `python -m evoxels_amd.build` "
            "(hipcc --offload-arch=gfx950) before using the 'hip' backend.")
    lib = C.CDLL(os.fspath(LIB_PATH))
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, C.c_int32)
    _lib = lib
    return lib


def last_error() -> str:
    msg = load().evx_last_error()
    return msg.decode() if msg else ""


def check(status: int, what: str = "") -> None:
    """Map a C status to the reference's exception types."""
    if status == EVX_OK:
        return
    msg = f"{what}: {last_error()}" if what else last_error()
    if status == EVX_ERR_INVALID:
        raise ValueError(msg)
    if status == EVX_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    if status == EVX_ERR_NOMEM:
        raise MemoryError(msg)
    raise HipBackendError(msg)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)


def stream_handle(device=None) -> C.c_void_p:
    """hipStream_t of torch's current stream on ``device``."""
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def ptr(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def dtype_code(dtype) -> int:
    if dtype == torch.float32:
        return EVX_F32
    if dtype == torch.float64:
        return EVX_F64
    raise ValueError(f"hip backend supports float32 and float64 fields, got {dtype}")


def make_grid(shape4, div_dx, div_dx2) -> EvxGrid:
    g = EvxGrid()
    g.channels, g.nx, g.ny, g.nz = (int(v) for v in shape4)
    for a in range(3):
        g.div_dx[a] = float(div_dx[a])
        g.div_dx2[a] = float(div_dx2[a])
    return g


BC_NODE_VALUE, BC_NODE_EXTRAP = 3, 4


def make_bc(bc, convention="cell_center") -> EvxBC:
    """bc: normalised ((kind, values), ...) triple.  On a staggered_x grid the
    x rules are the boundary-node ones (Dirichlet: the node holds the value,
    Neumann: 4/3 u1 - 1/3 u2)."""
    b = EvxBC()
    for a, (kind, values) in enumerate(bc):
        if kind not in BC_KIND:
            raise ValueError(f"Unsupported BC type: {kind}")
        b.kind[a] = BC_KIND[kind]
        if kind == "dirichlet":
            b.lo[a] = float(values[0])
            b.hi[a] = float(values[1])
        if convention == "staggered_x" and a == 0 and kind != "periodic":
            b.kind[a] = BC_NODE_VALUE if kind == "dirichlet" else BC_NODE_EXTRAP
    return b


def make_epilogue(mode=OUT_RHS, coef=0.0, base=None, acc_in=None, acc_out=None) -> EvxEpilogue:
    e = EvxEpilogue()
    e.mode = mode
    e.coef = float(coef)
    e.base = base.data_ptr() if base is not None else None
    e.acc_in = acc_in.data_ptr() if acc_in is not None else None
    e.acc_out = acc_out.data_ptr() if acc_out is not None else None
    return e


@contextlib.contextmanager
def gc_paused():
    """No automatic garbage collection inside the block.  Used around HIP
    graph captures: a collected reference cycle may run a finalizer that
    frees device or pinned memory (an FFT plan's tables), and such a call
    invalidates the capture.  torch.cuda.graph itself collects on entry."""
    was_on = gc.isenabled()
    gc.disable()
    try:
        yield
    finally:
        if was_on:
            gc.enable()
