"""evx_sum_f64_noalloc (no allocation, static device scratch) against
evx_sum_f64 (caller workspace) through the C ABI: bitwise equal results and
flags on two calls in a row, the second with a smaller input; more channels
than the static scratch holds is refused before any launch."""
import ctypes as C

import numpy as np
import pytest
import torch

from evoxels_amd import _lib

pytestmark = pytest.mark.gpu
MAXC = 64


def sums(lib, x, noalloc):
    c, n = x.shape[0], x[0].numel()
    out = torch.empty(c, dtype=torch.float64, device=x.device)
    flags = torch.zeros(2, dtype=torch.int32, device=x.device)
    dt = _lib.dtype_code(x.dtype)
    s = torch.cuda.current_stream().cuda_stream
    if noalloc:
        st = lib.evx_sum_f64_noalloc(dt, x.data_ptr(), c, n, out.data_ptr(), flags.data_ptr(), s)
    else:
        nb = C.c_size_t()
        assert lib.evx_reduce_workspace_bytes(c, C.byref(nb)) == 0
        ws = torch.empty(nb.value, dtype=torch.uint8, device=x.device)
        st = lib.evx_sum_f64(dt, x.data_ptr(), c, n, out.data_ptr(), flags.data_ptr(), ws.data_ptr(), s)
    assert st == 0, _lib.last_error()
    torch.cuda.synchronize()
    return out.cpu().numpy(), flags.cpu().numpy()


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_noalloc_sum_equals_workspace_sum_twice(dtype):
    lib = _lib.load()
    rng = np.random.default_rng(17)
    big = torch.tensor(rng.standard_normal((MAXC, 37, 41, 43)) * 1e3, dtype=dtype, device="cuda")
    big.view(MAXC, -1)[5, 7] = float("nan")
    small = torch.tensor(rng.standard_normal((3, 5, 7, 9)), dtype=dtype, device="cuda")
    small.view(3, -1)[2, -1] = float("inf")
    for x in (big, small):  # the second call reuses the scratch the first one filled
        a, fa = sums(lib, x, noalloc=True)
        b, fb = sums(lib, x, noalloc=False)
        assert np.array_equal(a, b, equal_nan=True), (a, b)
        assert np.array_equal(fa, fb)
    assert sums(lib, big, True)[1].tolist() == [1, 0]
    assert sums(lib, small, True)[1].tolist() == [0, 1]
    ok = torch.ones((2, 4, 4, 4), dtype=dtype, device="cuda")
    assert sums(lib, ok, True)[0].tolist() == [64.0, 64.0]


def test_noalloc_sum_refuses_more_channels_than_its_scratch():
    lib = _lib.load()
    x = torch.zeros((MAXC + 1, 4), dtype=torch.float32, device="cuda")
    out = torch.empty(MAXC + 1, dtype=torch.float64, device="cuda")
    st = lib.evx_sum_f64_noalloc(0, x.data_ptr(), MAXC + 1, 4, out.data_ptr(), None,
                                 torch.cuda.current_stream().cuda_stream)
    assert st == _lib.EVX_ERR_INVALID
    assert "static scratch" in _lib.last_error()
