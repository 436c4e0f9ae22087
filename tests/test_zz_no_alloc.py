"""pepsmg_vv_dot_device_noalloc: the allocation-free variant of
pepsmg_vv_dot (partial sums in a static device array), through the C-ABI,
against the existing entry point on two calls in a row, the second with a
smaller grid (the shared scratch then holds the first call's longer list of
partial sums, of which only the new count may be read)."""
import ctypes as C

import numpy as np
import pytest

from peps_mg_amd import _lib

pytestmark = pytest.mark.gpu


def padded(rng, nx, ny, nz):
    a = np.zeros((nz + 2, ny + 2, nx + 2))
    a[1:-1, 1:-1, 1:-1] = rng.uniform(-1, 1, (nz, ny, nx))
    return a


def test_noalloc_dot_matches_vv_dot_bitwise():
    import torch
    rng = np.random.default_rng(17)
    L = _lib.lib()
    out = torch.zeros(2, dtype=torch.float64, device="cuda")
    for q, (nx, ny, nz) in enumerate([(200, 130, 90), (33, 17, 9)]):   # 2.3e6 cells (1024 blocks), then 5049
        x, y = padded(rng, nx, ny, nz), padded(rng, nx, ny, nz)
        ref = C.c_double(0)
        _lib.check(L.pepsmg_vv_dot(nx, ny, nz, _lib.dptr(x), _lib.dptr(y), C.byref(ref)))
        dx, dy = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
        stream = torch.cuda.current_stream().cuda_stream
        _lib.check(L.pepsmg_vv_dot_device_noalloc(nx, ny, nz, C.c_void_p(dx.data_ptr()), C.c_void_p(dy.data_ptr()),
                                                  C.c_void_p(out.data_ptr() + 8 * q), C.c_void_p(stream)))
        torch.cuda.synchronize()
        got = out[q].item()
        assert got == ref.value, (q, got, ref.value)
        assert abs(got - float(np.sum(x * y))) <= 1e-12 * float(np.sum(np.abs(x * y)))


def test_noalloc_dot_rejects_bad_sizes():
    import torch
    L = _lib.lib()
    d = torch.zeros(27, dtype=torch.float64, device="cuda")
    p = C.c_void_p(d.data_ptr())
    for dims in [(0, 1, 1), (1, -1, 1), (1, 65536, 65536)]:
        rc = L.pepsmg_vv_dot_device_noalloc(*dims, p, p, p, None)
        assert rc == _lib.ERR_ARG, dims
    assert L.pepsmg_vv_dot_device_noalloc(1, 1, 1, None, p, p, None) == _lib.ERR_ARG
