"""The allocation-free caustic forward (owe_hip_caustic_forward_noalloc, include/owe_hip_caustic.h)
against the handle's forward (owe_hip_caustic_create + owe_hip_caustic_forward) through the C ABI:
the same kernels over library-held bins, so the same irradiance bit for bit — on two calls in a row,
the second with a smaller grid and fewer particles (the bins keep the first call's values, which the
second must clear) — and grids beyond the array's 512 × 512 refused before any launch."""
from __future__ import annotations

import ctypes

import pytest

from owe_hip import _lib
from owe_hip import caustic as C

pytestmark = pytest.mark.gpu


def _noalloc():
    f = _lib.load().owe_hip_caustic_forward_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]
    return f


def test_noalloc_forward_matches_the_handle_forward(cuda):
    import torch

    f = _noalloc()
    for nodes, pixels, particles, pass_index in ((256, 512, 1 << 20, 3), (64, 96, 1 << 16, 5)):
        print(f"noalloc forward: {nodes}^2 heights, {pixels}^2 pixels, {particles} particles, pass {pass_index}",
              flush=True)
        s = C.caustic_settings(nodes=nodes, pixels=pixels, particles=particles)
        H = torch.tensor(C.initial_heights(nodes), device=cuda)
        want = torch.empty((pixels, pixels), dtype=torch.float32, device=cuda)
        c = C.Caustic(s)
        c.forward_into(H, want, pass_index)
        got = torch.full((pixels, pixels), float("nan"), dtype=torch.float32, device=cuda)
        stream = torch.cuda.current_stream().cuda_stream
        _lib.check(f(0, ctypes.byref(s), H.data_ptr(), got.data_ptr(), pass_index, stream))
        torch.cuda.synchronize()
        c.close()
        assert float(want.sum()) > 0
        assert torch.equal(got, want)


def test_noalloc_forward_refuses_grids_beyond_its_array(cuda):
    import torch

    f = _noalloc()
    s = C.caustic_settings(nodes=64, pixels=513, particles=1024)
    H = torch.zeros((64, 64), dtype=torch.float32, device=cuda)
    out = torch.zeros((513, 513), dtype=torch.float32, device=cuda)
    rc = f(0, ctypes.byref(s), H.data_ptr(), out.data_ptr(), 0, None)
    assert rc == _lib.ERR_INVALID
    assert "512 x 512" in _lib.load().owe_hip_last_error().decode()
