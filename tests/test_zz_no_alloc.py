"""imago_mlp_forward_noalloc against imago_mlp_forward: device inputs and
outputs (allocated by the test through the HIP runtime the library itself
uses), hidden activations in the library's fixed device array; two calls in
a row, the second with a smaller batch, equal to the allocating entry point
bit for bit; a batch beyond the array is refused."""
import ctypes

import numpy as np
import pytest

from imago_amd import learned, lib

pytestmark = pytest.mark.gpu


def _hip():
    h = ctypes.CDLL("libamdhip64.so", mode=ctypes.RTLD_GLOBAL)
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


def test_noalloc_forward_matches_forward(gpu_ctx):
    hip = _hip()
    rng = np.random.default_rng(8)
    m = lib.Mlp(gpu_ctx, learned.WIDTHS, seed=5)
    L = lib.load()
    win, wout = learned.WIDTHS[0], learned.WIDTHS[-1]
    cap = 5000
    xd, yd = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(xd), cap * win * 4) == 0
    assert hip.hipMalloc(ctypes.byref(yd), cap * wout * 4) == 0
    try:
        for batch in (300, 37):
            x = rng.normal(size=(batch, win)).astype(np.float32)
            print("[gpu] noalloc forward batch %d" % batch, flush=True)
            ref = m.forward(x)
            nan = np.full((batch, wout), np.nan, dtype=np.float32)
            assert hip.hipMemcpy(xd, x.ctypes.data, x.nbytes, H2D) == 0
            assert hip.hipMemcpy(yd, nan.ctypes.data, nan.nbytes, H2D) == 0
            rc = L.imago_mlp_forward_noalloc(m._h, batch, xd, yd)
            assert rc == 0, L.imago_last_error()
            assert hip.hipDeviceSynchronize() == 0
            got = np.empty((batch, wout), dtype=np.float32)
            assert hip.hipMemcpy(got.ctypes.data, yd, got.nbytes, D2H) == 0
            assert np.array_equal(got, ref)
        # 2 x 128 x 5000 floats exceed the 4 MiB array: refused before any launch
        rc = L.imago_mlp_forward_noalloc(m._h, cap, xd, yd)
        assert rc == -1 and b"limit" in L.imago_last_error()
    finally:
        hip.hipFree(xd)
        hip.hipFree(yd)
        m.close()
