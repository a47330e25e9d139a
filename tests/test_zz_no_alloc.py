"""The allocation-free forward pass of the warm-start network
(shx_mlp_forward_noalloc): inputs and outputs in caller-owned device memory,
the hidden activations in the library's file-scope device scratch.  It must
give exactly the outputs of shx_mlp_forward (the same kernels in the same
order), on two calls in a row with the second one smaller (the scratch keeps
the first call's contents), and refuse a batch beyond its scratch limit.
Device memory comes from the HIP runtime libshellhip itself links."""
import ctypes

import numpy as np
import pytest

import shell_amd as S
from shell_amd import surrogate as SG

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _hip():
    S.lib()                                   # libshellhip maps the runtime first
    H = ctypes.CDLL("libamdhip64.so.7")
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipFree.argtypes = [ctypes.c_void_p]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    H.hipDeviceSynchronize.argtypes = []
    return H


def _noalloc(L):
    f = L.shx_mlp_forward_noalloc
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    return f


def test_noalloc_forward_equals_forward():
    H = _hip()
    L = S.lib()
    fwd = _noalloc(L)
    widths = [6, 256, 256, 256, 64]
    net = SG.MLP(widths, seed=3)
    rng = np.random.default_rng(4)
    nmax = 1024
    dx, dy = ctypes.c_void_p(), ctypes.c_void_p()
    assert H.hipMalloc(ctypes.byref(dx), nmax * widths[0] * 4) == 0
    assert H.hipMalloc(ctypes.byref(dy), nmax * widths[-1] * 4) == 0
    try:
        for B in (1024, 300):
            say("case noalloc forward batch", B)
            x = np.ascontiguousarray(rng.normal(size=(B, widths[0])), dtype=np.float32)
            ref = net.forward(x)
            assert H.hipMemcpy(dx, x.ctypes.data, x.nbytes, 1) == 0          # host to device
            rc = fwd(net.h, B, dx, dy)
            assert rc == 0, L.shx_mlp_last_error(net.h)
            y = np.empty((B, widths[-1]), dtype=np.float32)
            assert H.hipMemcpy(y.ctypes.data, dy, y.nbytes, 2) == 0          # device to host
            say("  max |y - y_forward|", float(np.abs(y - ref).max()))
            assert np.array_equal(y, ref)
        # beyond the scratch limit (2048 rows at hidden width 256): refused, nothing launched
        rc = fwd(net.h, 4096, dx, dy)
        assert rc == S.E_ARG
        assert b"scratch limit" in L.shx_mlp_last_error(net.h)
    finally:
        H.hipFree(dx)
        H.hipFree(dy)
        net.close()
