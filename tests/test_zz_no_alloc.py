"""mpm_surrogate_gradient_noalloc (device rows, fixed library scratch, no
allocation) against mpm_surrogate_gradient (host rows) through the C ABI:
float for float, on two calls in a row, the second smaller, and a size
beyond the fixed scratch rejected before any launch.

The device buffers the caller passes are made here with the HIP runtime the
library itself links (libamdhip64.so.7), so that no second runtime is loaded
into the process."""
import ctypes
import os
import re

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _params(seed):
    from oracle import mlp_oracle as mo
    rng = np.random.default_rng(seed)
    u = mo.unpack(rng.normal(0, 0.3, mo.PARAM_COUNT))
    u["W1"][13:] = 0.0
    u["W4"][:, 7:] = 0.0
    u["b4"][7:] = 0.0
    return mo.pack(u).astype(np.float32)


def _max_rows():
    """MPM_SURROGATE_NOALLOC_MAX_ROWS as declared in include/mpm.h."""
    text = open(os.path.join(ROOT, "include", "mpm.h")).read()
    return int(re.search(r"#define MPM_SURROGATE_NOALLOC_MAX_ROWS (\d+)", text).group(1))


class _Hip:
    """Device buffers through the library's own HIP runtime."""
    H2D, D2H = 1, 2

    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.lib.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), nbytes) == 0
        self.bufs.append(p)
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, self.H2D) == 0
        return p

    def filled(self, n, value):
        return self.upload(np.full(n, value, np.float32))

    def download(self, p, n):
        out = np.empty(n, np.float32)
        assert self.lib.hipMemcpy(out.ctypes.data, p, out.nbytes, self.D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.lib.hipFree(p)
        self.bufs = []


def test_noalloc_gradient_equals_gradient():
    from code_mpm_amd import capi
    from code_mpm_amd.surrogate import Surrogate
    lib = capi.load()
    s = Surrogate(seed=7)
    s.set_params(_params(7))
    hip = _Hip()
    rng = np.random.default_rng(8)
    nparams = int(lib.mpm_surrogate_param_count())
    try:
        for n in (_max_rows(), 1000):
            print(f"noalloc gradient: {n} rows", flush=True)
            x = rng.normal(0, 1, (n, 13)).astype(np.float32)
            t = rng.normal(0, 1, (n, 7)).astype(np.float32)
            g_ref = np.empty(nparams, np.float32)
            l_ref = np.zeros(1, np.float32)
            assert lib.mpm_surrogate_gradient(s.handle, x.ctypes.data, t.ctypes.data, n, g_ref.ctypes.data,
                                              l_ref.ctypes.data) == 0
            xd, td = hip.upload(x), hip.upload(t)
            gd, ld = hip.filled(nparams, np.nan), hip.filled(1, np.nan)
            assert lib.mpm_surrogate_gradient_noalloc(s.handle, xd, td, n, gd, ld) == 0
            g = hip.download(gd, nparams)
            assert np.array_equal(g, g_ref), f"{n} rows: max diff {np.nanmax(np.abs(g - g_ref))}"
            assert hip.download(ld, 1)[0] == l_ref[0]
        # beyond the fixed scratch: rejected on the host, nothing launched
        big = _max_rows() + 1
        assert lib.mpm_surrogate_gradient_noalloc(s.handle, gd, gd, big, gd, None) == -5
    finally:
        hip.free()
        s.close()
