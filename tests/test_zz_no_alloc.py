"""The allocation-free entry point eqd_mlp_forward_noalloc (device arrays in
and out, the network's operands in the library's per-device scratch array)
against eqd_mlp_forward, which allocates its own buffers: the same kernels,
so the outputs must be identical.  Two calls in a row, the second with fewer
rows and other parameters, so a scratch array holding the first call's
operands would show.

The test's own device arrays and stream come from the HIP runtime libeqd.so
is linked against (libamdhip64.so.7, through ctypes), so that no second HIP
runtime shares the process."""
import ctypes as C

import numpy as np
import pytest

from eqdyna_amd import solver
from test_gpu_mlp import _data, _params

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class Hip:
    def __init__(self):
        solver.lib()  # libeqd.so first: its HIP runtime is the one to share
        L = C.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        L.hipStreamSynchronize.argtypes = [C.c_void_p]
        L.hipStreamDestroy.argtypes = [C.c_void_p]
        L.hipSetDevice.argtypes = [C.c_int]
        self.L = L
        self.ptrs = []
        assert L.hipSetDevice(0) == 0

    def upload(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a, dtype=np.float32)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 4)) == 0
        self.ptrs.append(p.value)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p.value

    def download(self, ptr: int, n: int) -> np.ndarray:
        out = np.empty(n, dtype=np.float32)
        assert self.L.hipMemcpy(out.ctypes.data, C.c_void_p(ptr), out.nbytes, D2H) == 0
        return out

    def close(self):
        for p in self.ptrs:
            self.L.hipFree(C.c_void_p(p))
        self.ptrs = []


@pytest.mark.gpu
def test_noalloc_forward_matches_allocating_forward():
    hip = Hip()
    st = C.c_void_p()
    assert hip.L.hipStreamCreate(C.byref(st)) == 0
    try:
        for n, seed in ((5000, 1), (777, 2)):
            X, y = _data(n, seed)
            p = _params(X, y, seed + 10)
            ref = solver.mlp_forward(p, X)
            dp, dX = hip.upload(p), hip.upload(X)
            dout = hip.upload(np.full(n, np.nan, dtype=np.float32))
            solver.mlp_forward_noalloc(dp, dX, n, dout, 0, st.value)
            assert hip.L.hipStreamSynchronize(st) == 0
            got = hip.download(dout, n)
            print(f"n={n}: max |noalloc - forward| = {np.max(np.abs(got - ref)):.3e}", flush=True)
            assert np.array_equal(got, ref)
    finally:
        hip.L.hipStreamDestroy(st)
        hip.close()


@pytest.mark.gpu
def test_noalloc_forward_rejects_bad_sizes():
    hip = Hip()
    try:
        dp = hip.upload(np.zeros(solver.MLP_NPARAMS, dtype=np.float32))
        with pytest.raises(solver.EqdError):
            solver.mlp_forward_noalloc(dp, dp, -1, dp, 0, 0)
        with pytest.raises(solver.EqdError):
            solver.mlp_forward_noalloc(dp, dp, 64 * (2 ** 32), dp, 0, 0)
    finally:
        hip.close()
