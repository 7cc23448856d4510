"""lh_divergence_noalloc (the allocation-free variant of lh_divergence) through the C ABI.

The caller owns the inputs and the output: device buffers allocated here with hipMalloc of the HIP
runtime the library itself links (libamdhip64.so.7, loaded once per process), filled with a
scene's fields.  Two calls in a row, the second on a smaller grid, must each give the mean, the
L-inf norm and the count of lh_divergence on the same fields bit for bit (the same kernels, block
count and summation order); sizes beyond the limits and null pointers are refused with a message.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import config_file

lh = pytest.importorskip("liquidhairinteration_amd")
pytestmark = pytest.mark.gpu


class Hip:
    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), a.nbytes) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0   # hipMemcpyHostToDevice
        return p

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), nbytes) == 0
        self.bufs.append(p)
        return p

    def download(self, p, shape, dtype):
        a = np.empty(shape, dtype)
        assert self.L.hipMemcpy(a.ctypes.data, p, a.nbytes, 2) == 0   # hipMemcpyDeviceToHost
        return a

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def scene_after_steps(N, steps):
    import json
    cfg = json.load(open(config_file("dam_break_48.json")))
    cfg["gridDimension"] = [N, N, N]
    cfg["Boundary"]["child"]["halfExtent"] = [0.5 - 3.0 / N] * 3
    sc = lh.Scene(cfg)
    m = sc.GetLiquidModel()
    m.Init()
    dt = 0.01
    for _ in range(steps):
        dt, _ = m.RunSteps(1, dt)
    return sc, m


def test_divergence_noalloc_matches_lh_divergence():
    L = lh.load_library()
    hip = Hip()
    try:
        out = hip.alloc(3 * 8)
        for N, steps in ((40, 4), (24, 3)):          # the second call on a smaller grid
            print("case divergence noalloc %d^3" % N, flush=True)
            sc, m = scene_after_steps(N, steps)
            mean, linf, cnt = m.ComputeDivergence()
            ptrs = [hip.upload(m.get_field(f)) for f in ("vx", "vy", "vz", "liquidWx", "liquidWy", "liquidWz",
                                                          "liquidPhi")]
            dims = (C.c_int * 3)(N, N, N)
            dx = float(sc.dx[0])
            rc = L.lh_divergence_noalloc(*ptrs, dims, dx, out, None)
            assert rc == 0, L.lh_last_error(None)
            assert hip.L.hipDeviceSynchronize() == 0
            got = hip.download(out, (3,), np.float64)
            assert cnt > 0 and linf > 0
            assert got[2] == cnt and got[1] == linf and got[0] == mean, (got, (mean, linf, cnt))
            sc.close()
        # refused sizes and pointers: nothing launched, a message for lh_last_error(NULL)
        big = (C.c_int * 3)(4096, 4096, 4096)
        assert L.lh_divergence_noalloc(*ptrs, big, 0.1, out, None) == 6
        assert b"2^31" in L.lh_last_error(None)
        bad = (C.c_int * 3)(0, 8, 8)
        assert L.lh_divergence_noalloc(*ptrs, bad, 0.1, out, None) == 6
        assert L.lh_divergence_noalloc(None, *ptrs[1:], dims, 0.1, out, None) == 6
        assert b"null pointer" in L.lh_last_error(None)
    finally:
        hip.free()
