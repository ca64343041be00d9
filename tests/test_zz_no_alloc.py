"""rin_eval_points_noalloc (include/rin.h): the allocation-free variant of
rin_eval_points for callers that cannot allocate.  Through the C ABI, with
device buffers from the HIP runtime the library itself uses: two calls in a
row (the second with a smaller input) give, bit for bit, what
rin_eval_points gives on the same points, and an input beyond the fixed
array's size is an argument error."""
import ctypes

import numpy as np
import pytest

import support as S
import workloads as W

pytestmark = pytest.mark.gpu

HIP = "/opt/rocm/lib/libamdhip64.so.7"


class DeviceBuffers:
    def __init__(self):
        self.hip = ctypes.CDLL(HIP)
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.live = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), max(8, nbytes)) == 0
        self.live.append(p)
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, shape, dtype):
        out = np.zeros(shape, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.live:
            self.hip.hipFree(p)
        self.live = []


def _eval_both(ctx, bufs, pts, types, params):
    import rin

    L = rin.lib()
    L.rin_eval_points.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_uint64,
                                  ctypes.POINTER(rin.FuncDesc), ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]
    L.rin_eval_points_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                                          ctypes.c_uint32, ctypes.c_void_p]
    n, F = len(pts), len(types)
    d = rin.descs(types, params)
    want = np.zeros((n, F))
    ctx.check(L.rin_eval_points(ctx.h, pts.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n, d, F,
                                want.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    dp = bufs.upload(pts)
    dd = bufs.upload(np.frombuffer(bytes(d), dtype=np.uint8))
    dv = bufs.alloc(8 * n * F)
    ctx.check(L.rin_eval_points_noalloc(ctx.h, dp, n, dd, F, dv))
    got = bufs.download(dv, (n, F), np.float64)
    return want, got, (dp, dd, dv)


def test_noalloc_eval_matches_the_allocating_entry_point(ctx):
    bufs = DeviceBuffers()
    try:
        rng = np.random.default_rng(5)
        types, params = W.mixed(rng, 4, 4, 3, 3).arrays()
        for n in (50_000, 7_001):  # the second call smaller: its rows must not see the first's
            pts = np.ascontiguousarray(rng.uniform(-1, 1, (n, 3)))
            want, got, _ = _eval_both(ctx, bufs, pts, types, params)
            np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))
            np.testing.assert_array_equal(want.view(np.uint64), S.evaluate(types, params, pts).view(np.uint64))
        # beyond the fixed array: refused before any launch
        import rin

        L = rin.lib()
        too_many = (2 << 20) // len(types) + 1
        rc = L.rin_eval_points_noalloc(ctx.h, ctypes.c_void_p(16), too_many, ctypes.c_void_p(16), len(types),
                                       ctypes.c_void_p(16))
        assert rc == -1
        assert b"exceed" in L.rin_last_error(ctx.h)
    finally:
        bufs.free()
