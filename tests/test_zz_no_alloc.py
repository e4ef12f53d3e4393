"""lcr_depth_to_cloud_noalloc (through the C ABI) against lcr_depth_to_cloud on the
same depth maps: identical points, normals and colours, bitwise, on two calls in a
row (the second one smaller, so it runs over the first call's leftover scratch).
Device buffers come from the HIP runtime liblcr itself uses (hipMalloc through
ctypes); the entry point under test allocates nothing."""
import ctypes as C

import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu

INTR = np.array([525.0, 525.0, 319.5, 239.5], np.float64)


@pytest.fixture(scope="module")
def hip(lcr):
    # liblcr already loaded libamdhip64.so.7: dlopen by that name returns the same
    # runtime, so its device pointers are liblcr's
    L = C.CDLL("libamdhip64.so.7")
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.hipDeviceSynchronize.argtypes = []
    return L


H2D, D2H = 1, 2


class Dev:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(int(nbytes), 1)) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, dtype, count):
        out = np.empty(count, dtype)
        if out.nbytes:
            assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def depth_map(w, h, seed):
    """A synthetic 16-bit millimetre depth map: a tilted plane with a bump, a few
    holes, and depth steps larger than the 20 mm neighbour rule."""
    rng = np.random.default_rng(seed)
    jj, ii = np.meshgrid(np.arange(w), np.arange(h))
    z = 800 + 0.3 * jj + 0.2 * ii + 40 * np.exp(-((jj - w / 2) ** 2 + (ii - h / 2) ** 2) / (0.05 * w * h))
    z = z + rng.normal(scale=1.5, size=z.shape)
    z[rng.random(z.shape) < 0.03] = 0
    z[:, w // 3] += 60
    d = np.clip(np.rint(z), 0, 65535).astype(np.uint16)
    color = rng.integers(0, 256, size=(h, w, 3), dtype=np.uint8)
    return d, color


def test_noalloc_matches_depth_to_cloud_on_two_calls(lcr, hip):
    lcr.init(0)
    cap_w, cap_h = 640, 480
    cap = cap_w * cap_h
    bufs = {k: Dev(hip, n) for k, n in (("depth", 2 * cap), ("color", 3 * cap),
                                        ("xyz", 12 * cap), ("nrm", 12 * cap), ("rgb", 3 * cap),
                                        ("n", 8))}
    try:
        for call, (w, h, crop) in enumerate(((640, 480, None), (320, 240, (10, 200, 5, 300)))):
            d, col = depth_map(w, h, call)
            print_case(test="noalloc", call=call, w=w, h=h)
            ref = lcr.depth_to_cloud(d, INTR, 500, 1500, crop=crop, color=col)
            bufs["depth"].put(d)
            bufs["color"].put(col)
            cp = (C.c_int32 * 4)(*crop) if crop is not None else None
            rc = lcr.lib.lcr_depth_to_cloud_noalloc(
                bufs["depth"].p, w, h, bufs["color"].p, INTR.ctypes.data_as(C.POINTER(C.c_double)),
                500, 1500, cp, bufs["xyz"].p, bufs["nrm"].p, bufs["rgb"].p, bufs["n"].p, None)
            assert rc == 0, lcr.lib.lcr_last_error()
            assert hip.hipDeviceSynchronize() == 0
            n = int(bufs["n"].get(np.int64, 1)[0])
            xyz, nrm, rgb = ref[0], ref[1], ref[2]
            assert n == len(xyz) and n > 0
            np.testing.assert_array_equal(bufs["xyz"].get(np.float32, 3 * n).reshape(-1, 3), xyz)
            np.testing.assert_array_equal(bufs["nrm"].get(np.float32, 3 * n).reshape(-1, 3), nrm)
            np.testing.assert_array_equal(bufs["rgb"].get(np.uint8, 3 * n).reshape(-1, 3), rgb)
        # beyond the fixed scratch: rejected on the host, nothing launched
        rc = lcr.lib.lcr_depth_to_cloud_noalloc(
            bufs["depth"].p, 2048, 1024, None, INTR.ctypes.data_as(C.POINTER(C.c_double)), 500,
            1500, None, bufs["xyz"].p, None, None, bufs["n"].p, None)
        assert rc == -5 and b"LCR_NOALLOC_MAX_PIXELS" in lcr.lib.lcr_last_error()
    finally:
        for b in bufs.values():
            b.free()
