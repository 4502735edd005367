"""octlb_build_links_noalloc: the link build on device buffers with its
scratch in a static device array (no allocation in the call), checked
against octlb_build_links (bit-equal links and q) on two calls in a row, the
second on a smaller forest and soup, and its size limit refused before any
device work."""
import ctypes as C

import numpy as np
import pytest

from octlb_amd import abi, cases, geometry

pytestmark = pytest.mark.gpu


def _inputs(big):
    if big:
        case = cases.cylinder_channel(bricks_x=16, bricks_yz=4, base_level=0, fine=True, xc=1.0, radius=0.125)
        xc, r, segs, rings = 1.0, 0.125, 512, 8
    else:
        case = cases.cylinder_channel(bricks_x=2, base_level=1, fine=True, xc=0.5, radius=0.1,
                                      refine_lo=0.25, refine_hi=1.0)
        xc, r, segs, rings = 0.5, 0.1, 96, 3
    topo = case.topo
    mat = geometry.material_field(topo, lambda x, y, z: (x - xc) ** 2 + (y - 0.5) ** 2 < r * r)
    tris = geometry.cylinder_soup(xc, 0.5, r, 0.0, 1.0, segs, rings)
    return np.ascontiguousarray(topo.bounds, np.float64), np.ascontiguousarray(mat, np.uint8), \
        np.ascontiguousarray(tris, np.float64).reshape(-1, 9)


class _Hip:
    """Device buffers through the HIP runtime that liboctlb_gpu.so itself
    links (already loaded with it, so the same runtime instance)."""

    def __init__(self):
        abi.lib()
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.live = []

    def _ok(self, rc, what):
        assert rc == 0, f"{what} failed with {rc}"

    def alloc(self, nbytes):
        p = C.c_void_p()
        self._ok(self.rt.hipMalloc(C.byref(p), max(int(nbytes), 8)), "hipMalloc")
        self.live.append(p)
        return p

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        self._ok(self.rt.hipMemcpy(p, arr.ctypes.data_as(C.c_void_p), arr.nbytes, 1), "hipMemcpy H2D")
        return p

    def download(self, p, dtype, count):
        out = np.empty(count, dtype)
        if count:
            self._ok(self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def free(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live.clear()


def _noalloc(hip, bounds, mat, tris, capacity):
    cap = max(capacity, 1)
    d_b = hip.upload(bounds.reshape(-1))
    d_m = hip.upload(mat.reshape(-1))
    d_t = hip.upload(tris.reshape(-1))
    d_l = hip.upload(np.full(cap * 3, -1, np.int32))
    d_q = hip.upload(np.full(cap, -1.0, np.float64))
    cnt = C.c_int64(-1)
    nb = bounds.shape[0]
    rc = abi.lib().octlb_build_links_noalloc(nb, 8, d_b, d_m, tris.shape[0], d_t, capacity, C.byref(cnt),
                                             d_l, d_q, None)
    abi.check(rc, "octlb_build_links_noalloc")
    L = cnt.value
    return hip.download(d_l, np.int32, 3 * L).reshape(-1, 3), hip.download(d_q, np.float64, L)


def test_noalloc_link_build_matches_the_allocating_one():
    hip = _Hip()
    try:
        for big in (True, False):  # the second call is smaller: stale scratch must not leak into it
            bounds, mat, tris = _inputs(big)
            ref_l, ref_q = geometry.build_links(bounds, 8, mat, tris)
            assert len(ref_q) > 0
            got_l, got_q = _noalloc(hip, bounds, mat, tris, len(ref_q))
            print(f"{'big' if big else 'small'}: {len(ref_q)} links", flush=True)
            np.testing.assert_array_equal(got_l, ref_l)
            np.testing.assert_array_equal(got_q, ref_q)
    finally:
        hip.free()


def test_noalloc_link_build_refuses_sizes_beyond_its_scratch():
    hip = _Hip()
    try:
        p = hip.alloc(128)
        cnt = C.c_int64(-1)
        rc = abi.lib().octlb_build_links_noalloc(20000, 8, p, p, 1, p, 0, C.byref(cnt), None, None, None)
        assert rc != 0
        assert b"16384" in abi.lib().octlb_last_error()
        assert cnt.value == 0
    finally:
        hip.free()
