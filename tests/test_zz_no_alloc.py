"""iftrDrawNoAlloc against iftrDrawLegacy.

The no-allocation draw takes device-resident vertices and indices and writes
the caller's device outputs, with every intermediate buffer in the library's
static scratch array.  Its primitive IDs, depth bits and RGBA8 colour must be
identical to the ordinary draw's on the same inputs, on two calls in a row
(the second with fewer triangles, so the array holds the first call's data),
and a draw too large for the array must be refused before any launch.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _DeviceArrays:
    """Device buffers through the HIP runtime the library itself loaded (hipMalloc, hipMemcpy,
    hipFree by ctypes), so that the test needs nothing else on the GPU."""

    def __init__(self):
        import ifrit_hip

        ifrit_hip.lib()  # loads libamdhip64 with the library
        self.H = C.CDLL("libamdhip64.so.7")
        self.H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.H.hipFree.argtypes = [C.c_void_p]
        self.ptrs = []

    def put(self, arr):
        a = np.ascontiguousarray(arr)
        p = C.c_void_p()
        assert self.H.hipMalloc(C.byref(p), max(a.nbytes, 4)) == 0
        self.ptrs.append(p)
        assert self.H.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.H.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.ptrs:
            self.H.hipFree(p)
        self.ptrs = []


def _normal_draw(r, n_idx):
    r.L.iftrDrawLegacy(r.h, n_idx, 1)
    r._check()
    assert r.stats()["status"] == 0
    ids, depth = r.visibility()
    return ids, depth, r.color()


def _noalloc_draw(r, scene, n_idx, shade=True):
    import ifrit_hip

    W, H, S = scene.width, scene.height, scene.samples
    dev = _DeviceArrays()
    try:
        v = dev.put(scene.vertices.astype(np.float32))
        i = dev.put(scene.indices.astype(np.int32))
        ids0 = np.full((W * H * S,), 0x5A5A5A5A, np.uint32)
        d0 = np.full((W * H * S,), -7.0, np.float32)
        c0 = np.full((H, W, 4), 17, np.uint8)
        ids, depth = dev.put(ids0), dev.put(d0)
        rgba = dev.put(c0) if shade else None
        d = ifrit_hip.NoAllocDraw(v, scene.vertices.shape[1], scene.vertices.shape[0], i, n_idx, ids, depth, rgba)
        print(f"iftrDrawNoAlloc {W}x{H}x{S} indices={n_idx}", flush=True)
        r.L.iftrDrawNoAlloc(r.h, C.byref(d))
        r._check()
        return dev.get(ids, ids0), dev.get(depth, d0), (dev.get(rgba, c0) if shade else None)
    finally:
        dev.free()


@pytest.mark.parametrize("samples", [1, 4])
def test_noalloc_matches_draw_legacy_twice(samples):
    from ifrit_hip import Renderer, scenes

    s = scenes.config1(320, 240)
    s.samples = samples
    r = Renderer()
    try:
        r.load_scene(s)
        full = s.indices.size
        small = (full // 3 // 2) * 3  # half the triangles
        for n in (full, small):
            ids0, d0, c0 = _normal_draw(r, n)
            ids1, d1, c1 = _noalloc_draw(r, s, n)
            assert np.array_equal(ids0, ids1), f"{n} indices: primitive ids differ"
            assert np.array_equal(d0.view(np.uint32), d1.view(np.uint32)), f"{n} indices: depth differs"
            assert np.array_equal(c0, c1), f"{n} indices: colour differs"
            assert (ids0 != 0xFFFFFFFF).sum() > 1000
        # the ordinary draw still works after the no-allocation draws
        ids2, d2, c2 = _normal_draw(r, small)
        assert np.array_equal(ids2, ids1) and np.array_equal(c2, c1)
    finally:
        r.close()


def test_noalloc_refuses_draws_beyond_the_scratch_array():
    import ifrit_hip
    from ifrit_hip import IfritError, Renderer, scenes

    s = scenes.config1(64, 48)
    r = Renderer()
    dev = _DeviceArrays()
    try:
        r.load_scene(s)
        r.set_target(4096, 4096, 8)  # 134M samples of record slots: far beyond the array
        v = dev.put(s.vertices.astype(np.float32))
        i = dev.put(s.indices.astype(np.int32))
        depth = dev.put(np.zeros(16, np.float32))  # never written: the draw is refused first
        d = ifrit_hip.NoAllocDraw(v, s.vertices.shape[1], s.vertices.shape[0], i, s.indices.size, None, depth, None)
        r.L.iftrDrawNoAlloc(r.h, C.byref(d))
        with pytest.raises(IfritError, match="bytes of scratch"):
            r._check()
    finally:
        dev.free()
        r.close()
