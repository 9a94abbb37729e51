"""gs_sort_pairs_static, the allocation-free variant of gs_sort, against
gs_sort itself: the same visible (key, id) list, handed over in device
buffers the test owns, must come back in gs_sort's order bit for bit, on two
calls in a row (the second with fewer pairs and many equal keys), and a count
above the static scratch's maximum is refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from gamesplat_amd import abi, scenes

pytestmark = pytest.mark.gpu


class DeviceBuffers:
    """Four u32 device arrays owned by the test (the entry point allocates nothing)."""

    def __init__(self, n):
        self.hip = C.CDLL("libamdhip64.so")
        self.n = n
        self.ptrs = [C.c_void_p() for _ in range(4)]
        for p in self.ptrs:
            assert self.hip.hipMalloc(C.byref(p), C.c_size_t(max(4, 4 * n))) == 0

    def put(self, k, a):
        a = np.ascontiguousarray(a, np.uint32)
        assert a.size <= self.n
        assert self.hip.hipMemcpy(self.ptrs[k], a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0

    def get(self, k, count):
        out = np.empty(count, np.uint32)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptrs[k], C.c_size_t(out.nbytes), 2) == 0
        return out

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)


def _reference_sort(ctx, g, cam):
    """The visible list and its order from gs_cull + gs_sort."""
    ctx.upload(g)
    ctx.cull(cam)
    ctx.sort()
    vis = ctx.stats()["visible_count"]
    keys = ctx.read_buffer(abi.GS_BUFFER_VISIBLE_KEYS)
    ids = ctx.read_buffer(abi.GS_BUFFER_VISIBLE_IDS)
    sk = ctx.read_buffer(abi.GS_BUFFER_SORTED_KEYS)[:vis]
    si = ctx.read_buffer(abi.GS_BUFFER_SORTED_IDS)[:vis]
    return keys, ids, sk, si


def test_static_sort_equals_gs_sort(gs, c1_scene):
    g, cfg, cam = c1_scene
    small = gs.generate_scene(7, 30000, 4, 0.0)
    small[:, 0:3] = np.round(small[:, 0:3] * 2.0) / 2.0  # coarse positions: many equal depth keys
    cam2 = scenes.random_view_camera(11, 640, 480)
    with gs.Context(0) as ctx:
        lib = ctx._lib
        cases = [(g, cam), (small, cam2)]
        ref = [_reference_sort(ctx, gg, cc) for gg, cc in cases]
        n = max(r[0].size for r in ref)
        buf = DeviceBuffers(n)
        try:
            sizes = []
            for k, (keys, ids, sk, si) in enumerate(ref):
                v = keys.size
                sizes.append(v)
                print(f"case: static sort call {k + 1}, {v} pairs", flush=True)
                buf.put(0, keys)
                buf.put(1, ids)
                abi.check(lib.gs_sort_pairs_static(ctx.handle, buf.ptrs[0], buf.ptrs[1], buf.ptrs[2], buf.ptrs[3], v))
                ctx.synchronize()
                np.testing.assert_array_equal(buf.get(2, v), sk)
                np.testing.assert_array_equal(buf.get(3, v), si)
                np.testing.assert_array_equal(buf.get(0, v), keys)  # inputs untouched
                np.testing.assert_array_equal(buf.get(1, v), ids)
            assert sizes[1] < sizes[0]
            assert np.unique(ref[1][0]).size < ref[1][0].size  # ties exercised
            over = abi.GS_SORT_STATIC_MAX_PAIRS + 1
            code = lib.gs_sort_pairs_static(ctx.handle, buf.ptrs[0], buf.ptrs[1], buf.ptrs[2], buf.ptrs[3], over)
            assert code == abi.GS_ERR_LIMIT
            assert b"exceed" in lib.gs_last_error()
        finally:
            buf.close()
