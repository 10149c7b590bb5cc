"""cv_bvh_intersect_nearest_noalloc against cv_bvh_intersect_nearest: the
allocation-free query (fixed per-GPU scratch array) returns the same bits
as the allocating entry point with device pointers, call after call, with a
smaller input and another scene in between."""
import ctypes

import numpy as np
import pytest

from convector_amd import scenes
from helpers import model_scene

pytestmark = pytest.mark.gpu


class _Dev:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(max(nbytes, 16))) == 0
        self.nbytes = nbytes

    def put(self, a):
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p),
                                  ctypes.c_size_t(a.nbytes), 1) == 0

    def get(self, dtype, count):
        out = np.zeros(count, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p,
                                  ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def fill(self, byte):
        assert self.hip.hipMemset(self.p, byte, ctypes.c_size_t(self.nbytes)) == 0

    def free(self):
        self.hip.hipFree(self.p)


def _both(cv, scene, rays, bufs, mode):
    """(allocating entry point's hits/isects, allocation-free ones) for rays."""
    n = len(rays)
    dr, dh, di = bufs
    dr.put(np.ascontiguousarray(rays, np.float32))
    want_h = want_i = got_h = got_i = None
    hp = dh.p.value if mode in ("hits", "both") else None
    ip = di.p.value if mode in ("isects", "both") else None
    for noalloc in (False, True):
        dh.fill(0xAB)
        di.fill(0xCD)
        if noalloc:
            scene.intersect_nearest_noalloc(dr.p.value, n, hp, ip)
        else:
            cv.check(cv.lib().cv_bvh_intersect_nearest(
                scene._h, dr.p, n, ctypes.c_void_p(hp) if hp else None,
                ctypes.c_void_p(ip) if ip else None, cv.PTR_DEVICE))
        h = dh.get(np.uint32, 4 * n) if hp else None
        i = di.get(np.uint32, 16 * n) if ip else None
        if noalloc:
            got_h, got_i = h, i
        else:
            want_h, want_i = h, i
    return want_h, want_i, got_h, got_i


@pytest.mark.parametrize("mode", ["hits", "isects", "both"])
def test_noalloc_equals_allocating_query(cv, oracle, models_dir, mode):
    hip = ctypes.CDLL("libamdhip64.so")
    big = 70000
    bufs = (_Dev(hip, 24 * big), _Dev(hip, 16 * big), _Dev(hip, 64 * big))
    try:
        big_sd = scenes.config1()  # 1,330,746 triangles, BVH depth 36
        deep = cv.Scene.from_scene_data(big_sd)
        depth = deep.bvh_info()["depth"]
        # deeper than the 32 LDS entries hold: the fixed spill area is used
        assert 32 < depth + 4 and depth <= cv._native.CV_NOALLOC_MAX_DEPTH
        sd = model_scene(models_dir, "stanford_dragon")
        dragon = cv.Scene.from_scene_data(sd)
        suz = cv.Scene.from_scene_data(model_scene(models_dir, "suzanne"))
        rng = np.random.default_rng(91)
        d = rng.normal(size=(big, 3))
        rays_a = np.concatenate([rng.uniform(-0.9, 0.9, (big, 3)),
                                 d / np.linalg.norm(d, axis=1, keepdims=True)], 1)
        rays_a = rays_a.astype(np.float32)
        rays_b = rays_a[:5003]  # the second call: a smaller input
        rays_s = scenes.rays_inward(20011, 92, radius=5, target_scale=1.0)
        lo, hi = sd.tris.reshape(-1, 3).min(0), sd.tris.reshape(-1, 3).max(0)
        c, r = (lo + hi) / 2, float((hi - lo).max())
        rays_d = scenes.rays_inward(4099, 93, radius=2 * r, target_scale=0.5 * r)
        rays_d[:, :3] += c
        for scene, rays in ((deep, rays_a), (deep, rays_b), (suz, rays_s), (dragon, rays_d),
                            (deep, rays_b)):
            print("noalloc case: mode %s, %d rays" % (mode, len(rays)), flush=True)
            wh, wi, gh, gi = _both(cv, scene, rays, bufs, mode)
            if wh is not None:
                np.testing.assert_array_equal(gh, wh)
                assert (wh.reshape(-1, 4)[:, 0] != 0xFFFFFFFF).any()
            if wi is not None:
                np.testing.assert_array_equal(gi, wi)
        # and against the oracle on the smaller call
        ref = oracle.Scene.from_scene_data(sd).intersect(rays_d)
        _, _, gh, gi = _both(cv, dragon, rays_d, bufs, "both")
        np.testing.assert_array_equal(gh.reshape(-1, 4)[:, 0], ref["tri"])
        np.testing.assert_array_equal(gi.reshape(-1, 16)[:, 13], ref["numi_aabb"])
    finally:
        for b in bufs:
            b.free()


def test_noalloc_argument_errors(cv, models_dir):
    suz = cv.Scene.from_scene_data(model_scene(models_dir, "suzanne"))
    L = cv.lib()
    assert L.cv_bvh_intersect_nearest_noalloc(suz._h, None, 0, None, None) == 0
    assert L.cv_bvh_intersect_nearest_noalloc(suz._h, None, 5, None, None) == 1
    assert L.cv_bvh_intersect_nearest_noalloc(None, None, 0, None, None) == 1
