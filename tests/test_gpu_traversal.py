"""GPU nearest-hit queries (k_intersect) against the oracle's single-ray
traversal: hit ids, t/u/v, every MIntersection field and the debug
counters bit-identical; ids against FP64 brute force within the tolerance
the reference path allows; special rays (axis-aligned, origins on box
planes) that exercise x86 NaN and signed-zero rules."""
import ctypes

import numpy as np
import pytest

from convector_amd import scenes
from helpers import bits, indoor_scene, model_scene

pytestmark = pytest.mark.gpu


def _check_exact(hits, isects, ref):
    assert (hits["tri_id"] == ref["tri"]).all()
    assert (bits(hits["t"]) == bits(ref["tuv"][:, 0])).all()
    assert (bits(hits["u"]) == bits(ref["tuv"][:, 1])).all()
    assert (bits(hits["v"]) == bits(ref["tuv"][:, 2])).all()
    if isects is not None:
        o = ref["isect"]
        assert (bits(isects["position"]) == bits(o[:, 0:3])).all()
        assert (bits(isects["normal"]) == bits(o[:, 3:6])).all()
        assert (bits(isects["distance"]) == bits(o[:, 6])).all()
        assert (isects["material"] == bits(o[:, 7])).all()
        assert (bits(isects["tex_coords"]) == bits(o[:, 8:10])).all()
        assert (isects["numi_aabb"] == ref["numi_aabb"]).all()
        assert (isects["numi_tri"] == ref["numi_tri"]).all()


def _ray_families(center, extent, seed):
    c = np.asarray(center, np.float64)
    fams = [
        scenes.rays_inward(4096, seed, radius=3 * extent, target_scale=0.5 * extent, center=c),
        scenes.rays_inward(4096, seed + 1, radius=3 * extent, coherent=True,
                           target_scale=0.5 * extent, center=c),
        scenes.rays_inward(2048, seed + 2, radius=0.3 * extent, target_scale=0.9 * extent,
                           center=c),
    ]
    # axis-aligned rays: zero direction components give infinite reciprocals
    # and 0 * inf in the slab test (x86 default NaN rules)
    rng = np.random.default_rng(seed)
    o = c + rng.uniform(-1.5, 1.5, (3000, 3)) * extent
    d = np.zeros((3000, 3))
    axis = rng.integers(0, 3, 3000)
    d[np.arange(3000), axis] = rng.choice([-1.0, 1.0], 3000)
    fams.append(np.concatenate([o, d], 1).astype(np.float32))
    return np.concatenate(fams).astype(np.float32)


@pytest.mark.parametrize("name", ["suzanne", "stanford_bunny", "indoor"])
def test_models_exact(cv, oracle, models_dir, name):
    sd = model_scene(models_dir, name) if name != "indoor" else indoor_scene(models_dir)
    print("traversal case", name, sd.ntris, flush=True)
    osc = oracle.Scene.from_scene_data(sd)
    gsc = cv.Scene.from_scene_data(sd)
    lo, hi = sd.tris.reshape(-1, 3).min(0), sd.tris.reshape(-1, 3).max(0)
    rays = _ray_families((lo + hi) / 2, float((hi - lo).max()), 3)
    hits, isects = gsc.intersect_nearest(rays, full=True)
    ref = osc.intersect(rays, isect=True)
    _check_exact(hits, isects, ref)
    assert (hits["tri_id"] != cv.NO_HIT).mean() > 0.2


def test_axis_rays_on_box_planes(cv, oracle):
    """Rays starting exactly on node-box planes with zero direction
    components: -0.0 / +0.0 and NaN choices of minps/maxps decide the pushes."""
    sd = scenes.config0(subdivisions=3)
    osc = oracle.Scene.from_scene_data(sd)
    gsc = cv.Scene.from_scene_data(sd)
    nodes, _ = osc.bvh()
    rng = np.random.default_rng(12)
    pick = rng.integers(0, len(nodes), 4000)
    o = np.where(rng.random((4000, 3)) < 0.5, nodes["min"][pick], nodes["max"][pick])
    d = np.zeros((4000, 3), np.float32)
    ax = rng.integers(0, 3, 4000)
    d[np.arange(4000), ax] = rng.choice([-1.0, 1.0], 4000)
    rays = np.concatenate([o, d], 1).astype(np.float32)
    hits, isects = gsc.intersect_nearest(rays, full=True)
    _check_exact(hits, isects, osc.intersect(rays, isect=True))


def _edge_distance64(tris, tri_id, rays):
    """FP64 world-space distance from each ray's hit point on triangle
    tri_id to the nearest edge of that triangle."""
    v = tris[tri_id].astype(np.float64).reshape(-1, 3, 3)
    o, d = rays[:, :3].astype(np.float64), rays[:, 3:].astype(np.float64)
    a, b = v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]
    n = np.cross(a, b)
    t = ((v[:, 0] - o) * n).sum(1) / (d * n).sum(1)
    p = o + t[:, None] * d
    dist = np.full(len(p), np.inf)
    for i, j in ((0, 1), (1, 2), (2, 0)):
        e = v[:, j] - v[:, i]
        q = p - v[:, i]
        dist = np.minimum(dist, np.linalg.norm(np.cross(e, q), axis=1) / np.linalg.norm(e, axis=1))
    return dist


def test_config0_camera_and_bounce_rays(cv, oracle):
    sd = scenes.config0()
    osc = oracle.Scene.from_scene_data(sd)
    gsc = cv.Scene.from_scene_data(sd)
    rays = np.concatenate([scenes.rays_inward(20000, 21, radius=0.9, target_scale=0.45),
                           scenes.rays_inward(20000, 22, radius=0.95, target_scale=0.95)])
    hits, isects = gsc.intersect_nearest(rays, full=True)
    ref = osc.intersect(rays, isect=True)
    _check_exact(hits, isects, ref)
    # Against an FP64 brute-force re-intersection of every triangle, on
    # 20,000 rays: hit ids equal for >= 99.99 % (at most 2 differ), every
    # difference within 1e-6 of an edge of one of the two triangles.
    n = 20000
    sub = np.concatenate([rays[:n // 2], rays[20000:20000 + n // 2]])
    gid = np.concatenate([hits["tri_id"][:n // 2], hits["tri_id"][20000:20000 + n // 2]])
    b = osc.intersect_brute_f64(sub)
    mism = np.nonzero(gid != b["tri"])[0]
    assert len(mism) <= n // 10000
    for k in mism:
        near = []
        for tid in (gid[k], b["tri"][k]):
            if tid != cv.NO_HIT:
                near.append(float(_edge_distance64(sd.tris, np.array([tid]), sub[k:k + 1])[0]))
        assert near and min(near) <= 1e-6, (k, near)
    both = (gid == b["tri"]) & (gid != cv.NO_HIT)
    g = {k: np.concatenate([hits[k][:n // 2], hits[k][20000:20000 + n // 2]])[both]
         for k in ("t", "u", "v")}
    t64, u64, v64 = b["tuv"][both, 0], b["tuv"][both, 1], b["tuv"][both, 2]
    # t: within 1e-5 relative
    assert (np.abs(g["t"] - t64) <= 1e-5 * np.abs(t64)).all()
    # u, v (triangle.rs:100-112: u = (d x (v0 - o)) . e2 / (d . n), v with e1):
    # within 1e-5 relative to the magnitude of the terms that formula sums,
    # |d| |v0 - o| |e| / |d . n| -- the scale of the FP32 formula's own
    # forward error (on the sphere's 0.02-wide triangles the quotient and the
    # cross product cancel: plain relative differences reach 5e-5 there)
    tri = sd.tris[gid[both]].astype(np.float64).reshape(-1, 3, 3)
    o, d = sub[both, :3].astype(np.float64), sub[both, 3:].astype(np.float64)
    e1, e2 = tri[:, 0] - tri[:, 2], tri[:, 1] - tri[:, 0]
    dn = np.abs((d * np.cross(e1, e2)).sum(1))
    cmag = np.linalg.norm(d, axis=1) * np.linalg.norm(tri[:, 0] - o, axis=1)
    su = cmag * np.linalg.norm(e2, axis=1) / dn
    sv = cmag * np.linalg.norm(e1, axis=1) / dn
    assert (np.abs(g["u"] - u64) <= 1e-5 * su).all()
    assert (np.abs(g["v"] - v64) <= 1e-5 * sv).all()


def test_config1_sample_exact(cv, oracle):
    sd = scenes.config1(subdivisions=7, n_free=5000)
    osc = oracle.Scene.from_scene_data(sd)
    gsc = cv.Scene.from_scene_data(sd)
    rays = np.concatenate([scenes.rays_inward(30000, 31, radius=0.95, target_scale=0.45),
                           scenes.rays_inward(30000, 32, radius=0.9, target_scale=0.95)])
    hits, isects = gsc.intersect_nearest(rays, full=True)
    _check_exact(hits, isects, osc.intersect(rays, isect=True))


def test_imported_bvh(cv, oracle, models_dir):
    """cv_scene_create_with_bvh: traversal of a caller-provided tree."""
    sd = model_scene(models_dir, "suzanne")
    osc = oracle.Scene.from_scene_data(sd)
    nodes, order = osc.bvh()
    gsc = cv.Scene.from_scene_data(sd, with_bvh=(nodes, order))
    rays = scenes.rays_inward(5000, 41, radius=5, target_scale=1.0)
    hits, isects = gsc.intersect_nearest(rays, full=True)
    _check_exact(hits, isects, osc.intersect(rays, isect=True))


def test_sizes_and_device_pointers(cv, oracle, models_dir):
    sd = model_scene(models_dir, "suzanne")
    osc = oracle.Scene.from_scene_data(sd)
    gsc = cv.Scene.from_scene_data(sd)
    for n in (0, 1, 63, 65, 1000):
        rays = scenes.rays_inward(max(n, 1), 50 + n, radius=5, target_scale=1.0)[:n]
        hits = gsc.intersect_nearest(rays)
        if n:
            ref = osc.intersect(rays)
            assert (hits["tri_id"] == ref["tri"]).all()
    # CV_PTR_DEVICE with hipMalloc'd buffers
    hip = ctypes.CDLL("libamdhip64.so")
    n = 4096
    rays = scenes.rays_inward(n, 60, radius=5, target_scale=1.0)
    dr, dh = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dr), ctypes.c_size_t(rays.nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(dh), ctypes.c_size_t(16 * n)) == 0
    try:
        assert hip.hipMemcpy(dr, rays.ctypes.data_as(ctypes.c_void_p),
                             ctypes.c_size_t(rays.nbytes), 1) == 0
        gsc.intersect_nearest_device(dr.value, n, dh.value)
        out = np.zeros(n, cv.HIT_DTYPE)
        assert hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), dh, ctypes.c_size_t(16 * n),
                             2) == 0
        assert (out["tri_id"] == osc.intersect(rays)["tri"]).all()
    finally:
        hip.hipFree(dr)
        hip.hipFree(dh)


def test_stack_sizes_exact(cv, oracle, models_dir):
    """Every per-lane LDS stack size gives the same hits: deep trees spill
    the stack to global memory at 8 entries, never at 32."""
    sd = model_scene(models_dir, "stanford_dragon")
    osc = oracle.Scene.from_scene_data(sd)
    gsc = cv.Scene.from_scene_data(sd)
    assert gsc.bvh_info()["depth"] > 8
    lo, hi = sd.tris.reshape(-1, 3).min(0), sd.tris.reshape(-1, 3).max(0)
    rays = _ray_families((lo + hi) / 2, float((hi - lo).max()), 70)
    ref = osc.intersect(rays, isect=True)
    for stack in (8, 16, 24, 32):
        print("stack case", stack, sd.ntris, flush=True)
        gsc.set_tuning(stack_entries=stack)
        assert gsc.tuning()["stack_entries"] == stack
        hits, isects = gsc.intersect_nearest(rays, full=True)
        _check_exact(hits, isects, ref)


def test_tuning_validation(cv, models_dir):
    gsc = cv.Scene.from_scene_data(model_scene(models_dir, "suzanne"))
    before = gsc.tuning()
    for bad in (dict(stack_entries=12), dict(extend_waves_per_cu=65),
                dict(refill_threshold=66), dict(fetch_batches=65), dict(block_waves=3),
                dict(frames_per_step=65), dict(refill_interval=65), dict(leaf_batch=66)):
        with pytest.raises(cv.CvError) as e:
            gsc.set_tuning(**bad)
        assert e.value.name == "CV_ERR_INVALID_ARGUMENT"
    assert gsc.tuning() == before


def test_hits_only_refill_path_exact(cv, oracle, models_dir):
    """Calls without cv_isect output run the lane-refill traversal (rays as
    a one-bounce extend queue, hits mapped to source indices): the same
    hit bits as the oracle for every knob that traversal reads, on a tree
    deep enough to spill the stack at 8 entries."""
    sd = model_scene(models_dir, "stanford_dragon")
    osc = oracle.Scene.from_scene_data(sd)
    gsc = cv.Scene.from_scene_data(sd)
    lo, hi = sd.tris.reshape(-1, 3).min(0), sd.tris.reshape(-1, 3).max(0)
    rays = _ray_families((lo + hi) / 2, float((hi - lo).max()), 90)
    ref = osc.intersect(rays)
    cases = [{}, {"stack_entries": 16}, {"refill_threshold": 65},
             {"refill_threshold": 1, "refill_interval": 1}, {"leaf_batch": 65},
             {"leaf_batch": 1}, {"block_waves": 4, "extend_waves_per_cu": 7},
             {"fetch_batches": 4, "refill_threshold": 64}]
    for kw in cases:
        print("hits-only case", kw, len(rays), flush=True)
        gsc.set_tuning(**kw)
        _check_exact(gsc.intersect_nearest(rays), None, ref)
    # a smaller call after a larger one reuses the queue; a ragged tail
    part = rays[:1001]
    _check_exact(gsc.intersect_nearest(part), None, osc.intersect(part))
