"""GPU builds with 32 and 16 centroid bins (CV_BUILDER_BINS_32 / _16): the
node array and triangle order bit-identical to the oracle's build with the
same bins, on models, configuration scenes and adversarial inputs; the
traversal and renders over such a tree bit-identical to the oracle's over
the same tree; and against the reference's 64-bin tree, the nearest hits
equal except at exact ties or within 1e-6 of an edge."""
import numpy as np
import pytest

from convector_amd import scenes
from helpers import bits, check_tolerance, model_scene

pytestmark = pytest.mark.gpu


def _compare(cv, oracle, tris, bins, materials=None, label=""):
    print("bins %d case %s: %d triangles" % (bins, label, len(tris)), flush=True)
    materials = np.zeros(len(tris), np.uint32) if materials is None else materials
    st, on, oo, (ratio, per_leaf) = oracle.bvh_build(tris, bins)
    assert st == 0
    s = cv.Scene.from_meshes([cv.Mesh.from_triangles(tris, materials)], bins=bins)
    gn, go = s.bvh()
    info = s.bvh_info()
    assert info["bins"] == bins
    assert len(gn) == len(on), label
    assert gn.tobytes() == on.tobytes(), label
    assert (go == oo).all(), label
    assert info["avg_tris_per_leaf"] == np.float32(per_leaf)
    assert info["avg_area_ratio"] == np.float32(ratio)
    return s, info


@pytest.mark.parametrize("bins", [16, 32])
@pytest.mark.parametrize("name", ["suzanne", "indoor", "stanford_bunny", "stanford_dragon"])
def test_models(cv, oracle, models_dir, name, bins):
    sd = model_scene(models_dir, name)
    _compare(cv, oracle, sd.tris, bins, sd.materials, name)


@pytest.mark.parametrize("bins", [16, 32])
def test_config_scenes(cv, oracle, bins):
    sd = scenes.config0()
    _compare(cv, oracle, sd.tris, bins, sd.materials, "config0")
    sd = scenes.config1(subdivisions=6, n_free=3000)
    _compare(cv, oracle, sd.tris, bins, sd.materials, "config1-small")


@pytest.mark.parametrize("bins", [16, 32])
@pytest.mark.parametrize("n", [2, 3, 7, 8, 9, 15, 16, 17, 31, 33, 65, 100, 257])
def test_boundary_sizes(cv, oracle, n, bins):
    """Around bins / 2 (sorted single-triangle bins below, centroid bins
    above) and bins (the fallback test needs n > bins)."""
    rng = np.random.default_rng(1000 + n)
    c = rng.uniform(-1, 1, (n, 3))
    tris = np.concatenate([c + rng.normal(0, 0.05, (n, 3)) for _ in range(3)], 1)
    _compare(cv, oracle, tris.astype(np.float32), bins, label="random-%d" % n)


@pytest.mark.parametrize("bins", [16, 32])
@pytest.mark.parametrize("n", [10000, 40000])
def test_sorted_fallback(cv, oracle, n, bins):
    """A dense cluster and outliers: bins overflow n/8 (more often with
    fewer bins) and nodes take the sorted chunks, chunked above 4096."""
    rng = np.random.default_rng(n + bins)
    k = n * 9 // 10
    c = np.concatenate([rng.normal(0, 0.001, (k, 3)), rng.uniform(-50, 50, (n - k, 3))])
    tris = np.concatenate([c + rng.normal(0, 0.0005, c.shape) for _ in range(3)], 1)
    _compare(cv, oracle, tris.astype(np.float32), bins, label="fallback")


@pytest.mark.parametrize("bins", [16, 32])
def test_grid_ties(cv, oracle, bins):
    """Coincident barycenter coordinates on a grid: many equal costs."""
    g = np.stack(np.meshgrid(np.arange(20), np.arange(20), np.arange(4)), -1).reshape(-1, 3)
    c = g.astype(np.float32) * 0.1
    tris = np.concatenate([c, c + [0.05, 0, 0], c + [0, 0.05, 0]], 1).astype(np.float32)
    _compare(cv, oracle, tris, bins, label="grid")


@pytest.mark.parametrize("bins", [16, 32])
def test_traversal_and_render_over_binned_tree(cv, oracle, bins):
    """Exact against the oracle on the same tree; within the tolerance
    against the reference's tree."""
    sd = scenes.config0()
    g = cv.Scene.from_scene_data(sd, bins=bins)
    nodes, order = g.bvh()
    o_same = oracle.Scene.from_scene_data(sd, nodes=nodes.view(oracle.NODE_DTYPE), order=order)
    o_ref = oracle.Scene.from_scene_data(sd)
    rays = np.concatenate([scenes.rays_inward(20000, 51, radius=0.9, target_scale=0.45),
                           scenes.rays_inward(20000, 52, radius=0.95, target_scale=0.95)])
    hits, isects = g.intersect_nearest(rays, full=True)
    same = o_same.intersect(rays)
    assert (hits["tri_id"] == same["tri"]).all()
    assert (bits(hits["t"]) == bits(same["tuv"][:, 0])).all()
    # the counters of this tree, as the oracle counts them on it
    assert (isects["numi_aabb"] == same["numi_aabb"]).all()
    assert (isects["numi_tri"] == same["numi_tri"]).all()
    ref = o_ref.intersect(rays)
    check_tolerance(sd, hits, ref, rays)
    W, H = 320, 192
    r = cv.Renderer(g, W, H)
    hdr = r.new_buffer_f32()
    r.accumulate_frames_f32(hdr, 1, 2)
    ohdr, _ = o_same.accumulate_frames(W, H, 1, 2, oracle.RNG_PIXEL, 16)
    assert hdr.tobytes() == ohdr.tobytes()
    g_ref = cv.Scene.from_scene_data(sd)
    r_ref = cv.Renderer(g_ref, W, H)
    hdr_ref = r_ref.new_buffer_f32()
    r_ref.accumulate_frames_f32(hdr_ref, 1, 2)
    a = oracle.hdr_blocks_to_rows(W, H, hdr)
    b = oracle.hdr_blocks_to_rows(W, H, hdr_ref)
    frac = (a == b).all(axis=2).mean()
    print("bins %d: %.5f of the pixels equal the 64-bin render" % (bins, frac), flush=True)
    assert frac >= 0.999


def test_builder_option_errors(cv):
    tris = scenes.procedural_scene(500, 3).tris
    with pytest.raises(cv.CvError) as e:
        cv.Scene.from_meshes([cv.Mesh.from_triangles(tris, np.zeros(len(tris), np.uint32))],
                             bins=48)
    assert e.value.name == "CV_ERR_INVALID_ARGUMENT"
