"""GPU BVH build (build.hip) against the oracle's restatement of
Bvh::build: the node array, the triangle order and the statistics must be
bit-identical, on the reference's models, the configuration scenes and
adversarial inputs (sorted-bin fallbacks, ties, tiny scenes, errors)."""
import numpy as np
import pytest

from convector_amd import scenes
from helpers import model_scene

pytestmark = pytest.mark.gpu


def _compare(cv, oracle, tris, materials=None, label=""):
    print("bvh case", label, len(tris), flush=True)
    materials = np.zeros(len(tris), np.uint32) if materials is None else materials
    st, on, oo, (ratio, per_leaf) = oracle.bvh_build(tris)
    assert st == 0
    mesh = cv.Mesh.from_triangles(tris, materials)
    s = cv.Scene.from_meshes([mesh])
    gn, go = s.bvh()
    info = s.bvh_info()
    assert len(gn) == len(on), label
    assert gn.tobytes() == on.tobytes(), label
    assert (go == oo).all(), label
    assert info["avg_tris_per_leaf"] == np.float32(per_leaf)
    assert info["avg_area_ratio"] == np.float32(ratio)
    return info


@pytest.mark.parametrize("name", ["suzanne", "indoor", "box_windows", "suzannes_in_box",
                                  "stanford_bunny", "stanford_dragon"])
def test_models(cv, oracle, models_dir, name):
    sd = model_scene(models_dir, name)
    _compare(cv, oracle, sd.tris, sd.materials, name)


@pytest.mark.parametrize("sub", [2, 5])
def test_config0(cv, oracle, sub):
    sd = scenes.config0(subdivisions=sub)
    _compare(cv, oracle, sd.tris, sd.materials, "config0-%d" % sub)


def test_config1_small(cv, oracle):
    sd = scenes.config1(subdivisions=6, n_free=3000)
    _compare(cv, oracle, sd.tris, sd.materials, "config1-small")


def test_config1_full(cv, oracle):
    """The 1,330,746-triangle scene of the benchmark."""
    sd = scenes.config1()
    info = _compare(cv, oracle, sd.tris, sd.materials, "config1")
    assert info["num_nodes"] == 2602928


@pytest.mark.parametrize("n", [2, 3, 5, 17, 31, 32, 33, 40, 63, 64, 65, 100, 257])
def test_tiny_and_boundary_sizes(cv, oracle, n):
    rng = np.random.default_rng(n)
    c = rng.uniform(-1, 1, (n, 3))
    tris = np.concatenate([c + rng.normal(0, 0.05, (n, 3)) for _ in range(3)], 1)
    _compare(cv, oracle, tris.astype(np.float32), label="random-%d" % n)


def test_sorted_fallback_large(cv, oracle):
    """A dense cluster and far outliers: centroid bins overflow n/8 and the
    node falls back to sorted bins (bvh.rs:159-169), at several levels."""
    rng = np.random.default_rng(8)
    c = np.concatenate([rng.normal(0, 0.001, (9000, 3)), rng.uniform(-50, 50, (1000, 3))])
    tris = np.concatenate([c + rng.normal(0, 0.0005, c.shape) for _ in range(3)], 1)
    _compare(cv, oracle, tris.astype(np.float32), label="fallback")


def test_sorted_fallback_wide(cv, oracle):
    """The same cluster and outliers at 40000 triangles: the top nodes are
    binned and partitioned in chunks (nodes of more than 4096 triangles)
    and fall back to sorted bins there."""
    rng = np.random.default_rng(9)
    c = np.concatenate([rng.normal(0, 0.001, (36000, 3)), rng.uniform(-50, 50, (4000, 3))])
    tris = np.concatenate([c + rng.normal(0, 0.0005, c.shape) for _ in range(3)], 1)
    _compare(cv, oracle, tris.astype(np.float32), label="fallback wide")


def test_grid_ties(cv, oracle):
    """A regular grid of equal triangles: many equal barycenter coordinates,
    so the stable order of the sorts decides the tree."""
    g = np.stack(np.meshgrid(np.arange(40), np.arange(40), np.arange(3)), -1).reshape(-1, 3)
    c = g.astype(np.float64) * 0.1
    tris = np.concatenate([c, c + [0.05, 0, 0], c + [0, 0.05, 0]], 1).astype(np.float32)
    _compare(cv, oracle, tris, label="grid")


def test_flat_wall(cv, oracle):
    """Coplanar triangles (zero extent on one axis: NaN centroid bins on it)."""
    rng = np.random.default_rng(4)
    xy = rng.uniform(-1, 1, (500, 2))
    tris = np.zeros((500, 9))
    for k in range(3):
        tris[:, 3 * k:3 * k + 2] = xy + rng.normal(0, 0.02, (500, 2))
        tris[:, 3 * k + 2] = 0.25
    _compare(cv, oracle, tris.astype(np.float32), label="flat")


def test_build_errors(cv):
    one = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32)
    for tris, status in [(one, "CV_ERR_BVH_NO_SPLIT"), (np.repeat(one, 40, 0), "CV_ERR_BVH_NO_SPLIT")]:
        mesh = cv.Mesh.from_triangles(tris, np.zeros(len(tris), np.uint32))
        with pytest.raises(cv.CvError) as e:
            cv.Scene.from_meshes([mesh])
        assert e.value.name == status
    bad = np.repeat(one, 3, 0)
    bad[1] += np.array([0, 0, 0, 0, 0, 0, 0, 0, 1], np.float32)
    bad[2, 0] = np.nan
    mesh = cv.Mesh.from_triangles(bad, np.zeros(3, np.uint32))
    with pytest.raises(cv.CvError) as e:
        cv.Scene.from_meshes([mesh])
    assert e.value.name == "CV_ERR_NAN"


def test_multiple_meshes_concatenate(cv, oracle, models_dir):
    """Bvh::from_meshes concatenates meshes in order (bvh.rs:547-568)."""
    a = model_scene(models_dir, "suzanne")
    b = model_scene(models_dir, "indoor")
    ma = cv.Mesh.from_triangles(a.tris, a.materials)
    mb = cv.Mesh.from_triangles(b.tris, b.materials)
    s = cv.Scene.from_meshes([ma, mb])
    gn, go = s.bvh()
    st, on, oo, _ = oracle.bvh_build(np.concatenate([a.tris, b.tris]))
    assert gn.tobytes() == on.tobytes() and (go == oo).all()
    assert s.bvh_info()["num_direct"] == 8
