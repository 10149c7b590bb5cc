"""The AVX2 8-wide baseline (oracle/cv_oracle_avx.c) against the scalar
oracle.

With each lane traversed alone (per_lane) the AVX2 code must reproduce the
scalar oracle bit for bit, in both approximation modes: camera, Rng draw
order, traversal, triangle picks, shading and accumulation.  The
reference's packet traversal (the timed baseline) differs from a single-ray
traversal for rare rays: a box is visited when any lane's slab test hits
it, so a lane can hit a triangle in a box its own slab test missed (with
rcpps reciprocals the slabs are only ~12-bit accurate), and ties between
triangles at the same distance are decided in another visit order
(bvh.rs:589-657).  The packet tests bound those differences."""
import numpy as np
import pytest

from convector_amd import scenes
from oracle import oracle as _O

pytestmark = pytest.mark.skipif(not _O.avx_available(), reason="host without AVX2 + FMA")

APPROX = [_O.APPROX_PORTABLE, _O.APPROX_HOST_X86]


def _scenes():
    return [("config0", scenes.config0(subdivisions=3)),
            ("procedural", scenes.procedural_scene(6000, 7))]


@pytest.mark.parametrize("approx", APPROX)
def test_per_lane_identical(oracle, approx):
    for name, sd in _scenes():
        sc = oracle.Scene.from_scene_data(sd, approx=approx)
        W, H = 96, 64
        ref, _ = sc.accumulate_frames(W, H, 1, 2, oracle.RNG_REF_PATCH, 8)
        got, _ = sc.avx_accumulate_frames(W, H, 1, 2, per_lane=True, nthreads=8)
        assert np.asarray(ref).tobytes() == got.tobytes(), name
        assert np.abs(got).sum() > 0
        _, q = sc.avx_accumulate_frames(W, H, 3, 1, per_lane=True, nthreads=4)
        nq, _, _ = sc.frame_work(W, H, 3, oracle.RNG_REF_PATCH, 4)
        assert q == nq, name


@pytest.mark.parametrize("approx", APPROX)
def test_packet_differences_bounded(oracle, approx):
    for name, sd in _scenes():
        sc = oracle.Scene.from_scene_data(sd, approx=approx)
        W, H = 128, 64
        ref, _ = sc.accumulate_frames(W, H, 1, 2, oracle.RNG_REF_PATCH, 8)
        got, q = sc.avx_accumulate_frames(W, H, 1, 2, nthreads=8)
        _, q1 = sc.avx_accumulate_frames(W, H, 1, 2, per_lane=True, nthreads=8)
        ref = oracle.hdr_blocks_to_rows(W, H, np.asarray(ref))
        rows = oracle.hdr_blocks_to_rows(W, H, got)
        differ = (ref != rows).any(axis=-1)
        print(name, approx, "pixels differing from single-ray traversal:", int(differ.sum()),
              "of", differ.size, "queries", q, "vs", q1)
        assert differ.mean() < 2e-3
        assert abs(q - q1) <= 1e-3 * q1


def test_sample_patches_matches_frames(oracle):
    """A sample of patches does the packet path's work for those patches."""
    sd = scenes.config0(subdivisions=3)
    sc = oracle.Scene.from_scene_data(sd, approx=oracle.APPROX_PORTABLE)
    W, H = 128, 64
    q_all = sum(sc.avx_sample_patches(W, H, 1, [p], nthreads=1) for p in range(8))
    _, q_frame = sc.avx_accumulate_frames(W, H, 1, 1, nthreads=2)
    assert q_all == q_frame > 0
    qs, _, _ = sc.sample_patches(W, H, 1, list(range(8)), rng_mode=oracle.RNG_REF_PATCH,
                                 nthreads=2)
    assert abs(q_all - qs) <= 1e-3 * qs
    with pytest.raises(RuntimeError):
        sc.avx_sample_patches(W, H, 1, [1000], nthreads=1)
