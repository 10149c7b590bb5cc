"""Queue partitions of the traversal (cv_tuning.queue_parts = 8): the queue
is cut into chunks dealt to 8 partitions in turn, and each wave takes rays
from the partition of its XCD before the others'.  Only which lane
traverses which ray changes, so hits, renders, gbuffers and the summed
test counters are the same bits as with one cursor -- at sizes where some
partitions are empty, where the last chunk is partial, for one scene, for
multi-job batches, under the wide and sorted traversals, in graphs and on
the server."""
import numpy as np
import pytest

from convector_amd import scenes

pytestmark = pytest.mark.gpu


def _render(cv, g, w, h, n, flags=0):
    r = cv.Renderer(g, w, h)
    hdr = r.new_buffer_f32()
    gb = np.zeros(w * h, np.uint32)
    r.accumulate_frames_f32(hdr, 1, n, flags=flags, gbuffer=gb)
    return hdr, gb, r.stats()


@pytest.mark.parametrize("w,h,n", [(320, 192, 2), (64, 48, 1), (1280, 736, 1)])
def test_partitioned_render_equals_one_cursor(cv, w, h, n):
    """122,880 paths: chunks of 8192, partitions of 1-2 chunks; 3,072
    paths: one partial chunk, seven empty partitions; 942,080 paths: chunks
    of 2^16."""
    print("render %dx%d x %d frames" % (w, h, n), flush=True)
    g = cv.Scene.from_scene_data(scenes.config0())
    g.set_tuning(queue_parts=1)
    ref, gb_ref, st_ref = _render(cv, g, w, h, n, cv.RENDER_COUNT)
    g.set_tuning(queue_parts=8)
    assert g.tuning()["queue_parts"] == 8
    hdr, gb, st = _render(cv, g, w, h, n, cv.RENDER_COUNT)
    assert hdr.tobytes() == ref.tobytes()
    assert (gb == gb_ref).all()
    for k in ("extend_rays", "sum_numi_aabb", "sum_numi_tri"):
        assert st[k] == st_ref[k], k
    hdr_g, _, _ = _render(cv, g, w, h, n, cv.RENDER_GRAPH)
    hdr_g2, _, _ = _render(cv, g, w, h, n, cv.RENDER_GRAPH)
    assert hdr_g.tobytes() == ref.tobytes() and hdr_g2.tobytes() == ref.tobytes()


@pytest.mark.parametrize("nrays", [1, 100, 4096 * 8, 4096 * 8 + 1, 65536 * 8 + 4095, 2500000])
def test_partitioned_queries(cv, nrays):
    print("queries: %d rays" % nrays, flush=True)
    sd = scenes.config0()
    g = cv.Scene.from_scene_data(sd)
    rays = scenes.rays_inward(nrays, 7 + nrays % 97, radius=0.9, target_scale=0.6)
    g.set_tuning(queue_parts=1)
    ref = g.intersect_nearest(rays)
    g.set_tuning(queue_parts=8)
    got = g.intersect_nearest(rays)
    assert got.tobytes() == ref.tobytes()


def test_partitioned_wide_and_sorted(cv):
    g = cv.Scene.from_scene_data(scenes.procedural_scene(40000, 77))
    g.set_tuning(bvh_width=4, ray_sort_bits=12, queue_parts=1)
    ref, _, _ = _render(cv, g, 256, 128, 2)
    g.set_tuning(queue_parts=8)
    hdr, _, _ = _render(cv, g, 256, 128, 2)
    assert hdr.tobytes() == ref.tobytes()


@pytest.mark.parametrize("njobs", [3, 40])
def test_partitioned_batch(cv, njobs):
    sds = [scenes.procedural_scene(3000 + 1500 * (k % 5), 300 + k) for k in range(njobs)]
    gs = [cv.Scene.from_scene_data(sd) for sd in sds]
    sizes = [(64, 32), (32, 48), (96, 16), (48, 64)]
    specs = [(gs[k], sizes[k % 4][0], sizes[k % 4][1], 1 + k % 3) for k in range(njobs)]

    def run(parts, flags):
        b = cv.Batch()
        b.set_tuning(queue_parts=parts)
        assert b.tuning()["queue_parts"] == parts
        jobs = [cv.Job(g, w, h, 1, n, gbuffer=np.zeros(w * h, np.uint32))
                for (g, w, h, n) in specs]
        b.render(jobs, flags)
        return jobs

    ref = run(1, 0)
    for flags in (0, cv.RENDER_GRAPH):
        out = run(8, flags)
        for a, r in zip(out, ref):
            assert a.hdr.tobytes() == r.hdr.tobytes(), flags
            assert (a.gbuffer == r.gbuffer).all(), flags


def test_partitioned_server(cv):
    g = cv.Scene.from_scene_data(scenes.procedural_scene(20000, 91))
    pos, q = (0.1, 0.7, 0.0), (0.70710677, -0.70710677, 0.0, 0.0)
    out = []
    for parts in (1, 8):
        sv = cv.Server()
        sid = sv.add_scene(g)
        sv.set_tuning(queue_parts=parts)
        t, hdr = sv.submit(sid, 64, 32, 3, pos, q)
        sv.wait(t)
        sv.close()
        out.append(hdr)
    assert out[0].tobytes() == out[1].tobytes()


def test_queue_parts_validation(cv):
    g = cv.Scene.from_scene_data(scenes.procedural_scene(2000, 5))
    for bad in (2, 4, 7, 9, 16):
        with pytest.raises(cv.CvError) as e:
            g.set_tuning(queue_parts=bad)
        assert e.value.name == "CV_ERR_INVALID_ARGUMENT"
