"""The ray sort (cv_tuning.ray_sort_bits): before every bounce after the
camera rays the paths are traversed in the order of a key (job, direction
octant, Morton code of the origin).  Only the assignment of rays to lanes
changes, so renders, gbuffers and the summed test counters are the same
bits as in queue order -- for one scene, for multi-job batches (where the
job index takes key bits from the Morton code), in graphs and on the
server."""
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


@pytest.mark.parametrize("bits", [3, 12, 24])
def test_sorted_render_equals_queue_order(cv, bits):
    sd = scenes.config0()
    g = cv.Scene.from_scene_data(sd)
    W, H, n = 320, 192, 2
    ref, gb_ref, st_ref = _render(cv, g, W, H, n, cv.RENDER_COUNT)
    g.set_tuning(ray_sort_bits=bits)
    assert g.tuning()["ray_sort_bits"] == bits
    hdr, gb, st = _render(cv, g, W, H, n, cv.RENDER_COUNT)
    print("sort %d bits: %d rays, %.3f ms extend" % (bits, st["extend_rays"], st["extend_ms"]),
          flush=True)
    assert hdr.tobytes() == ref.tobytes()
    assert (gb == gb_ref).all()
    for k in ("extend_rays", "sum_numi_aabb", "sum_numi_tri"):
        assert st[k] == st_ref[k], k
    # timed launches report the sort's device time; graphs replay it
    _, _, st_t = _render(cv, g, W, H, n, cv.RENDER_TIMING)
    assert st_t["sort_ms"] > 0
    hdr_g, _, _ = _render(cv, g, W, H, n, cv.RENDER_GRAPH)
    hdr_g2, _, _ = _render(cv, g, W, H, n, cv.RENDER_GRAPH)
    assert hdr_g.tobytes() == ref.tobytes() and hdr_g2.tobytes() == ref.tobytes()
    g.set_tuning(ray_sort_bits=cv.RAY_SORT_NONE)
    assert g.tuning()["ray_sort_bits"] == cv.RAY_SORT_NONE


def test_sorted_wide_render(cv):
    """The sort under the wide traversal equals the wide traversal alone."""
    g = cv.Scene.from_scene_data(scenes.procedural_scene(40000, 77))
    g.set_tuning(bvh_width=4)
    ref, _, _ = _render(cv, g, 256, 128, 2)
    g.set_tuning(ray_sort_bits=24)
    hdr, _, _ = _render(cv, g, 256, 128, 2)
    assert hdr.tobytes() == ref.tobytes()


@pytest.mark.parametrize("njobs", [3, 40])
def test_sorted_batch(cv, njobs):
    """Multi-job steps: 40 jobs take 6 key bits, so a 24-bit key keeps 5
    Morton bits per axis."""
    sds = [scenes.procedural_scene(3000 + 1500 * (k % 5), 300 + k) for k in range(njobs)]
    gs = [cv.Scene.from_scene_data(sd) for sd in sds]
    sizes = [(64, 32), (32, 48), (96, 16), (48, 64)]
    specs = [(gs[k], sizes[k % 4][0], sizes[k % 4][1], 1 + k % 3) for k in range(njobs)]

    def run(bits, flags):
        b = cv.Batch()
        b.set_tuning(ray_sort_bits=bits)
        jobs = [cv.Job(g, w, h, 1, n, gbuffer=np.zeros(w * h, np.uint32))
                for (g, w, h, n) in specs]
        b.render(jobs, flags)
        return jobs

    ref = run(cv.RAY_SORT_NONE, 0)
    for bits in (12, 24):
        for flags in (0, cv.RENDER_GRAPH):
            out = run(bits, flags)
            for a, r in zip(out, ref):
                assert a.hdr.tobytes() == r.hdr.tobytes(), (bits, flags)
                assert (a.gbuffer == r.gbuffer).all(), (bits, flags)


def test_sorted_server(cv):
    g = cv.Scene.from_scene_data(scenes.procedural_scene(20000, 91))
    pos, q = (0.1, 0.7, 0.0), (0.70710677, -0.70710677, 0.0, 0.0)
    out = []
    for bits in (cv.RAY_SORT_NONE, 12):
        sv = cv.Server()
        sid = sv.add_scene(g)
        sv.set_tuning(ray_sort_bits=bits)
        t, hdr = sv.submit(sid, 64, 32, 3, pos, q)
        sv.wait(t)
        sv.close()
        out.append(hdr)
    assert out[0].tobytes() == out[1].tobytes()


def test_sort_bits_validation(cv):
    g = cv.Scene.from_scene_data(scenes.procedural_scene(2000, 5))
    for bad in (1, 13, 27, 30):
        with pytest.raises(cv.CvError) as e:
            g.set_tuning(ray_sort_bits=bad)
        assert e.value.name == "CV_ERR_INVALID_ARGUMENT"
