"""cv_batch: many jobs of different scenes, sizes and sample counts in
shared launches.  Each job's HDR accumulation and gbuffer must equal, bit
for bit, the single-job renderer on the same scene, size and frames (which
the renderer tests tie to the oracle), alone or in any mix, split over
several steps or replayed from a captured HIP graph."""
import ctypes

import numpy as np
import pytest

from convector_amd import scenes

pytestmark = pytest.mark.gpu

# (scene index, width, height, first frame, frames)
SPECS = [(0, 64, 48, 1, 3), (1, 160, 96, 5, 2), (2, 32, 16, 1, 4), (0, 96, 64, 2, 1),
         (1, 64, 48, 9, 3), (2, 48, 32, 3, 2)]


@pytest.fixture(scope="module")
def world(cv):
    sds = [scenes.config0(subdivisions=3), scenes.procedural_scene(3000, 7),
           scenes.procedural_scene(1200, 8)]
    gs = [cv.Scene.from_scene_data(sd) for sd in sds]
    expect = []
    for (si, w, h, first, n) in SPECS:
        r = cv.Renderer(gs[si], w, h)
        hdr = r.new_buffer_f32()
        gb = np.zeros(w * h, np.uint32)
        r.accumulate_frames_f32(hdr, first, n, gbuffer=gb)
        expect.append((hdr, gb))
    return sds, gs, expect


def _jobs(cv, gs, specs=SPECS):
    return [cv.Job(gs[si], w, h, first, n, gbuffer=np.zeros(w * h, np.uint32))
            for (si, w, h, first, n) in specs]


def _check(jobs, expect):
    for j, (hdr, gb) in zip(jobs, expect):
        assert j.hdr.tobytes() == hdr.tobytes()
        assert (j.gbuffer == gb).all()


def test_batch_equals_renderer(cv, world):
    sds, gs, expect = world
    b = cv.Batch()
    jobs = _jobs(cv, gs)
    b.render(jobs)
    _check(jobs, expect)
    st = b.stats()
    assert st["paths"] == sum(w * h * n for (_, w, h, _, n) in SPECS)
    assert st["extend_rays"] >= st["paths"]


def test_batch_against_oracle(cv, oracle, world):
    sds, gs, expect = world
    b = cv.Batch()
    jobs = _jobs(cv, gs, SPECS[:2])
    b.render(jobs)
    for (si, w, h, first, n), j in zip(SPECS[:2], jobs):
        ohdr, ogb = oracle.Scene.from_scene_data(sds[si]).accumulate_frames(
            w, h, first, n, oracle.RNG_PIXEL, 8)
        assert j.hdr.tobytes() == ohdr.tobytes()
        assert (j.gbuffer == ogb).all()


def test_batch_split_steps(cv, world):
    """A queue bound below the step's pixel count splits steps by jobs."""
    sds, gs, expect = world
    b = cv.Batch(max_paths_per_step=10000)
    jobs = _jobs(cv, gs)
    b.render(jobs)
    _check(jobs, expect)


def test_batch_packed_frames_and_features(cv, world):
    """Late steps hold several frames of the jobs that still have frames
    (each further frame in its own buffer, folded in frame order); jobs with
    features take one frame per step.  Both give the renderer's bits, also
    over an image that already holds samples."""
    sds, gs, expect = world
    specs = [(0, 64, 48, 1, 9), (1, 32, 16, 2, 1), (2, 48, 32, 3, 6), (1, 64, 32, 1, 4)]
    ref = []
    for k, (si, w, h, first, n) in enumerate(specs):
        r = cv.Renderer(gs[si], w, h)
        hdr = np.full(w * h * 3, 0.25, np.float32)
        feat = r.new_features() if k == 3 else None
        if feat is not None:
            r.accumulate_aov_f32(hdr, feat, first, n)
        else:
            r.accumulate_frames_f32(hdr, first, n)
        ref.append((hdr, feat))
    for bound in (0, 5000):
        print("queue bound", bound, flush=True)
        b = cv.Batch(max_paths_per_step=bound)
        jobs = [cv.Job(gs[si], w, h, first, n, hdr=np.full(w * h * 3, 0.25, np.float32),
                       features=np.zeros(w * h * 8, np.float32) if k == 3 else None)
                for k, (si, w, h, first, n) in enumerate(specs)]
        b.render(jobs)
        for j, (hdr, feat) in zip(jobs, ref):
            assert j.hdr.tobytes() == hdr.tobytes()
            if feat is not None:
                assert j.features.tobytes() == feat.tobytes()


def test_batch_alone_and_in_mix(cv, world):
    sds, gs, expect = world
    b = cv.Batch()
    for k in (3, 0, 5):
        jobs = _jobs(cv, gs, [SPECS[k]])
        b.render(jobs)
        _check(jobs, [expect[k]])


def test_batch_graph_replays(cv, world):
    """CV_RENDER_GRAPH: each step replays one captured launch sequence.
    Several calls in a row, with eager renderer work between them, must
    each give the eager result."""
    sds, gs, expect = world
    b = cv.Batch()
    for rep in range(3):
        print("graph replay", rep, flush=True)
        jobs = _jobs(cv, gs)
        b.render(jobs, flags=cv.RENDER_GRAPH)
        _check(jobs, expect)
        r = cv.Renderer(gs[1], 32, 16)
        h = r.new_buffer_f32()
        r.accumulate_frames_f32(h, 1, 1)


def test_batch_device_pointers(cv, world):
    sds, gs, expect = world
    hip = ctypes.CDLL("libamdhip64.so")
    (si, w, h, first, n) = SPECS[1]
    nb = w * h * 12
    ptr = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(nb)) == 0
    try:
        assert hip.hipMemset(ptr, 0, ctypes.c_size_t(nb)) == 0
        b = cv.Batch()
        b.render([cv.Job(gs[si], w, h, first, n, hdr=ptr.value)], flags=cv.PTR_DEVICE)
        out = np.zeros(w * h * 3, np.float32)
        assert hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ptr, ctypes.c_size_t(nb), 2) == 0
        assert out.tobytes() == expect[1][0].tobytes()
    finally:
        hip.hipFree(ptr)


def test_batch_tuning_variants(cv, world):
    sds, gs, expect = world
    b = cv.Batch()
    for t in (dict(stack_entries=16, refill_threshold=8), dict(stack_entries=32, refill_threshold=65),
              dict(stack_entries=8, extend_waves_per_cu=4, refill_threshold=64),
              dict(stack_entries=24, extend_waves_per_cu=8, refill_threshold=32, block_waves=4),
              dict(stack_entries=8, refill_threshold=16, leaf_batch=65),
              dict(stack_entries=16, refill_threshold=8, leaf_batch=3)):
        print("batch tuning", t, flush=True)
        b.set_tuning(**t)
        jobs = _jobs(cv, gs)
        b.render(jobs)
        _check(jobs, expect)


def test_batch_errors(cv, world):
    sds, gs, expect = world
    b = cv.Batch()
    with pytest.raises(cv.CvError) as e:
        b.render([cv.Job(gs[0], 40, 16, 1, 1)])
    assert e.value.name == "CV_ERR_NOT_MULTIPLE_OF_16"
    with pytest.raises(cv.CvError) as e:
        b.render([cv.Job(gs[0], 32, 16, 1, 0)])
    assert e.value.name == "CV_ERR_INVALID_ARGUMENT"
    with pytest.raises(cv.CvError) as e:
        b.set_tuning(stack_entries=12)
    assert e.value.name == "CV_ERR_INVALID_ARGUMENT"


def test_batch_autotune(cv, world):
    """The autotuner picks a valid launch shape, keeps it on the batch, and
    the batch's results do not change with it."""
    sds, gs, expect = world
    b = cv.Batch()
    jobs = _jobs(cv, gs)
    t, ms = b.autotune(jobs, max_sample_paths=20000)
    print("autotune", t, ms, flush=True)
    assert t["stack_entries"] in (8, 16, 24, 32) and ms > 0
    # without the tolerance flag the reference's tree is kept
    assert t["bvh_width"] == 2 and t["ray_sort_bits"] in (cv.RAY_SORT_NONE, 12, 24)
    assert b.tuning() == t
    b.render(jobs)
    _check(jobs, expect)
