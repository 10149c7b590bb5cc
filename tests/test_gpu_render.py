"""GPU renderer (k_generate / k_extend / k_shade) against the oracle's
restatement of Renderer in PIXEL + portable mode: accumulated radiance,
gbuffer and 8-bit bitmaps bit-identical, through the frame and the patch
entry points; size-independent properties at the benchmark's full size."""
import ctypes

import numpy as np
import pytest

from convector_amd import scenes
from helpers import indoor_scene

pytestmark = pytest.mark.gpu


def _pair(cv, oracle, sd):
    return cv.Scene.from_scene_data(sd), oracle.Scene.from_scene_data(sd)


@pytest.mark.parametrize("case", ["config0", "indoor", "config1-small"])
def test_accumulate_exact(cv, oracle, models_dir, case):
    if case == "config0":
        sd, W, H, n = scenes.config0(), 320, 192, 3
    elif case == "indoor":
        sd, W, H, n = indoor_scene(models_dir), 256, 144, 3
    else:
        sd, W, H, n = scenes.config1(subdivisions=6, n_free=2000), 192, 112, 2
    print("render case", case, W, H, n, flush=True)
    g, o = _pair(cv, oracle, sd)
    r = cv.Renderer(g, W, H)
    hdr = r.new_buffer_f32()
    gb = np.zeros(W * H, np.uint32)
    r.accumulate_frames_f32(hdr, 1, n, gbuffer=gb)
    ohdr, ogb = o.accumulate_frames(W, H, 1, n, oracle.RNG_PIXEL, 8)
    assert hdr.tobytes() == ohdr.tobytes()
    assert (gb == ogb).all()
    assert hdr.max() > 0
    st = r.stats()
    assert st["paths"] == W * H * n and st["extend_rays"] >= W * H * n


def test_render_frame_u8_exact(cv, oracle):
    sd = scenes.config0(subdivisions=4)
    g, o = _pair(cv, oracle, sd)
    W, H = 160, 96
    r = cv.Renderer(g, W, H)
    for frame in (1, 7):
        bm, gb = r.render_frame_u8(frame)
        obm, ogb = o.render_frame_u8(W, H, frame, oracle.RNG_PIXEL, 8)
        assert (bm == obm).all() and (gb == ogb).all()


def test_patch_entry_points(cv, oracle):
    """render_patch_u8 / accumulate_patch_f32 (renderer.rs:306-362) write
    only their patch and equal the oracle's patches."""
    sd = scenes.config0(subdivisions=4)
    g, o = _pair(cv, oracle, sd)
    W, H = 96, 64
    r = cv.Renderer(g, W, H)
    bm = np.full(W * H, 0xDEADBEEF, np.uint32)
    gb = np.full(W * H, 0xDEADBEEF, np.uint32)
    obm = bm.copy()
    ogb = gb.copy()
    for (x, y, pw) in [(0, 0, 32), (32, 32, 32), (64, 0, 16), (80, 48, 16)]:
        r.render_patch_u8(bm, gb, pw, x, y, 5)
        assert o.render_patch_u8(W, H, obm, ogb, pw, x, y, 5, oracle.RNG_PIXEL) == 0
    assert (bm == obm).all() and (gb == ogb).all()
    hdr = np.full(W * H * 3, 0.5, np.float32)
    ohdr = hdr.copy()
    gb2, ogb2 = np.zeros(W * H, np.uint32), np.zeros(W * H, np.uint32)
    for (x, y, pw) in [(0, 0, 32), (32, 32, 32)]:
        r.accumulate_patch_f32(hdr, gb2, pw, x, y, 2)
        assert o.accumulate_patch_f32(W, H, ohdr, ogb2, pw, x, y, 2, oracle.RNG_PIXEL) == 0
    assert hdr.tobytes() == ohdr.tobytes() and (gb2 == ogb2).all()


def test_buffer_f32_into_render_buffer(cv, oracle):
    sd = scenes.config0(subdivisions=3)
    g, o = _pair(cv, oracle, sd)
    W, H = 64, 48
    r = cv.Renderer(g, W, H)
    hdr = r.new_buffer_f32()
    r.accumulate_frames_f32(hdr, 1, 5)
    bm = r.buffer_f32_into_render_buffer(hdr, 5)
    assert (bm == oracle.buffer_f32_into_bitmap(W, H, hdr, 5)).all()


def test_reference_sky_emission(cv, oracle, models_dir):
    """Without an emission override every emitter returns
    sky_intensity(direction) (material.rs:205-215, renderer.rs:452-460)."""
    sd = scenes.config0(subdivisions=3, emission=None)
    g, o = _pair(cv, oracle, sd)
    W, H = 96, 64
    r = cv.Renderer(g, W, H)
    hdr = r.new_buffer_f32()
    r.accumulate_frames_f32(hdr, 1, 2)
    assert hdr.tobytes() == o.accumulate_frames(W, H, 1, 2, oracle.RNG_PIXEL, 8)[0].tobytes()


def test_renderer_preconditions(cv, models_dir):
    sd = scenes.config0(subdivisions=2)
    g = cv.Scene.from_scene_data(sd)
    with pytest.raises(cv.CvError) as e:
        cv.Renderer(g, 100, 64)
    assert e.value.name == "CV_ERR_NOT_MULTIPLE_OF_16"
    r = cv.Renderer(g, 64, 64)
    bm = r.new_bitmap()
    with pytest.raises(cv.CvError) as e:
        r.render_patch_u8(bm, None, 24, 0, 0, 1)
    assert e.value.name == "CV_ERR_NOT_MULTIPLE_OF_16"
    with pytest.raises(cv.CvError) as e:
        r.render_patch_u8(bm, None, 32, 48, 0, 1)
    assert e.value.name == "CV_ERR_OUT_OF_RANGE"
    # fewer than 8 direct-sample triangles (scene.rs:147-160)
    no_light = scenes.SceneData(tris=sd.tris[sd.materials != scenes.MAT_SKY],
                                materials=sd.materials[sd.materials != scenes.MAT_SKY])
    g2 = cv.Scene.from_scene_data(no_light)
    with pytest.raises(cv.CvError) as e:
        cv.Renderer(g2, 64, 64)
    assert e.value.name == "CV_ERR_LIGHT_COUNT"


def test_full_size_properties(cv):
    """The benchmark workload (1.33M triangles, 1920x1080, 16 spp): finite,
    non-negative, deterministic, and the 16-frame accumulation equals the
    frame-by-frame sum in frame order."""
    sd = scenes.config1()
    g = cv.Scene.from_scene_data(sd)
    W, H = 1920, 1080
    r = cv.Renderer(g, W, H)
    a = r.new_buffer_f32()
    r.accumulate_frames_f32(a, 1, 16)
    st = r.stats()
    assert st["paths"] == W * H * 16
    assert np.isfinite(a).all() and (a >= 0).all()
    b = r.new_buffer_f32()
    r.accumulate_frames_f32(b, 1, 16)
    assert a.tobytes() == b.tobytes()
    c = r.new_buffer_f32()
    for f in range(1, 17):
        r.accumulate_frames_f32(c, f, 1)
    assert a.tobytes() == c.tobytes()
    assert 0.005 < float(a.mean()) / 16 < 1.0


def test_tuning_variants_exact(cv, oracle):
    """Launch shapes change only the schedule: every stack size, refill
    threshold (lane refill and whole-batch fetch), refill interval, leaf
    batch (lanes waiting at leaves for a common leaf round, or none) and
    fetch width renders the same bits."""
    sd = scenes.config0(subdivisions=4)
    g, o = _pair(cv, oracle, sd)
    W, H, n = 160, 96, 2
    ohdr, ogb = o.accumulate_frames(W, H, 1, n, oracle.RNG_PIXEL, 8)
    for t in (dict(stack_entries=8, refill_threshold=1, refill_interval=1, leaf_batch=1),
              dict(stack_entries=16, refill_threshold=32, refill_interval=64, leaf_batch=64),
              dict(stack_entries=8, refill_threshold=16, refill_interval=16, leaf_batch=65),
              dict(stack_entries=8, refill_threshold=16, refill_interval=4, leaf_batch=7),
              dict(stack_entries=24, refill_threshold=65, fetch_batches=1),
              dict(stack_entries=32, refill_threshold=65, fetch_batches=4),
              dict(stack_entries=8, extend_waves_per_cu=1, refill_threshold=64, refill_interval=5),
              dict(stack_entries=16, extend_waves_per_cu=8, refill_threshold=32, block_waves=4),
              dict(stack_entries=8, extend_waves_per_cu=6, refill_threshold=16, block_waves=2,
                   refill_interval=16)):
        print("tuning case", t, flush=True)
        g.set_tuning(**t)
        r = cv.Renderer(g, W, H)
        hdr = r.new_buffer_f32()
        gb = np.zeros(W * H, np.uint32)
        r.accumulate_frames_f32(hdr, 1, n, gbuffer=gb)
        assert hdr.tobytes() == ohdr.tobytes()
        assert (gb == ogb).all()


def test_frames_per_step_exact(cv, oracle):
    """Frames rendered in the same launches (each into its own buffer, added
    to the image in frame order) give the bits of one frame per step, on top
    of an image that already holds samples, with partial last steps, and
    with the last frame's gbuffer."""
    sd = scenes.config0(subdivisions=4)
    g, o = _pair(cv, oracle, sd)
    W, H, n = 160, 96, 5
    ohdr, ogb = o.accumulate_frames(W, H, 1, 3 + n, oracle.RNG_PIXEL, 8)
    ref = None
    for k in (1, 2, 3, 5, 64, cv.FRAMES_AUTO):
        print("frames_per_step", k, flush=True)
        g.set_tuning(frames_per_step=k)
        assert g.tuning()["frames_per_step"] == k
        r = cv.Renderer(g, W, H)
        hdr = r.new_buffer_f32()
        gb = np.zeros(W * H, np.uint32)
        r.accumulate_frames_f32(hdr, 1, 3)
        r.accumulate_frames_f32(hdr, 4, n, gbuffer=gb)
        assert hdr.tobytes() == ohdr.tobytes()
        assert (gb == ogb).all()
        if ref is None:
            ref = hdr
        # device buffers and graph replay take the same path
        hip = ctypes.CDLL("libamdhip64.so")
        p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(hdr.nbytes)) == 0
        try:
            assert hip.hipMemset(p, 0, ctypes.c_size_t(hdr.nbytes)) == 0
            r.accumulate_frames_device(p.value, 1, 3 + n, cv.RENDER_GRAPH)
            out = np.zeros_like(hdr)
            assert hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p,
                                 ctypes.c_size_t(hdr.nbytes), 2) == 0
            assert out.tobytes() == ohdr.tobytes()
        finally:
            hip.hipFree(p)
    with pytest.raises(cv.CvError):
        g.set_tuning(frames_per_step=65)
    g.set_tuning(frames_per_step=cv.FRAMES_AUTO)


def test_trace_export(cv, tmp_path):
    """The device trace log: every launch of the frames as a Chrome trace
    span (trace.rs:122-146 format), bounded by the ring limit."""
    import json
    sd = scenes.config0(subdivisions=3)
    g = cv.Scene.from_scene_data(sd)
    g.set_tuning(frames_per_step=1)  # one frame per step: its own launches
    r = cv.Renderer(g, 64, 32)
    hdr = r.new_buffer_f32()
    plain = r.new_buffer_f32()
    r.accumulate_frames_f32(plain, 1, 3)
    cv.trace_enable(1000)
    try:
        r.accumulate_frames_f32(hdr, 1, 3)
        path = str(tmp_path / "trace.json")
        cv.trace_export(path)
        ev = json.load(open(path))["traceEvents"]
        assert len(ev) == 3 * (1 + 2 * 5)
        assert [e["name"] for e in ev[:3]] == ["generate", "extend", "shade"]
        assert all(e["ph"] == "X" and e["dur"] > 0 and e["ts"] >= 0 for e in ev)
        assert [e["args"]["frame"] for e in ev[::11]] == [1, 2, 3]
        assert sorted(e["ts"] for e in ev) == [e["ts"] for e in ev]
        cv.trace_enable(5)
        r.accumulate_frames_f32(r.new_buffer_f32(), 1, 1)
        cv.trace_export(path)
        ev = json.load(open(path))["traceEvents"]
        assert len(ev) == 5 and ev[-1]["name"] == "shade" and ev[-1]["args"]["id"] == 4
        # the three frames in the same launches: one generate, 5 x (extend, shade)
        g.set_tuning(frames_per_step=cv.FRAMES_AUTO)
        cv.trace_enable(1000)
        together = r.new_buffer_f32()
        r.accumulate_frames_f32(together, 1, 3)
        cv.trace_export(path)
        ev = json.load(open(path))["traceEvents"]
        assert len(ev) == 1 + 2 * 5
        assert together.tobytes() == plain.tobytes()
    finally:
        cv.trace_disable()
    # tracing changes nothing in the image
    assert hdr.tobytes() == plain.tobytes()


@pytest.mark.parametrize("W,H", [(64, 52), (80, 44), (48, 12)])
def test_leftover_band_exact(cv, oracle, W, H):
    """Heights of 4 mod 8: the camera paths' 8 x 8 tiles leave a last band
    of 4 rows, rendered and compared pixel for pixel (the oracle's PIXEL
    frames take heights that are multiples of 4)."""
    sd = scenes.config0(subdivisions=4)
    g, o = _pair(cv, oracle, sd)
    r = cv.Renderer(g, W, H)
    hdr = r.new_buffer_f32()
    gb = np.full(W * H, 0xDEADBEEF, np.uint32)
    r.accumulate_frames_f32(hdr, 1, 2, gbuffer=gb)
    ohdr, ogb = o.accumulate_frames(W, H, 1, 2, oracle.RNG_PIXEL, 8)
    assert hdr.tobytes() == ohdr.tobytes()
    np.testing.assert_array_equal(gb, ogb)
    # every pixel of the band received its paths
    rows = oracle.hdr_blocks_to_rows(W, H, hdr)
    assert (rows[H - 4:].max(axis=2) > 0).mean() > 0.3
    bm, gbu = r.render_frame_u8(3)
    obm, ogbu = o.render_frame_u8(W, H, 3, oracle.RNG_PIXEL, 8)
    np.testing.assert_array_equal(bm, obm)
    np.testing.assert_array_equal(gbu, ogbu)


def test_config0_full_size_exact(cv, oracle):
    """The small configuration at its own size: 1280 x 736, 1 spp."""
    sd = scenes.config0()
    g, o = _pair(cv, oracle, sd)
    W, H = 1280, 736
    r = cv.Renderer(g, W, H)
    hdr = r.new_buffer_f32()
    gb = np.zeros(W * H, np.uint32)
    r.accumulate_frames_f32(hdr, 1, 1, gbuffer=gb)
    ohdr, ogb = o.accumulate_frames(W, H, 1, 1, oracle.RNG_PIXEL, 16)
    assert hdr.tobytes() == ohdr.tobytes()
    np.testing.assert_array_equal(gb, ogb)


def test_config1_full_size_exact(cv, oracle):
    """The large configuration at its own size: 1,330,746 triangles at
    1920 x 1080 (135 rows of 8), 1 spp, against the oracle traversing the
    same tree (the GPU's BVH, itself pinned to the oracle's build)."""
    sd = scenes.config1()
    g = cv.Scene.from_scene_data(sd)
    nodes, order = g.bvh()
    o = oracle.Scene.from_scene_data(sd, nodes=nodes.view(oracle.NODE_DTYPE), order=order)
    W, H = 1920, 1080
    r = cv.Renderer(g, W, H)
    hdr = r.new_buffer_f32()
    gb = np.zeros(W * H, np.uint32)
    r.accumulate_frames_f32(hdr, 1, 1, gbuffer=gb)
    print("config1 full size: GPU done", flush=True)
    ohdr, ogb = o.accumulate_frames(W, H, 1, 1, oracle.RNG_PIXEL, 16)
    assert hdr.tobytes() == ohdr.tobytes()
    np.testing.assert_array_equal(gb, ogb)
