"""CV_RNG_REF_PATCH on the GPU: the reference's own random streams -- one
Rng per 32 x 32 patch seeded with (x, y, frame), its 4 x u64 state shared
by the 8 lanes of a group, groups of a patch in order with the break when
all 8 lanes hit emitters (random.rs:43-136, renderer.rs:292-468) -- bit for
bit against the oracle's restatement, through the patch and the frame entry
points."""
import numpy as np
import pytest

from convector_amd import scenes
from helpers import indoor_scene

pytestmark = pytest.mark.gpu


def _pair(cv, oracle, sd, W, H):
    g = cv.Scene.from_scene_data(sd)
    o = oracle.Scene.from_scene_data(sd)
    return g, o, cv.Renderer(g, W, H, rng_mode=cv.RNG_REF_PATCH)


@pytest.mark.parametrize("case", ["config0", "indoor", "config0-sky"])
def test_patch_calls_exact(cv, oracle, models_dir, case):
    if case == "indoor":
        sd = indoor_scene(models_dir)
    else:
        sd = scenes.config0(subdivisions=4, emission=None if case == "config0-sky" else
                            (10.0, 10.0, 10.0))
    W, H = 128, 96
    g, o, r = _pair(cv, oracle, sd, W, H)
    for (x, y, pw, frame) in [(0, 0, 32, 1), (32, 16, 16, 2), (16, 32, 64, 3), (96, 64, 32, 9)]:
        print("ref patch case %s: patch (%d, %d) width %d frame %d" % (case, x, y, pw, frame),
              flush=True)
        bm = np.full(W * H, 0xDEADBEEF, np.uint32)
        gb = np.full(W * H, 0xDEADBEEF, np.uint32)
        obm, ogb = bm.copy(), gb.copy()
        r.render_patch_u8(bm, gb, pw, x, y, frame)
        assert o.render_patch_u8(W, H, obm, ogb, pw, x, y, frame, oracle.RNG_REF_PATCH) == 0
        np.testing.assert_array_equal(bm, obm)
        np.testing.assert_array_equal(gb, ogb)
        hdr = np.full(W * H * 3, 0.25, np.float32)
        gb2 = np.zeros(W * H, np.uint32)
        ohdr, ogb2 = hdr.copy(), gb2.copy()
        r.accumulate_patch_f32(hdr, gb2, pw, x, y, frame)
        assert o.accumulate_patch_f32(W, H, ohdr, ogb2, pw, x, y, frame,
                                      oracle.RNG_REF_PATCH) == 0
        assert hdr.tobytes() == ohdr.tobytes()
        np.testing.assert_array_equal(gb2, ogb2)
        assert (hdr != 0.25).any()


@pytest.mark.parametrize("W,H,frames", [(96, 48, 3), (80, 64, 2), (320, 192, 2)])
def test_frame_calls_exact(cv, oracle, W, H, frames):
    """Whole frames: 32 x 32 patches, with 16 x 16 patches where fewer than
    32 columns or rows remain (80 = 2 x 32 + 16, 48 = 32 + 16)."""
    sd = scenes.config0(subdivisions=5)
    g, o, r = _pair(cv, oracle, sd, W, H)
    print("ref patch frames %dx%d x %d" % (W, H, frames), flush=True)
    hdr = r.new_buffer_f32()
    gb = np.zeros(W * H, np.uint32)
    r.accumulate_frames_f32(hdr, 1, frames, gbuffer=gb)
    ohdr, ogb = o.accumulate_frames(W, H, 1, frames, oracle.RNG_REF_PATCH, 8)
    assert hdr.tobytes() == ohdr.tobytes()
    np.testing.assert_array_equal(gb, ogb)
    bm, gbu = r.render_frame_u8(5)
    obm, ogbu = o.render_frame_u8(W, H, 5, oracle.RNG_REF_PATCH, 8)
    np.testing.assert_array_equal(bm, obm)
    np.testing.assert_array_equal(gbu, ogbu)


def test_config0_full_frame_exact(cv, oracle):
    """The small configuration at its own size, 1280 x 736, one frame."""
    sd = scenes.config0()
    W, H = 1280, 736
    g, o, r = _pair(cv, oracle, sd, W, H)
    hdr = r.new_buffer_f32()
    gb = np.zeros(W * H, np.uint32)
    r.accumulate_frames_f32(hdr, 1, 1, gbuffer=gb)
    ohdr, ogb = o.accumulate_frames(W, H, 1, 1, oracle.RNG_REF_PATCH, 16)
    assert hdr.tobytes() == ohdr.tobytes()
    np.testing.assert_array_equal(gb, ogb)
    assert hdr.max() > 0


def test_ref_patch_differs_from_pixel_streams_but_estimates_the_same_image(cv, oracle):
    sd = scenes.config0(subdivisions=4)
    W, H, n = 64, 64, 48
    g = cv.Scene.from_scene_data(sd)
    a = cv.Renderer(g, W, H, rng_mode=cv.RNG_REF_PATCH)
    b = cv.Renderer(g, W, H)
    ha, hb = a.new_buffer_f32(), b.new_buffer_f32()
    a.accumulate_frames_f32(ha, 1, n)
    b.accumulate_frames_f32(hb, 1, n)
    assert ha.tobytes() != hb.tobytes()
    ma, mb = ha.sum() / (W * H * n), hb.sum() / (W * H * n)
    assert abs(ma - mb) < 0.03 * max(ma, mb)


def test_ref_patch_errors(cv):
    sd = scenes.config0(subdivisions=2)
    g = cv.Scene.from_scene_data(sd)
    r = cv.Renderer(g, 64, 40, rng_mode=cv.RNG_REF_PATCH)  # 40 rows: not whole 16 x 16 patches
    with pytest.raises(cv.CvError) as e:
        r.accumulate_frames_f32(r.new_buffer_f32(), 1, 1)
    assert e.value.status == 6
    r2 = cv.Renderer(g, 64, 32, rng_mode=cv.RNG_REF_PATCH)
    with pytest.raises(cv.CvError) as e:
        r2.accumulate_aov_f32(r2.new_buffer_f32(), r2.new_features(), 1, 1)
    assert e.value.status == 1
    import ctypes
    o = cv._native.RenderOpts(cv.RNG_REF_PATCH, 0, 5)
    h = ctypes.c_void_p()
    assert cv.lib().cv_batch_create(ctypes.byref(o), 0, ctypes.byref(h)) == 1
