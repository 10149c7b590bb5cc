"""The oracle's renderer (renderer.rs) against closed forms and layout
restatements: the 16x4 block layout, the 8-bit packing, sky radiance of
rays that miss everything, patch/frame consistency and the reference's
preconditions."""
import math

import numpy as np
import pytest

from convector_amd import scenes
from helpers import model_scene


def _block_index(x, y, width):
    # renderer.rs:154-170 + 356-358: group g = 2*(dx//4) + dy//2, lane l = dx%4 + 4*(dy%2)
    dx, dy = x % 16, y % 4
    g = (dx // 4) * 2 + dy // 2
    lane = dx % 4 + 4 * (dy % 2)
    blk = (y // 4) * (width // 16) + x // 16
    return blk * 192 + g * 24 + lane


def test_block_layout_roundtrip(oracle):
    W, H = 64, 32
    hdr = np.arange(W * H * 3, dtype=np.float32)
    rows = oracle.hdr_blocks_to_rows(W, H, hdr)
    for y in range(H):
        for x in range(W):
            o = _block_index(x, y, W)
            assert tuple(rows[y, x]) == (hdr[o], hdr[o + 8], hdr[o + 16])


def _pack(c):
    f = np.float32
    v = [min(f(1), f(ch) * f(2)) * f(255) for ch in c]
    r = [int(np.rint(x)) for x in v]
    return r[0] | (r[1] << 8) | (r[2] << 16)


def test_pack_color_known_answers(oracle):
    W, H = 16, 4
    hdr = np.zeros(W * H * 3, np.float32)
    cols = [(0.25, 0.1, 1.0), (0.0, 0.0, 0.0), (0.5, 0.49, 0.001), (10, 0.3, 0.2)]
    for k, c in enumerate(cols):
        o = _block_index(k, 0, W)
        hdr[o], hdr[o + 8], hdr[o + 16] = c
    bm = oracle.buffer_f32_into_bitmap(W, H, hdr, 1)
    for k, c in enumerate(cols):
        assert bm[k] == _pack(c)
    assert _pack((0.25, 0.1, 1.0)) == 128 | (51 << 8) | (255 << 16)  # 127.5 rounds to even
    # num_samples divides (renderer.rs:384-390)
    bm2 = oracle.buffer_f32_into_bitmap(W, H, hdr * np.float32(4), 4)
    assert (bm2 == bm).all()


def _sky_scene():
    """8 light triangles far behind the camera, nothing else: every camera ray
    misses and returns sky_intensity(direction) (material.rs:205-215)."""
    tris, mats = [], []
    for k in range(8):
        z = 50.0 + k
        tris.append([0, z, -0.5, 1, z, -0.5, 0, z, 0.5])
        mats.append(0xC0000000)
    sd = scenes.SceneData(tris=np.array(tris, np.float32), materials=np.array(mats, np.uint32),
                          camera_pos=(0.0, 0.0, 0.0), camera_q=(1.0, 0.0, 0.0, 0.0))
    return sd


def _camera_dir_f64(x, y, screen_distance, q):
    v = np.array([x, y, -screen_distance], np.float64)
    v /= np.linalg.norm(v)
    a, b, c, d = q
    r = np.array([[1 - 2 * (c * c + d * d), 2 * (b * c - a * d), 2 * (b * d + a * c)],
                  [2 * (b * c + a * d), 1 - 2 * (b * b + d * d), 2 * (c * d - a * b)],
                  [2 * (b * d - a * c), 2 * (c * d + a * b), 1 - 2 * (b * b + c * c)]])
    return r @ v


def test_sky_radiance_closed_form(oracle):
    """PIXEL mode end to end: AA offsets from the pixel stream, the camera ray
    and sky_intensity, against FP64 evaluations."""
    import json
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    from make_rng_vectors import pixel_draws, unit  # independent restatement

    sd = _sky_scene()
    s = oracle.Scene.from_scene_data(sd)
    W, H = 64, 32
    hdr, gb = s.accumulate_frames(W, H, 1, 1, oracle.RNG_PIXEL, 2)
    rows = oracle.hdr_blocks_to_rows(W, H, hdr)
    scale = 2.0 / W
    worst = 0.0
    for y in range(0, H, 3):
        for x in range(0, W, 5):
            u = pixel_draws(y * W + x, 1, 2)
            xs = scale * (x - W * 0.5) + unit(u[0]) * scale
            ys = scale * (y - H * 0.5) + unit(u[1]) * scale
            d = _camera_dir_f64(xs, ys, sd.screen_distance, (1, 0, 0, 0))
            dd = 0.5 * d[2] + 0.5
            exp = np.array([dd, dd * dd, dd ** 3]) * 2 + 0.5
            worst = max(worst, float(np.abs(rows[y, x] - exp).max() / exp.max()))
    assert worst < 1e-5
    assert (gb == 0).all()  # camera rays end on the sky: no first-bounce gbuffer


def test_render_requires_eight_lights(oracle):
    sd = _sky_scene()
    sd = scenes.SceneData(tris=sd.tris[:7], materials=sd.materials[:7])
    s = oracle.Scene.from_scene_data(sd)
    bm = np.zeros(32 * 32, np.uint32)
    assert oracle.lib().or_render_patch_u8(s._h, 32, 32, oracle.ptr(bm), None, 32, 0, 0, 1,
                                           oracle.RNG_REF_PATCH) == -12


def test_render_patch_preconditions(oracle):
    s = oracle.Scene.from_scene_data(scenes.config0(subdivisions=2))
    bm = np.zeros(64 * 64, np.uint32)
    L = oracle.lib()
    assert L.or_render_patch_u8(s._h, 64, 64, oracle.ptr(bm), None, 24, 0, 0, 1, 1) == -10
    assert L.or_render_patch_u8(s._h, 60, 64, oracle.ptr(bm), None, 16, 0, 0, 1, 1) == -10
    assert L.or_render_patch_u8(s._h, 64, 64, oracle.ptr(bm), None, 32, 48, 0, 1, 1) == -11


@pytest.mark.parametrize("mode", [0, 1])
def test_patch_equals_frame(oracle, mode):
    """render_patch_u8 of each 32x32 patch equals the frame render (the patch
    loop of main.rs:174-214)."""
    sd = scenes.config0(subdivisions=3)
    s = oracle.Scene.from_scene_data(sd)
    W, H = 64, 64
    frame_bm, frame_gb = s.render_frame_u8(W, H, 3, mode, 4)
    bm = np.zeros(W * H, np.uint32)
    gb = np.zeros(W * H, np.uint32)
    for x in (0, 32):
        for y in (0, 32):
            assert s.render_patch_u8(W, H, bm, gb, 32, x, y, 3, mode) == 0
    assert (bm == frame_bm).all() and (gb == frame_gb).all()


def test_accumulation_is_a_sum_of_frames(oracle):
    sd = scenes.config0(subdivisions=3)
    s = oracle.Scene.from_scene_data(sd)
    W, H = 32, 32
    a, _ = s.accumulate_frames(W, H, 1, 3, oracle.RNG_PIXEL, 4)
    f = [s.accumulate_frames(W, H, k, 1, oracle.RNG_PIXEL, 4)[0] for k in (1, 2, 3)]
    assert a.tobytes() == ((f[0] + f[1]) + f[2]).tobytes()


def test_pixel_and_patch_streams_estimate_the_same_image(oracle):
    """Both RNG modes are unbiased estimators of the same image: their means
    agree statistically (they share no random numbers)."""
    sd = scenes.config0(subdivisions=3)
    s = oracle.Scene.from_scene_data(sd, approx=oracle.APPROX_HOST_X86)
    W, H = 128, 64
    a, _ = s.accumulate_frames(W, H, 1, 24, oracle.RNG_PIXEL, 8)
    b, _ = s.accumulate_frames(W, H, 1, 24, oracle.RNG_REF_PATCH, 8)
    ma, mb = float(a.mean()), float(b.mean())
    assert abs(ma - mb) < 0.04 * max(ma, mb)


def test_gbuffer_texture_index(oracle, models_dir):
    """On the reference's own scene the first-bounce gbuffer carries the
    texture index of the floor (1) and wood (2) materials (renderer.rs:260-286)."""
    from helpers import indoor_scene
    sd = indoor_scene(models_dir)
    s = oracle.Scene.from_scene_data(sd)
    bm, gb = s.render_frame_u8(128, 64, 1, oracle.RNG_REF_PATCH, 4)
    tex = gb >> 24
    assert set(np.unique(tex).tolist()) <= {0, 1, 2}
    assert (tex == 1).any() and (tex == 2).any()
    assert (bm != 0).mean() > 0.3


def test_approx_modes_close(oracle):
    """HOST_X86 (rcpps/rsqrtps) and PORTABLE estimate the same image; the
    image RMSE between them is reported as the cost of the portable mode."""
    sd = scenes.config0(subdivisions=3)
    W, H = 96, 64
    a = oracle.Scene.from_scene_data(sd, approx=oracle.APPROX_PORTABLE)
    b = oracle.Scene.from_scene_data(sd, approx=oracle.APPROX_HOST_X86)
    ha, _ = a.accumulate_frames(W, H, 1, 16, oracle.RNG_PIXEL, 8)
    hb, _ = b.accumulate_frames(W, H, 1, 16, oracle.RNG_PIXEL, 8)
    assert abs(float(ha.mean()) - float(hb.mean())) < 0.05 * float(ha.mean())
