"""Pins the oracle's shading (material.rs, scene.rs) with checks that do not
restate the oracle's own formulas in its own precision:

* closed forms of textbook functions in FP64 (Blinn-Phong lobes with the
  (n+1)/2pi normalisation that material.rs:490 states, Schlick's
  approximation, the cosine pdf);
* normalisation invariants (each lobe integrates to 1 over the hemisphere;
  the direct-sampling pdf integrates to 1 over the lights' solid angle);
* the geometry of the light sampler (every sample on the light that its top
  three bits name, area-uniform first and second moments);
* an FP64 lane-by-lane restatement of continue_path;
* the one-sample MIS estimator: its Monte Carlo mean over seeded draws equals
  an FP64 quadrature of the same expectation, and stays within the
  documented bias of the rendering-equation integral.

A misreading of a constant (lobe factor, Schlick exponent, MIS choice
probability, the 1/8 light pick, the u + v fold) fails one of them."""
from __future__ import annotations

import math

import numpy as np
import pytest

from convector_amd import scenes

GLOSS_EXP = [0, 1, 2, 4, 8, 16]  # material.rs:23-25,490-512


def _unit(v):
    v = np.asarray(v, dtype=np.float64)
    with np.errstate(invalid="ignore", divide="ignore"):  # zero vectors of stopped lanes
        return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _f32unit(v):
    return _unit(v).astype(np.float32)


def _unit_f32(d):
    """sample_unit of raw draws as float64 (random.rs:104-116): the u32 is
    reinterpreted as i32, converted to f32, scaled by -2^-32 and offset by
    0.5 with one rounding (fma)."""
    i = np.asarray(d, dtype=np.uint32).view(np.int32).astype(np.float32).astype(np.float64)
    return (i * -(2.0 ** -32) + 0.5).astype(np.float32).astype(np.float64)


def _biunit_f32(d):
    i = np.asarray(d, dtype=np.uint32).view(np.int32).astype(np.float32).astype(np.float64)
    return i * -(2.0 ** -31)


# ---------------------------------------------------------------- lobes


@pytest.mark.parametrize("gloss", range(6))
def test_normal_dist_is_the_normalised_blinn_phong_lobe(oracle, gloss):
    n = np.array([0.0, 0.0, 1.0], np.float32)
    e = GLOSS_EXP[gloss]
    for c in [1.0, 0.96875, 0.75, 0.5, 0.25, 0.0, -0.25, -0.5]:
        c32 = np.float32(c)
        h = np.array([math.sqrt(1.0 - c * c), 0.0, c32], np.float32)
        got = oracle.microfacet_normal_dist(h, n, gloss)
        cc = float(c32)
        if gloss == 1:
            cc = max(cc, 0.0)  # the only odd power is clamped (material.rs:497)
        want = (e + 1) / (2 * math.pi) * cc ** e
        assert got == pytest.approx(want, rel=2e-6, abs=1e-12), (gloss, c)


@pytest.mark.parametrize("gloss", range(6))
def test_normal_dist_integrates_to_one_over_the_hemisphere(oracle, gloss):
    # int D(h.n) d(omega) = 2 pi int_0^{pi/2} D(cos t) sin t dt = 1 for the
    # (n+1)/2pi normalisation
    m = 4000
    th = (np.arange(m) + 0.5) * (math.pi / 2) / m
    n = np.array([0.0, 0.0, 1.0], np.float32)
    vals = np.array([oracle.microfacet_normal_dist(
        np.array([math.sin(t), 0.0, math.cos(t)], np.float32), n, gloss) for t in th])
    total = 2 * math.pi * np.sum(vals * np.sin(th)) * (math.pi / 2) / m
    assert total == pytest.approx(1.0, abs=2e-4)


def test_fresnel_is_schlick(oracle):
    color = np.array([0.2, 0.5, 0.9], np.float32)
    incoming = np.array([0.0, 0.0, 1.0], np.float32)
    for c in [1.0, 0.875, 0.5, 0.25, 0.0, -0.5, -1.0]:
        h = np.array([math.sqrt(1.0 - c * c), 0.0, c], np.float32)
        out, raw = oracle.microfacet_fresnel(incoming, h, color)
        ct5 = (1.0 - abs(float(h[2]))) ** 5
        want = color.astype(np.float64) + (1.0 - color.astype(np.float64)) * ct5
        assert raw == pytest.approx(ct5, rel=1e-6, abs=1e-9), c
        np.testing.assert_allclose(out, want, rtol=1e-6, atol=1e-9)
    # normal incidence gives the material colour, grazing incidence white
    out, raw = oracle.microfacet_fresnel(incoming, incoming, color)
    assert raw == 0.0 and (out == color).all()
    out, raw = oracle.microfacet_fresnel(incoming, np.array([1, 0, 0], np.float32), color)
    assert raw == 1.0
    np.testing.assert_allclose(out, [1, 1, 1], atol=1e-7)


def _brdf64(color, gloss, d_in, d_out, normal, white_fresnel=False):
    h = _unit(np.asarray(d_in, np.float64) - np.asarray(d_out, np.float64))
    c = float(h @ np.asarray(normal, np.float64))
    e = GLOSS_EXP[gloss]
    if gloss == 1:
        c = max(c, 0.0)
    dist = (e + 1) / (2 * math.pi) * c ** e
    ct5 = (1.0 - abs(float(h @ np.asarray(d_in, np.float64)))) ** 5
    f = np.ones(3) if white_fresnel else color + (1 - color) * ct5
    return f * dist, ct5


def test_brdf_decodes_colour_gloss_and_texture(oracle):
    rng = np.random.default_rng(5)
    for k in range(60):
        rgb = rng.uniform(0.05, 1.0, 3)
        gloss = k % 6
        tex = (k // 6) % 4
        mat = scenes.mat_with_glossiness(scenes.mat_diffuse(*rgb), gloss) | (tex << 24)
        color = np.array([(mat >> s) & 0xFF for s in (0, 8, 16)], np.float64) / 255.0
        normal = _f32unit(rng.normal(size=3))
        d_in = _f32unit(normal + rng.normal(size=3) * 0.5)  # new direction, mostly above
        d_out = _f32unit(-normal + rng.normal(size=3) * 0.5)  # incoming ray direction
        for ignore in (False, True):
            got, raw = oracle.microfacet_brdf(mat, d_in, d_out, normal, ignore_fresnel=ignore)
            want, ct5 = _brdf64(color, gloss, d_in, d_out, normal,
                                white_fresnel=ignore and tex != 0)
            np.testing.assert_allclose(got, want, rtol=4e-5, atol=1e-9, err_msg=str((k, ignore)))
            assert raw == pytest.approx(ct5, rel=1e-4, abs=1e-7)


def test_pd_brdf_is_the_cosine_pdf(oracle):
    rng = np.random.default_rng(6)
    for _ in range(200):
        n = _f32unit(rng.normal(size=3))
        d = _f32unit(rng.normal(size=3))
        want = max(0.0, float(n.astype(np.float64) @ d.astype(np.float64))) / math.pi
        assert oracle.pd_brdf(n, d) == pytest.approx(want, rel=1e-6, abs=1e-7)  # FP32 dot


def test_sky_intensity_closed_form(oracle):
    rng = np.random.default_rng(7)
    for _ in range(100):
        d = _f32unit(rng.normal(size=3))
        x = 0.5 * float(d[2]) + 0.5
        want = 2.0 * np.array([x, x * x, x ** 3]) + 0.5
        np.testing.assert_allclose(oracle.sky_intensity(d), want, rtol=1e-6)


# ------------------------------------------------------- direct sampling


@pytest.fixture(scope="module")
def box(oracle):
    sd = scenes.config0(subdivisions=2)
    s = oracle.Scene.from_scene_data(sd)
    nodes, order = s.bvh()
    mats = sd.materials[order]
    direct = np.nonzero(mats & 0x40000000)[0]  # BVH positions, scene.rs:117-122
    assert len(direct) == 8
    lights = sd.tris[order[direct]].astype(np.float64).reshape(8, 3, 3)
    return s, sd, lights


def _light_geometry(lights):
    v0, v1, v2 = lights[:, 0], lights[:, 1], lights[:, 2]
    e1, e2 = v0 - v2, v1 - v0
    nd = np.cross(e1, e2)
    area = 0.5 * np.linalg.norm(nd, axis=1)
    return v0, v1, v2, e1, e2, nd, area


def _ray_light_hits(o, d, lights):
    """FP64 Moller-Trumbore of rays (o, d) against the 8 lights: t, and the
    barycentric margin (distance of (u, v, 1-u-v) from the boundary)."""
    v0, v1, v2 = lights[:, 0], lights[:, 1], lights[:, 2]
    a, b = v1 - v0, v2 - v0
    p = np.cross(d[:, None, :], b[None])
    det = np.einsum("lk,nlk->nl", a, p)
    inv = 1.0 / det
    s = o[:, None, :] - v0[None]
    u = np.einsum("nlk,nlk->nl", s, p) * inv
    q = np.cross(s, a[None])
    v = np.einsum("nk,nlk->nl", d, q) * inv
    t = np.einsum("lk,nlk->nl", b, q) * inv
    margin = np.minimum(np.minimum(u, v), 1.0 - u - v)
    return t, margin


def _pd_direct64(o, d, lights):
    _, _, _, _, _, nd, area = _light_geometry(lights)
    t, margin = _ray_light_hits(o, d, lights)
    nhat = nd / np.linalg.norm(nd, axis=1, keepdims=True)
    cos = np.abs(np.einsum("lk,nk->nl", nhat, d))
    pd = t * t / (area[None] * (cos + 1e-4))
    hit = (margin > 0) & (t > 0)
    return np.where(hit, pd, 0.0).sum(1) / 8.0, margin, t


def test_pd_direct_matches_fp64_solid_angle_pdf(box):
    s, _, lights = box
    rng = np.random.default_rng(8)
    n = 4000
    o = rng.uniform(-0.9, 0.9, (n, 3)).astype(np.float32)
    # half toward random points of the light plane, half anywhere
    tgt = np.stack([rng.uniform(-0.35, 0.35, n), rng.uniform(-0.35, 0.35, n),
                    np.full(n, -1.0)], 1)
    d = _unit(tgt - o)
    d[n // 2:] = _unit(rng.normal(size=(n - n // 2, 3)))
    d = d.astype(np.float32)
    got = s.pd_direct(np.concatenate([o, d], 1))
    want, margin, t = _pd_direct64(o.astype(np.float64), d.astype(np.float64), lights)
    # decisive rays: no light within 1e-4 (barycentric) of its boundary
    robust = np.all(np.abs(margin) > 1e-4, axis=1) & np.all(np.abs(t) > 1e-4, axis=1)
    assert robust.sum() > 0.95 * n
    hit = want > 0
    assert hit[robust].sum() > n // 4 and (~hit[robust]).sum() > n // 4
    np.testing.assert_array_equal(got[robust & ~hit], 0.0)
    np.testing.assert_allclose(got[robust & hit], want[robust & hit], rtol=2e-5)


def test_pd_direct_integrates_to_one_over_the_lights(box):
    """int pd d(omega) over the lights = sum_k A_k E_{p ~ U(T_k)}[pd cos_k / r^2]
    = 1 for a normalised pdf (less the 1e-4 added to |cos|)."""
    s, _, lights = box
    v0, v1, v2, _, _, nd, area = _light_geometry(lights)
    nhat = nd / np.linalg.norm(nd, axis=1, keepdims=True)
    rng = np.random.default_rng(9)
    m = 2000
    for o in ([0.1, 0.2, 0.3], [-0.6, 0.5, -0.5], [0.0, 0.0, 0.9]):
        o = np.array(o, np.float32)
        total = 0.0
        ideal = 0.0
        for k in range(8):
            u, v = rng.random(m), rng.random(m)
            f = u + v > 1
            u[f], v[f] = 1 - u[f], 1 - v[f]
            p = v0[k] + u[:, None] * (v2[k] - v0[k]) + v[:, None] * (v1[k] - v0[k])
            rvec = p - o.astype(np.float64)
            r = np.linalg.norm(rvec, axis=1)
            d = (rvec / r[:, None]).astype(np.float32)
            pd = s.pd_direct(np.concatenate([np.repeat(o[None], m, 0), d], 1))
            cos = np.abs(d.astype(np.float64) @ nhat[k])
            w = pd * area[k] * cos / (r * r)
            _, margin, _ = _pd_direct64(np.repeat(o[None].astype(np.float64), m, 0),
                                        d.astype(np.float64), lights)
            inside = np.abs(margin[:, k]) > 1e-5
            # per sample: only light k lies along the ray (coplanar lights)
            np.testing.assert_allclose(w[inside], (cos / (cos + 1e-4))[inside] / 8.0,
                                       rtol=5e-5)
            total += w.mean()
            ideal += (cos / (cos + 1e-4)).mean() / 8.0
        assert total == pytest.approx(ideal, rel=2e-4)
        assert total == pytest.approx(1.0, abs=3e-3)


def test_direct_sample_lands_on_the_light_its_top_bits_name(box):
    s, _, lights = box
    v0, v1, v2, _, _, nd, area = _light_geometry(lights)
    rng = np.random.default_rng(10)
    n = 20000
    draws = rng.integers(0, 2 ** 32, (n, 3), dtype=np.uint64).astype(np.uint32)
    pos, nrm, ar = s.direct_sample(draws)
    k = draws[:, 0] >> 29
    # barycentrics of the sample in light k (FP64 least squares)
    a = v1[k] - v0[k]
    b = v2[k] - v0[k]
    p = pos.astype(np.float64) - v0[k]
    aa, ab, bb = (a * a).sum(1), (a * b).sum(1), (b * b).sum(1)
    pa, pb = (p * a).sum(1), (p * b).sum(1)
    den = aa * bb - ab * ab
    beta = (bb * pa - ab * pb) / den
    gamma = (aa * pb - ab * pa) / den
    resid = np.linalg.norm(p - beta[:, None] * a - gamma[:, None] * b, axis=1)
    assert resid.max() < 1e-6
    assert beta.min() > -1e-6 and gamma.min() > -1e-6 and (beta + gamma).max() < 1 + 1e-6
    # the exact point: p = v0 + u (v2 - v0) + v (v1 - v0), (u, v) folded when u + v > 1
    u, v = _unit_f32(draws[:, 1]), _unit_f32(draws[:, 2])
    f = (1.0 - u) - v < 0
    u[f], v[f] = 1 - u[f], 1 - v[f]
    want = v0[k] + u[:, None] * (v2[k] - v0[k]) + v[:, None] * (v1[k] - v0[k])
    np.testing.assert_allclose(pos, want, atol=2e-7)
    # the normal and area it computes (unused by continue_path)
    nhat = nd / np.linalg.norm(nd, axis=1, keepdims=True)
    np.testing.assert_allclose(nrm, nhat[k], atol=1e-6)
    np.testing.assert_allclose(ar, area[k], rtol=1e-6)
    # the index is uniform over the 8 lights
    counts = np.bincount(k, minlength=8)
    assert (np.abs(counts - n / 8) < 5 * math.sqrt(n / 8)).all()


def test_direct_sample_is_area_uniform(box):
    s, _, lights = box
    rng = np.random.default_rng(11)
    n = 1 << 17
    for k in (0, 3, 7):
        draws = rng.integers(0, 2 ** 32, (n, 3), dtype=np.uint64).astype(np.uint32)
        draws[:, 0] = (draws[:, 0] & 0x1FFFFFFF) | (k << 29)
        pos, _, _ = s.direct_sample(draws)
        p = pos.astype(np.float64)
        a, b, c = lights[k]
        S = a + b + c
        mean = S / 3
        m2 = (np.outer(a, a) + np.outer(b, b) + np.outer(c, c) + np.outer(S, S)) / 12
        sd = p.std(0) + 1e-12
        assert (np.abs(p.mean(0) - mean) < 5 * sd / math.sqrt(n) + 1e-7).all()
        cov = m2 - np.outer(mean, mean)
        emp = np.cov(p.T, bias=True)
        scale = np.sqrt(np.outer(np.diag(cov), np.diag(cov))) + 1e-12
        assert (np.abs(emp - cov) < 6 * scale / math.sqrt(n) + 1e-9).all(), (emp, cov)


# ------------------------------------------------------------ continue_path


def _cases(rng, n, lights, emissive_share=0.05):
    """Random shading lanes inside the box: (ray7, isect7, draws6, facts)."""
    pos = rng.uniform(-0.9, 0.9, (n, 3))
    normal = _unit(rng.normal(size=(n, 3)))
    d_in = _unit(-normal + 0.8 * rng.normal(size=(n, 3)))
    rgb = rng.uniform(0.05, 1.0, (n, 3))
    gloss = rng.integers(0, 6, n)
    tex = rng.integers(0, 4, n)
    mats = np.array([scenes.mat_with_glossiness(scenes.mat_diffuse(*c), int(g)) | (int(t) << 24)
                     for c, g, t in zip(rgb, gloss, tex)], np.uint32)
    em = rng.random(n) < emissive_share
    mats[em] |= 0x80000000
    ray7 = np.zeros((n, 7), np.float32)
    ray7[:, :3] = (pos - d_in).astype(np.float32)
    ray7[:, 3:6] = d_in.astype(np.float32)
    inactive = rng.random(n) < 0.03
    ray7[:, 6] = np.where(inactive, np.float32(-0.0), np.float32(0.0))
    isect7 = np.zeros((n, 7), np.float32)
    isect7[:, :3] = pos.astype(np.float32)
    isect7[:, 3:6] = normal.astype(np.float32)
    isect7[:, 6] = mats.view(np.float32)
    draws = rng.integers(0, 2 ** 32, (n, 6), dtype=np.uint64).astype(np.uint32)
    return ray7, isect7, draws, mats, inactive | em


def _light_point64(lights, draws):
    k = draws[:, 2] >> 29
    u, v = _unit_f32(draws[:, 3]), _unit_f32(draws[:, 4])
    f = (1.0 - u) - v < 0
    u[f], v[f] = 1 - u[f], 1 - v[f]
    v0, v1, v2 = lights[k, 0], lights[k, 1], lights[k, 2]
    return v0 + u[:, None] * (v2 - v0) + v[:, None] * (v1 - v0)


@pytest.mark.parametrize("ignore_fresnel", [False, True])
def test_continue_path_lane_by_lane_against_fp64(box, ignore_fresnel):
    s, _, lights = box
    rng = np.random.default_rng(12 + ignore_fresnel)
    n = 6000
    ray7, isect7, draws, mats, dead = _cases(rng, n, lights)
    out, cm, fr = s.continue_path(ray7, isect7, draws, ignore_fresnel=ignore_fresnel)
    # lanes that stop: an emissive hit or an inactive ray keep their ray and get white
    np.testing.assert_array_equal(out[dead, :6], ray7[dead, :6])
    np.testing.assert_array_equal(cm[dead], 1.0)
    act_bits = out[:, 6].view(np.uint32) >> 31
    np.testing.assert_array_equal(act_bits, dead.astype(np.uint32))
    live = ~dead
    pos = isect7[:, :3].astype(np.float64)
    nrm = isect7[:, 3:6].astype(np.float64)
    d_old = ray7[:, 3:6].astype(np.float64)
    d_new = out[:, 3:6].astype(np.float64)
    o_new = out[:, :3].astype(np.float64)
    # branch: the light sample direction, or a cosine-hemisphere direction
    p_light = _light_point64(lights, draws)
    d_light = _unit(p_light - pos)
    is_direct = np.linalg.norm(d_new - d_light, axis=1) < 1e-5
    rr = nrm[:, 2].astype(np.float32).astype(np.float64) * float(np.float32(0.8)) + \
        _biunit_f32(draws[:, 5])
    decisive = np.abs(rr) > 1e-6
    np.testing.assert_array_equal(is_direct[live & decisive], rr[live & decisive] < 0)
    brdf = live & ~is_direct
    # cosine hemisphere: n.d = sqrt(1 - unit), |d| = 1 within the reference's bound
    z = np.sqrt(1.0 - _unit_f32(draws[:, 1]))
    np.testing.assert_allclose((nrm * d_new).sum(1)[brdf], z[brdf], atol=2e-5)
    norm = np.linalg.norm(d_new[brdf], axis=1)
    assert norm.min() > 0.991 and norm.max() < 1.009
    np.testing.assert_allclose(o_new[live], pos[live] + 1e-4 * d_new[live], atol=2e-7)
    # colour modulation (material.rs:357-399) from the chosen ray in FP64
    pd_b = np.maximum((nrm * d_new).sum(1), 0.0) / math.pi
    pd_d, margin, _ = _pd_direct64(o_new, d_new, lights)
    nz = nrm[:, 2]
    p_choice = np.where(is_direct, 0.5 - 0.4 * nz, 0.5 + 0.4 * nz)
    modulation = 1.0 / ((pd_b + pd_d + 0.01) * p_choice)
    cos = np.maximum((nrm * d_new).sum(1), 0.0)
    color = np.stack([(mats >> s_) & 0xFF for s_ in (0, 8, 16)], 1).astype(np.float64) / 255.0
    gloss = (mats >> 26) & 7
    tex = (mats >> 24) & 3
    h = _unit(d_new - d_old)
    c1 = (h * nrm).sum(1)
    e = np.array(GLOSS_EXP, np.float64)[np.minimum(gloss, 5)]
    c1c = np.where(gloss == 1, np.maximum(c1, 0.0), c1)
    dist = (e + 1) / (2 * math.pi) * c1c ** e
    ct5 = (1.0 - np.abs((h * d_new).sum(1))) ** 5
    f = color + (1 - color) * ct5[:, None]
    if ignore_fresnel:
        f[tex != 0] = 1.0
    want = np.minimum(f * (dist * modulation * cos)[:, None], 2.0)
    robust = live & np.all(np.abs(margin) > 1e-4, axis=1)
    assert robust.sum() > 0.9 * live.sum()
    np.testing.assert_allclose(cm[robust], want[robust], rtol=3e-4, atol=1e-7)
    np.testing.assert_allclose(fr[live], ct5[live], rtol=1e-3, atol=2e-6)
    # the clamp is reached, and exactly at 2.0
    assert (cm[live] == 2.0).any() and cm[live].max() == 2.0


def test_continue_path_picks_brdf_with_probability_half_plus_04_nz(box):
    s, _, lights = box
    rng = np.random.default_rng(14)
    n = 100000
    for nz in (-0.9, -0.3, 0.6):
        normal = np.array([math.sqrt(1 - nz * nz), 0.0, nz])
        ray7 = np.zeros((n, 7), np.float32)
        ray7[:, :3] = [0.1, 0.1, 0.5]
        ray7[:, 3:6] = _f32unit(-normal + [0.1, 0.2, 0.0])
        isect7 = np.zeros((n, 7), np.float32)
        isect7[:, :3] = [0.1, 0.1, 0.2]
        isect7[:, 3:6] = normal.astype(np.float32)
        isect7[:, 6] = np.uint32(scenes.mat_diffuse(0.5, 0.5, 0.5)).view(np.float32)
        draws = rng.integers(0, 2 ** 32, (n, 6), dtype=np.uint64).astype(np.uint32)
        out, _, _ = s.continue_path(ray7, isect7, draws)
        p_light = _light_point64(lights, draws)
        d_light = _unit(p_light - isect7[:, :3].astype(np.float64))
        direct = np.linalg.norm(out[:, 3:6] - d_light, axis=1) < 1e-5
        p_b = 0.5 + 0.4 * nz
        sigma = math.sqrt(p_b * (1 - p_b) / n)
        assert abs((~direct).mean() - p_b) < 5 * sigma, (nz, (~direct).mean(), p_b)


# ------------------------------------------------ one-sample MIS estimator


def _light_only_scene(oracle):
    """The 2x2 light grid of the box alone (8 direct-sample triangles)."""
    sd = scenes.config0(subdivisions=1)
    light = (sd.materials & 0xC0000000) == 0xC0000000
    s = oracle.Scene(sd.tris[light], sd.materials[light])
    nodes, order = s.bvh()
    lights = sd.tris[light][order].astype(np.float64).reshape(8, 3, 3)
    return s, lights


def _hemisphere_basis(oracle, normal32):
    """Columns R e_x, R e_y, R e_z of rotate_hemisphere (vector3.rs:221-251)
    from the oracle's own (known-answer pinned) implementation."""
    ex = np.eye(3, dtype=np.float32)
    out = np.zeros((3, 3), np.float32)
    oracle.lib().or_rotate_hemisphere(oracle.ptr(ex), oracle.ptr(np.repeat(normal32[None], 3, 0)),
                                      3, oracle.APPROX_PORTABLE, oracle.ptr(out))
    return out.astype(np.float64).T  # column j = R e_j


def _sin_poly64(x):
    # simd.rs:216-233 (the coefficients as f32)
    a, b, c = (float(np.float32(v)) for v in (0.982006669405, -0.152175915294,
                                              0.00533738081738135917))
    return x * a + x ** 3 * b + x ** 5 * c


def _cos_poly64(x):
    # simd.rs:190-204
    a, b, c = (float(np.float32(v)) for v in (-0.496000299455, 0.0392596924214,
                                              -0.000966231179636657107))
    return 1 + x ** 2 * a + x ** 4 * b + x ** 6 * c


@pytest.mark.parametrize("gloss,rgb", [(0, (0.3, 0.25, 0.2)), (3, (0.15, 0.2, 0.25))])
def test_one_sample_mis_estimator_matches_quadrature(oracle, gloss, rgb):
    """E[color_mod * L * hit] over seeded draws (Monte Carlo through the
    oracle) against FP64 quadratures:
      (a) the same expectation in sample space -- the BRDF branch over the
          (angle, unit) square through the reference's polynomial sin/cos,
          the light branch over the lights' area -- to 5 sigma: the branch
          choice probability must cancel the 1/p compensation and the
          samplers must follow the draws;
      (b) the rendering-equation integral int_lights f_r L cos w/(w+0.01)
          d(omega) with w = pd_brdf + pd_direct, within 1.5 %: the MIS
          weights are the samplers' densities (the polynomial azimuth makes
          the BRDF sampler only approximately cosine-distributed)."""
    s, lights = _light_only_scene(oracle)
    pos = np.array([0.1, -0.05, -0.55], np.float32)
    normal = _f32unit([-0.2, 0.1, -0.95])
    d_in = _f32unit([0.3, -0.2, 0.9])
    mat = scenes.mat_with_glossiness(scenes.mat_diffuse(*rgb), gloss)
    color = np.array([(mat >> s_) & 0xFF for s_ in (0, 8, 16)], np.float64) / 255.0
    e = GLOSS_EXP[gloss]
    n64, din64, pos64 = normal.astype(np.float64), d_in.astype(np.float64), pos.astype(np.float64)

    def g_over_w(d, origin):
        """f_r cos * hit / (pd_b + pd_d + 0.01) for unit-ish directions d."""
        pd_b = np.maximum(d @ n64, 0.0) / math.pi
        pd_d, margin, t = _pd_direct64(origin, d, lights)
        hit = np.any((margin > 0) & (t > 0), axis=1)
        h = _unit(d - din64)
        c1 = h @ n64
        c1 = np.maximum(c1, 0.0) if gloss == 1 else c1
        dist = (e + 1) / (2 * math.pi) * c1 ** e
        ct5 = (1.0 - np.abs((h * d).sum(1))) ** 5
        f = color + (1 - color) * ct5[:, None]
        cos = np.maximum(d @ n64, 0.0)
        return f * (dist * cos * hit / (pd_b + pd_d + 0.01))[:, None], pd_b + pd_d

    # Monte Carlo through the oracle
    rng = np.random.default_rng(15 + gloss)
    n = 1 << 20
    ray7 = np.zeros((n, 7), np.float32)
    ray7[:, :3] = pos - d_in
    ray7[:, 3:6] = d_in
    isect7 = np.zeros((n, 7), np.float32)
    isect7[:, :3] = pos
    isect7[:, 3:6] = normal
    isect7[:, 6] = np.uint32(mat).view(np.float32)
    draws = rng.integers(0, 2 ** 32, (n, 6), dtype=np.uint64).astype(np.uint32)
    out, cm, _ = s.continue_path(ray7, isect7, draws)
    hit = np.zeros(n, bool)
    for c in np.array_split(np.arange(n), 16):
        _, margin, t = _pd_direct64(out[c, :3].astype(np.float64),
                                    out[c, 3:6].astype(np.float64), lights)
        hit[c] = np.any((margin > 0) & (t > 0), axis=1)
    # the clamp never acts where the ray reaches a light, so the expectation is exact
    assert cm[hit].max() < 2.0
    contrib = cm.astype(np.float64) * hit[:, None]
    mc, sem = contrib.mean(0), contrib.std(0) / math.sqrt(n)

    # (a) sample-space quadrature: E = E_phi,u[g/w](brdf) + E_lights[g/w](direct)
    R = _hemisphere_basis(oracle, normal)
    m = 1536
    phi = -math.pi + (np.arange(m) + 0.5) * (2 * math.pi / m)
    u = (np.arange(m) + 0.5) / m
    PH, U = np.meshgrid(phi, u, indexing="ij")
    r = np.sqrt(U.ravel())
    v = np.stack([_sin_poly64(PH.ravel()) * r, _cos_poly64(PH.ravel()) * r,
                  np.sqrt(1 - U.ravel())], 1)
    d_b = v @ R.T
    e_b = np.zeros(3)
    for chunk in np.array_split(np.arange(len(d_b)), 16):
        gw, _ = g_over_w(d_b[chunk], pos64 + 1e-4 * d_b[chunk])
        e_b += gw.sum(0)
    e_b /= len(d_b)
    v0, v1, v2, _, _, nd, area = _light_geometry(lights)
    q = 96  # light branch: centroid rule on q^2 sub-triangles per light
    i, j = np.meshgrid(np.arange(q), np.arange(q), indexing="ij")
    keep = i + j < q
    bu = np.concatenate([((i + 1 / 3) / q)[keep], ((i + 2 / 3) / q)[(i + j < q - 1)]])
    bv = np.concatenate([((j + 1 / 3) / q)[keep], ((j + 2 / 3) / q)[(i + j < q - 1)]])
    e_d = np.zeros(3)
    integral = np.zeros(3)
    nhat = nd / np.linalg.norm(nd, axis=1, keepdims=True)
    for k in range(8):
        p = v0[k] + bu[:, None] * (v2[k] - v0[k]) + bv[:, None] * (v1[k] - v0[k])
        rv = p - pos64
        rr = np.linalg.norm(rv, axis=1)
        d = rv / rr[:, None]
        gw, w = g_over_w(d, pos64 + 1e-4 * d)
        e_d += gw.mean(0) / 8.0
        # (b) int over light k of f_r cos w/(w+0.01) d(omega), d(omega) = |cos_L|/r^2 dA
        jac = np.abs(d @ nhat[k]) / (rr * rr) * area[k]
        integral += (gw * (w * jac)[:, None]).mean(0)
    expect = e_b + e_d
    assert np.all(np.abs(mc - expect) < 5 * sem + 1e-4 * expect), (mc, expect, sem)
    np.testing.assert_allclose(expect, integral, rtol=0.015)
