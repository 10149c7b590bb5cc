"""The north-star configurations at their own workloads:

* the many-jobs configuration: the 64 seeded jobs (1e3..2e6 triangles,
  64x64..1920x1080, 1..64 spp) in one cv_batch with captured HIP graphs --
  finite, deterministic, graph replay equal to eager launches, every 8th job
  equal to its own renderer, the two smallest equal to the oracle;
* the denoiser: the 1280 x 736 forward and one training step at batch 16 of
  256 x 256 crops against the FP32 PyTorch restatement;
* serving: 4 resident scenes (20k, 200k, 1.33M and 2M triangles) and
  requests at 256^2, 720p and 1080p with 1, 4 or 16 spp arriving at Poisson
  times -- each result the same alone and in the mix, one request per scene
  equal to the oracle under the request's camera."""
import math
import time

import numpy as np
import pytest

from convector_amd import scenes

import denoise_ref as R

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _oracle_for(cv, oracle, g, sd):
    """The oracle traversing the GPU's tree (itself the oracle's build, bit
    for bit) -- spares the host a multi-second build of large scenes."""
    nodes, order = g.bvh()
    return oracle.Scene.from_scene_data(sd, nodes=nodes.view(oracle.NODE_DTYPE), order=order)


# ------------------------------------------------------------ 64 jobs


@pytest.fixture(scope="module")
def jobs64(cv):
    specs = scenes.batch_jobs(64, 1000)
    sds = [scenes.procedural_scene(n, seed) for (n, seed, _, _, _) in specs]
    gs = [cv.Scene.from_scene_data(sd) for sd in sds]
    return specs, sds, gs


def _render64(cv, specs, gs, flags):
    jobs = [cv.Job(g, w, h, 1, spp, gbuffer=np.zeros(w * h, np.uint32))
            for g, (_, _, w, h, spp) in zip(gs, specs)]
    b = cv.Batch()
    b.render(jobs, flags)
    return jobs, b


def test_config2_64_jobs(cv, oracle, jobs64):
    specs, sds, gs = jobs64
    assert len(specs) == 64 and max(sd.ntris for sd in sds) > 1e6
    paths = sum(w * h * spp for (_, _, w, h, spp) in specs)
    print("64 jobs: %d triangles, %d paths" % (sum(sd.ntris for sd in sds), paths), flush=True)
    graph, b = _render64(cv, specs, gs, cv.RENDER_GRAPH)
    assert b.stats()["paths"] == paths
    replay, _ = _render64(cv, specs, gs, cv.RENDER_GRAPH)
    eager, _ = _render64(cv, specs, gs, 0)
    for a, r, e in zip(graph, replay, eager):
        assert np.isfinite(a.hdr).all()
        assert a.hdr.tobytes() == r.hdr.tobytes() == e.hdr.tobytes()
        assert (a.gbuffer == r.gbuffer).all() and (a.gbuffer == e.gbuffer).all()
    assert sum(float(j.hdr.sum()) for j in graph) > 0
    for k in range(0, 64, 8):
        _, _, w, h, spp = specs[k]
        r = cv.Renderer(gs[k], w, h)
        hdr = r.new_buffer_f32()
        gb = np.zeros(w * h, np.uint32)
        r.accumulate_frames_f32(hdr, 1, spp, gbuffer=gb)
        assert hdr.tobytes() == graph[k].hdr.tobytes(), k
        assert (gb == graph[k].gbuffer).all(), k
    small = sorted(range(64), key=lambda k: specs[k][2] * specs[k][3] * specs[k][4])[:2]
    for k in small:
        _, _, w, h, spp = specs[k]
        print("64 jobs: job %d (%dx%d, %d spp, %d triangles) against the oracle"
              % (k, w, h, spp, sds[k].ntris), flush=True)
        ohdr, ogb = _oracle_for(cv, oracle, gs[k], sds[k]).accumulate_frames(
            w, h, 1, spp, oracle.RNG_PIXEL, 16)
        assert graph[k].hdr.tobytes() == ohdr.tobytes(), k
        assert (graph[k].gbuffer == ogb).all(), k


# ------------------------------------------------------------ denoiser


def test_config3_forward_full_size(cv):
    """The network on one 1280 x 736 frame."""
    import torch
    d = cv.Denoiser(11)
    rng = np.random.default_rng(12)
    inp = rng.normal(0.0, 1.0, (1, 736, 1280, 10)).astype(np.float32)
    inp[..., 9] = 1.0
    out = d.forward(inp)
    ref = R.forward(torch.tensor(inp), R.unpack(d.params()), bf16=True).detach().numpy()
    rel = _rel(out, ref)
    print("1280x736 forward rel L2", rel, flush=True)
    assert np.isfinite(out).all() and rel < 1e-2


def test_config3_training_step_batch16(cv):
    """One training step at the configuration's shape: 16 crops of
    256 x 256, gradients of every layer against torch, then Adam."""
    d = cv.Denoiser(13)
    p = d.params()
    rng = np.random.default_rng(14)
    inp = rng.normal(0.0, 1.0, (16, 256, 256, 10)).astype(np.float32)
    inp[..., 9] = 1.0
    target = np.exp(rng.normal(-1.0, 1.0, (16, 256, 256, 3))).astype(np.float32)
    loss = d.train_step(inp, target, 0.0)
    g = d.grads()
    rloss, rg, _ = R.loss_and_grads(p, inp, target, bf16=True)
    assert abs(loss - rloss) <= 2e-3 * abs(rloss)
    for k, (ws, bs) in enumerate(R.layer_slices()):
        gw, rw = g[ws], rg[ws]
        cos = float(np.dot(gw, rw) / (np.linalg.norm(gw) * np.linalg.norm(rw) + 1e-30))
        print("batch-16 layer", k, "rel", _rel(gw, rw), "cos", cos, flush=True)
        assert cos > 0.999 and _rel(gw, rw) < 5e-2
        assert _rel(g[bs], rg[bs]) < 5e-2
    # Adam's first step from the same parameters (a fresh denoiser of the
    # same seed): every weight moves by lr * g / (|g| + eps)
    d2 = cv.Denoiser(13)
    assert (d2.params() == p).all()
    d2.train_step(inp, target, 1e-3)
    expect = p - 1e-3 * g / (np.abs(g) + 1e-8)
    np.testing.assert_allclose(d2.params(), expect, rtol=1e-5, atol=1e-7)


# ------------------------------------------------------------- serving


def _yaw_pitch_q(yaw, pitch):
    """Camera orientation: the default view (looking down -y) turned by yaw
    about z and pitch about x."""
    def qmul(a, b):
        a1, b1, c1, d1 = a
        a2, b2, c2, d2 = b
        return (a1 * a2 - b1 * b2 - c1 * c2 - d1 * d2, a1 * b2 + b1 * a2 + c1 * d2 - d1 * c2,
                a1 * c2 - b1 * d2 + c1 * a2 + d1 * b2, a1 * d2 + b1 * c2 - c1 * b2 + d1 * a2)
    q0 = (math.sqrt(0.5), -math.sqrt(0.5), 0.0, 0.0)
    qz = (math.cos(yaw / 2), 0.0, 0.0, math.sin(yaw / 2))
    qx = (math.cos(pitch / 2), math.sin(pitch / 2), 0.0, 0.0)
    return tuple(float(np.float32(c)) for c in qmul(qmul(qz, q0), qx))


def test_config4_serving_mix(cv, oracle):
    sds = [scenes.procedural_scene(20000, 4001), scenes.procedural_scene(200000, 4002),
           scenes.config1(), scenes.procedural_scene(2000000, 4004)]
    gs = [cv.Scene.from_scene_data(sd) for sd in sds]
    print("serving scenes:", [sd.ntris for sd in sds], flush=True)
    rng = np.random.default_rng(4005)
    sizes = [(256, 256), (1280, 720), (1920, 1080)]
    reqs = []
    for k in range(20):
        w, h = sizes[int(rng.integers(0, 3))]
        spp = int(rng.choice([1, 4, 16]))
        if w * h * spp > 1920 * 1080 * 4:  # keep the test's GPU time to seconds
            spp = 1
        pos = (float(rng.uniform(-0.6, 0.6)), float(rng.uniform(0.5, 0.9)),
               float(rng.uniform(-0.6, 0.6)))
        q = _yaw_pitch_q(rng.uniform(-0.4, 0.4), rng.uniform(-0.3, 0.3))
        reqs.append((k % 4, w, h, spp, pos, q, bool(rng.random() < 0.25)))
    d = cv.Denoiser(17)
    sv = cv.Server()
    ids = [sv.add_scene(g) for g in gs]
    sv.set_denoiser(d)
    # Poisson arrivals at 200 requests/s while earlier ones are on the GPU
    gaps = rng.exponential(1.0 / 200.0, len(reqs))
    mix = []
    for (si, w, h, spp, pos, q, dn), gap in zip(reqs, gaps):
        time.sleep(gap)
        mix.append(sv.submit(ids[si], w, h, spp, pos, q, denoise=dn))
    lat = [sv.wait(t) for t, _ in mix]
    st = sv.stats()
    print("serving mix: p50 %.1f ms, p99 %.1f ms" % (st["p50_ms"], st["p99_ms"]), flush=True)
    assert st["completed"] == len(reqs) and all(x > 0 for x in lat)
    # alone: a sample of the requests one at a time on a fresh server
    alone_sv = cv.Server()
    aids = [alone_sv.add_scene(g) for g in gs]
    alone_sv.set_denoiser(d)
    for k in range(0, len(reqs), 3):
        si, w, h, spp, pos, q, dn = reqs[k]
        t, hdr = alone_sv.submit(aids[si], w, h, spp, pos, q, denoise=dn)
        alone_sv.wait(t)
        assert hdr.tobytes() == mix[k][1].tobytes(), k
    alone_sv.close()
    sv.close()
    # one plain request per scene against the oracle under its camera
    for si in range(4):
        k = next(i for i, r in enumerate(reqs) if r[0] == si and not r[6])
        _, w, h, spp, pos, q, _ = reqs[k]
        w, h, spp = min(w, 256), min(h, 256), min(spp, 4)
        sv2 = cv.Server()
        t, hdr = sv2.submit(sv2.add_scene(gs[si]), w, h, spp, pos, q)
        sv2.wait(t)
        sv2.close()
        o = _oracle_for(cv, oracle, gs[si], sds[si])
        o.set_camera(pos, (0.0, 0.0, 0.0), q, (0.0, 0.0, 0.0), sds[si].screen_distance)
        ohdr, _ = o.accumulate_frames(w, h, 1, spp, oracle.RNG_PIXEL, 16)
        print("serving: scene %d request %dx%d x %d against the oracle" % (si, w, h, spp),
              flush=True)
        assert hdr.tobytes() == ohdr.tobytes(), si
