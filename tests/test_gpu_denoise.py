"""The learned denoiser: HIP forward and backward against the FP32 PyTorch
restatement in tests/denoise_ref.py (with BF16 rounding at the same
points), the Adam step, the renderer's feature outputs, and a short
training run showing what the network gains over the plain 1-sample
image."""
import numpy as np
import pytest

from convector_amd import scenes

import denoise_ref as R

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _random_batch(seed, nb=2, h=16, w=64):
    rng = np.random.default_rng(seed)
    inp = rng.normal(0.0, 1.0, (nb, h, w, 10)).astype(np.float32)
    inp[..., 9] = 1.0
    target = np.exp(rng.normal(-1.0, 1.0, (nb, h, w, 3))).astype(np.float32)
    return inp, target


def test_num_params(cv):
    d = cv.Denoiser(1)
    assert d.num_params == R.num_params()


def test_forward_matches_torch(cv):
    d = cv.Denoiser(3)
    p = d.params()
    inp, _ = _random_batch(4)
    out = d.forward(inp)
    import torch
    ref = R.forward(torch.tensor(inp), R.unpack(p), bf16=True).detach().numpy()
    assert np.isfinite(out).all()
    # BF16 activations: an accumulation-order difference can round a hidden
    # value to the neighbouring BF16 number (2^-8 relative), so agreement is
    # to a few 1e-3, not to FP32 precision
    print("forward rel L2 vs bf16-emulating torch", _rel(out, ref), flush=True)
    assert _rel(out, ref) < 1e-2
    # the network is not trivially zero
    assert np.abs(ref).mean() > 1e-3


def test_loss_is_deterministic(cv):
    """The loss is summed from block partials in a fixed order: with lr = 0
    (gradients only, the weights unchanged) the same batch gives the same
    loss bits on every call, and it equals the mean absolute error of the
    forward's output in log space to float precision."""
    d = cv.Denoiser(23)
    inp, tgt = _random_batch(24, nb=2, h=36, w=128)
    losses = [d.train_step(inp, tgt, 0.0) for _ in range(3)]
    assert np.isfinite(losses[0])
    assert losses[0] == losses[1] == losses[2]
    out = d.forward(inp, materialize=True)
    ref = float(np.mean(np.abs(out.astype(np.float64) - np.log1p(tgt.astype(np.float64)))))
    assert abs(losses[0] - ref) <= 1e-5 * max(ref, 1e-30)


def test_device_forward_is_exact(cv):
    """Caller device buffers (the input also at an offset of one float, the
    output written by the last layer in place) give the bits of the
    host-pointer call."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    nb, h, w = 2, 36, 128
    d = cv.Denoiser(17)
    inp, _ = _random_batch(18, nb=nb, h=h, w=w)
    ref = d.forward(inp)
    n_in, n_out = inp.nbytes, nb * h * w * 12
    di, do = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(di), ctypes.c_size_t(n_in + 16)) == 0
    assert hip.hipMalloc(ctypes.byref(do), ctypes.c_size_t(n_out)) == 0
    try:
        for off in (0, 4):
            assert hip.hipMemcpy(ctypes.c_void_p(di.value + off), inp.ctypes.data_as(ctypes.c_void_p),
                                 ctypes.c_size_t(n_in), 1) == 0
            d.forward_device(di.value + off, do.value, nb, h, w)
            out = np.empty_like(ref)
            assert hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), do, ctypes.c_size_t(n_out),
                                 2) == 0
            assert out.tobytes() == ref.tobytes(), off
    finally:
        hip.hipFree(di)
        hip.hipFree(do)


@pytest.mark.parametrize("nb,h,w", [(2, 16, 64), (1, 36, 128), (1, 736, 1280)])
def test_fused_upsampling_forward_is_exact(cv, nb, h, w):
    """The decoder convolutions reading the upsampled halves of their inputs
    from the lower level give the bits of the forward that writes them out
    first (borders, odd tile rows at 36, the configuration's 1280x736)."""
    d = cv.Denoiser(11)
    inp, _ = _random_batch(12, nb=nb, h=h, w=w)
    fused = d.forward(inp)
    ref = d.forward(inp, materialize=True)
    assert np.isfinite(fused).all()
    assert fused.tobytes() == ref.tobytes()


@pytest.mark.parametrize("nb,h,w", [(2, 16, 64), (3, 36, 64), (2, 256, 256)])
def test_fused_training_step_is_exact(cv, nb, h, w):
    """Pooling in the encoder epilogues and the upsampling gradients in the
    decoder's input-gradient convolutions (writing the layer below's dZ)
    give the gradients, loss and Adam step of the separate kernels, bit for
    bit."""
    inp, target = _random_batch(14, nb=nb, h=h, w=w)
    d1, d2 = cv.Denoiser(15), cv.Denoiser(15)
    l1 = d1.train_step(inp, target, 0.0)
    l2 = d2.train_step(inp, target, 0.0, materialize=True)
    assert d1.grads().tobytes() == d2.grads().tobytes()
    assert l1 == l2 or abs(l1 - l2) <= 1e-6 * abs(l2)  # the loss is a float atomic sum
    d1.train_step(inp, target, 1e-3)
    d2.train_step(inp, target, 1e-3, materialize=True)
    assert d1.params().tobytes() == d2.params().tobytes()


def test_gradients_match_torch(cv):
    d = cv.Denoiser(5)
    p = d.params()
    inp, target = _random_batch(6)
    loss = d.train_step(inp, target, 0.0)
    g = d.grads()
    rloss, rg, _ = R.loss_and_grads(p, inp, target, bf16=True)
    assert abs(loss - rloss) <= 2e-3 * abs(rloss)
    for k, (ws, bs) in enumerate(R.layer_slices()):
        gw, rw = g[ws], rg[ws]
        cos = float(np.dot(gw, rw) / (np.linalg.norm(gw) * np.linalg.norm(rw) + 1e-30))
        print("layer", k, "rel", _rel(gw, rw), "cos", cos, flush=True)
        assert cos > 0.999 and _rel(gw, rw) < 5e-2
        assert _rel(g[bs], rg[bs]) < 5e-2
    # lr = 0 leaves the parameters unchanged
    assert (d.params() == p).all()


def test_gradients_split_and_ragged(cv):
    """3 x 36 x 64 pixels: the full-size weight gradients split the pixels
    over 3 blocks, and at quarter size (3 x 9 x 16 = 432 pixels) the last
    64-pixel chunk is partial.  The splits are summed in a fixed order, so a
    second run gives the same gradient bits (the loss is a float atomic sum,
    equal to rounding)."""
    d = cv.Denoiser(9)
    p = d.params()
    inp, target = _random_batch(10, nb=3, h=36, w=64)
    loss = d.train_step(inp, target, 0.0)
    g = d.grads()
    assert abs(d.train_step(inp, target, 0.0) - loss) <= 1e-6 * abs(loss)
    assert d.grads().tobytes() == g.tobytes()
    rloss, rg, _ = R.loss_and_grads(p, inp, target, bf16=True)
    assert abs(loss - rloss) <= 2e-3 * abs(rloss)
    for k, (ws, bs) in enumerate(R.layer_slices()):
        gw, rw = g[ws], rg[ws]
        cos = float(np.dot(gw, rw) / (np.linalg.norm(gw) * np.linalg.norm(rw) + 1e-30))
        print("layer", k, "rel", _rel(gw, rw), "cos", cos, flush=True)
        assert cos > 0.999 and _rel(gw, rw) < 5e-2
        assert _rel(g[bs], rg[bs]) < 5e-2


def test_adam_first_step(cv):
    """Adam's first step moves every weight by lr * g / (|g| + eps)."""
    d = cv.Denoiser(7)
    p0 = d.params()
    inp, target = _random_batch(8)
    lr = 1e-3
    d.train_step(inp, target, lr)
    g = d.grads()
    p1 = d.params()
    expect = p0 - lr * g / (np.abs(g) + 1e-8)
    assert np.allclose(p1, expect, rtol=1e-5, atol=1e-7)


def _render_pair(cv, scene, w, h, frame, ref_spp):
    r = cv.Renderer(scene, w, h)
    hdr, feat = r.new_buffer_f32(), r.new_features()
    r.accumulate_aov_f32(hdr, feat, frame, 1)
    ref = r.new_buffer_f32()
    r.accumulate_frames_f32(ref, 10000, ref_spp)
    inp = cv.Denoiser.prepare_input(hdr, feat, w, h, 1)
    rf = cv.Denoiser.prepare_input(ref, feat, w, h, ref_spp)
    target = np.expm1(rf[..., :3])
    return inp, target, hdr, feat


def test_feature_outputs(cv, oracle):
    sd = scenes.config0(subdivisions=3)
    g = cv.Scene.from_scene_data(sd)
    W, H = 64, 48
    r = cv.Renderer(g, W, H)
    hdr, feat = r.new_buffer_f32(), r.new_features()
    r.accumulate_aov_f32(hdr, feat, 1, 3)
    f = feat.reshape(H, W, 8)
    assert (f[..., 7] == 3).all()
    # a mean of unit normals, albedo in [0, 1] per frame, positive distances
    assert (np.linalg.norm(f[..., 3:6] / 3, axis=-1) <= 1.0 + 1e-5).all()
    assert (f[..., 0:3] >= 0).all() and (f[..., 0:3] <= 3.0 + 1e-6).all()
    assert (f[..., 6] > 0).all()
    one = r.new_features()
    r.accumulate_aov_f32(r.new_buffer_f32(), one, 1, 1)
    n1 = one.reshape(H, W, 8)[..., 3:6]
    assert np.allclose(np.linalg.norm(n1, axis=-1), 1.0, atol=1e-5)
    # the radiance is the plain accumulation's
    ohdr, _ = oracle.Scene.from_scene_data(sd).accumulate_frames(W, H, 1, 3, oracle.RNG_PIXEL, 8)
    assert hdr.tobytes() == ohdr.tobytes()


def test_denoiser_beats_plain_1spp(cv):
    """A short training run on 1-sample / 256-sample pairs of procedural
    scenes; on a held-out scene the denoised 1-sample image has a lower
    RMSE against the 256-sample reference than the plain 1-sample image."""
    W, H = 64, 64
    train = []
    for k in range(6):
        sd = scenes.procedural_scene(4000 + 1000 * k, 100 + k)
        g = cv.Scene.from_scene_data(sd)
        for f in (1, 2):
            inp, target, _, _ = _render_pair(cv, g, W, H, f, 256)
            train.append((inp, target))
    xs = np.stack([t[0] for t in train])
    ys = np.stack([t[1] for t in train])
    d = cv.Denoiser(11)
    rng = np.random.default_rng(0)
    first = last = None
    for it in range(300):
        idx = rng.choice(len(train), 4, replace=False)
        loss = d.train_step(xs[idx], ys[idx], 1e-3)
        first = loss if first is None else first
        last = loss
    print("loss", first, "->", last, flush=True)
    assert last < first
    sd = scenes.procedural_scene(9000, 999)
    g = cv.Scene.from_scene_data(sd)
    inp, target, hdr, feat = _render_pair(cv, g, W, H, 3, 256)
    out = np.expm1(np.maximum(d.forward(inp[None])[0], 0.0))
    noisy = np.expm1(inp[..., :3])
    rmse_dn = float(np.sqrt(((out - target) ** 2).mean()))
    rmse_1 = float(np.sqrt(((noisy - target) ** 2).mean()))
    print("held-out RMSE: plain 1 spp %.4f, denoised %.4f" % (rmse_1, rmse_dn), flush=True)
    assert rmse_dn < rmse_1
