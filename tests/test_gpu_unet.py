"""The learned denoiser (pt_unet.h) on the device against an fp32 PyTorch
restatement of the same network (tests/unet_ref.py): forward values and
every parameter gradient, plus training and the frame path.

Tolerances: activations, activation gradients and weights are bf16 on the
device (8-bit mantissa, relative rounding 2^-9 per value) with fp32
accumulation, against fp32 everywhere in the reference, through 19 layers.
"""
import numpy as np
import pytest

import unet_ref as R
from webgpuraytracer_amd import LearnedDenoiser, SceneBuilder

pytestmark = pytest.mark.gpu


def _features(rng, n, h, w):
    """Noisy-render-like features: radiance (heavy-tailed), albedo, unit
    normal, hit distance."""
    f = np.empty((n, h, w, 10), np.float32)
    yy, xx = np.meshgrid(np.linspace(0, 1, h), np.linspace(0, 1, w), indexing="ij")
    for i in range(n):
        base = 0.5 + 0.4 * np.sin(6 * xx + i) * np.cos(5 * yy - i)
        alb = np.stack([base, 0.6 * base + 0.2, 1.0 - 0.5 * base], -1)
        f[i, :, :, 3:6] = alb
        nrm = rng.normal(size=(h, w, 3))
        f[i, :, :, 6:9] = nrm / np.linalg.norm(nrm, axis=-1, keepdims=True)
        f[i, :, :, 9] = 2.0 + 8.0 * base
        f[i, :, :, 0:3] = alb * rng.gamma(1.0, 1.0, size=(h, w, 3))
    return f


def _target(f):
    """A clean image the features predict: albedo times a smooth shading."""
    return (f[..., 3:6] * (0.5 + 0.5 * f[..., 9:10] / 10.0)).astype(np.float32)


@pytest.fixture
def net(renderer):
    return LearnedDenoiser(renderer, seed=3)


def test_param_layout_and_roundtrip(net):
    n = net.param_count()
    assert n == R.param_count()
    p = net.get_params()
    assert p.shape == (n,) and np.isfinite(p).all()
    sl = dict((name, (ws, bs)) for name, ws, bs in R.layer_slices())
    # padded input channels of the first layer and unused outputs of the last are zero
    w0 = p[sl["e0a"][0]].reshape(9, 32, 16)
    assert not w0[:, :, 10:].any() and w0[:, :, :10].std() > 0
    wo = p[sl["out"][0]].reshape(9, 16, 32)
    assert not wo[:, 3:, :].any() and wo[:, :3, :].std() > 0
    # He initialisation: std sqrt(2 / (9 fan_in)) per layer
    w1 = p[sl["e1b"][0]]
    assert abs(w1.std() / np.sqrt(2.0 / (9 * 48)) - 1.0) < 0.05
    q = p.copy()
    q[::7] *= -0.5
    net.set_params(q)
    assert np.array_equal(net.get_params(), q)
    net.set_params(p)


def test_forward_matches_torch_fp32(net):
    rng = np.random.default_rng(11)
    feat = _features(rng, 2, 64, 96)
    pred = net.forward(feat)
    import torch
    with torch.no_grad():
        ref = R.forward(R.unflatten(net.get_params()), R.pack_input(feat)).permute(0, 2, 3, 1).numpy()
    assert pred.shape == ref.shape
    err = np.linalg.norm(pred - ref) / np.linalg.norm(ref)
    print("forward relative L2 error %.3e, max abs %.3e (|ref| max %.3f)" % (err, np.abs(pred - ref).max(),
                                                                             np.abs(ref).max()))
    assert err < 2e-2
    assert np.abs(pred - ref).max() < 0.05 * np.abs(ref).max()


def _target_with_margin(rng, params, feat, margin=0.25):
    """Targets whose log(1 + t) sits at least `margin` above or below the
    reference prediction, so the sign of every L1 residual is the same on the
    device (bf16) and in the reference (fp32)."""
    import torch
    with torch.no_grad():
        ref = R.forward(R.unflatten(params), R.pack_input(feat)).permute(0, 2, 3, 1).numpy()
    below = (rng.random(ref.shape) < 0.5) & (ref - margin - 0.05 > 0.0)
    lt = np.where(below, ref - margin - 0.05 * rng.random(ref.shape), ref + margin + 0.05 * rng.random(ref.shape))
    return np.expm1(lt).astype(np.float32)


def test_gradients_match_torch_fp32(net):
    rng = np.random.default_rng(12)
    _check_gradients(net, rng, _features(rng, 2, 64, 64))


def test_gradients_match_torch_fp32_ragged_levels(net):
    # 48 x 80 gives levels 48x80, 24x40, 12x20, 6x10 and 3x5: the weight
    # gradient's pixel tiles (8, 16 or 32 wide) overhang the image at every level
    rng = np.random.default_rng(15)
    _check_gradients(net, rng, _features(rng, 2, 48, 80))


def _check_gradients(net, rng, feat):
    target = _target_with_margin(rng, net.get_params(), feat)
    loss, grad = net.loss_grad(feat, target)
    rloss, rgrad = R.loss_and_grads(net.get_params(), feat, target)
    assert abs(loss - rloss) <= 1e-2 * rloss
    for name, ws, bs in R.layer_slices():
        for part, s in (("w", ws), ("b", bs)):
            g, r = grad[s], rgrad[s]
            if name == "e0a" and part == "w":  # padded input channels have zero gradient
                assert not g.reshape(9, 32, 16)[:, :, 10:].any()
            nr = np.linalg.norm(r)
            rel = np.linalg.norm(g - r) / nr
            cos = float(g @ r) / (np.linalg.norm(g) * nr)
            print("%s.%s relative L2 %.3e cosine %.6f" % (name, part, rel, cos))
            assert rel < 0.06, (name, part, rel)
            assert cos > 0.999, (name, part, cos)


def test_loss_without_gradient_is_the_same_loss(net):
    rng = np.random.default_rng(13)
    feat = _features(rng, 1, 32, 48)
    target = _target(feat)
    l1, _ = net.loss_grad(feat, target, with_grad=False)
    l2, g = net.loss_grad(feat, target)
    assert l1 == l2 and np.isfinite(g).all()


def test_training_reduces_loss(net):
    rng = np.random.default_rng(14)
    feat = _features(rng, 16, 64, 64)
    target = _target(feat)
    net.set_dataset(feat, target)
    before, _ = net.loss_grad(feat[:8], target[:8], with_grad=False)
    losses = net.train(steps=80, batch=4, lr=2e-3, seed=5)
    after, _ = net.loss_grad(feat[:8], target[:8], with_grad=False)
    print("loss %.4f -> %.4f; steps: first %.4f last %.4f" % (before, after, losses[0], losses[-1]))
    assert np.isfinite(losses).all()
    assert after < 0.5 * before
    assert losses[-10:].mean() < 0.6 * losses[:5].mean()


def test_resumed_training_is_bit_identical(renderer):
    """A run split in two (state saved after the first part, restored into a
    fresh network) ends on the same parameter and moment bits as the run
    done in one piece."""
    rng = np.random.default_rng(15)
    feat = _features(rng, 8, 32, 32)
    target = _target(feat)
    a = LearnedDenoiser(renderer, seed=21)
    a.set_dataset(feat, target)
    a.train(steps=6, batch=4, lr=1e-3, seed=1)
    saved = a.get_state()
    assert int(saved["steps"][0]) == 6 and np.abs(saved["m"]).max() > 0 and saved["v"].min() >= 0
    a.train(steps=5, batch=4, lr=5e-4, seed=2)
    whole = a.get_state()
    b = LearnedDenoiser(renderer, seed=99)  # other weights, cleared moments
    b.set_dataset(feat, target)
    b.set_state(saved)
    b.train(steps=5, batch=4, lr=5e-4, seed=2)
    split = b.get_state()
    for k in ("params", "m", "v", "steps"):
        assert np.array_equal(whole[k], split[k]), k
    # without the moments (set_params alone) the continuation differs
    b.set_params(saved["params"])
    b.train(steps=5, batch=4, lr=5e-4, seed=2)
    assert not np.array_equal(b.get_params(), whole["params"])


def test_denoise_frame_matches_forward_on_padded_features(renderer, net):
    scene = SceneBuilder.config(1, seed=3)
    renderer.set_scene(scene, isgbuf=True)
    w, h = 72, 40  # not multiples of 16: the frame is padded to 80 x 48
    renderer.resize(w, h)
    renderer.render(spp=2)
    renderer.gbuffer()
    out = net.execute()
    fin = net.read_result()
    assert out.shape == (h, w, 4) and fin.shape == (h, w, 4)
    assert np.isfinite(fin).all() and (fin[..., :3] >= 0).all() and (fin[..., 3] == 1).all()
    f = net.frame_features()
    fp = np.zeros((1, 48, 80, 10), np.float32)
    fp[:, :h, :w] = f
    pred = net.forward(fp)[0, :h, :w]
    want = np.maximum(np.expm1(pred), 0.0)
    np.testing.assert_allclose(fin[..., :3], want, rtol=2e-6, atol=1e-7)
    t = want * net.exposure
    t = np.power(t / (t + 0.3), 0.4545)
    rgba = np.floor(np.clip(t, 0, 1) * 255.0 + 0.5)
    assert np.abs(out[..., :3].astype(np.int32) - rgba).max() <= 1
    assert (out[..., 3] == 255).all()


def test_argument_errors(net):
    from webgpuraytracer_amd import _abi as A
    with pytest.raises(A.PTError):
        net.forward(np.zeros((1, 24, 32, 10), np.float32))  # 24 is not a multiple of 16
    with pytest.raises(ValueError):
        net.forward(np.zeros((1, 32, 32, 9), np.float32))
    with pytest.raises(A.PTError):  # zero steps
        net.r._check(net.r.lib.pt_unet_train(net.r.ctx, 0, 1, 1e-3, 0, None))


def test_dataset_capture_from_frames_matches_host_features(renderer, net):
    """Crops captured on the device from a rendered frame train on exactly the
    features and targets the host would build from the same frame."""
    scene = SceneBuilder.config(1, seed=5)
    renderer.set_scene(scene, isgbuf=True)
    renderer.resize(96, 64)
    renderer.gbuffer()
    renderer.render(spp=8, base_seed=3)
    net.dataset_reserve(1, 32, 48)
    net.capture(0, target=True, x0=40, y0=16)
    target = renderer.read_accum()[16:48, 40:88, :3][None].copy()
    renderer.reset()
    renderer.render(spp=1, base_seed=9)
    net.capture(0, target=False, x0=40, y0=16)
    feat = net.frame_features()[:, 16:48, 40:88].copy()
    with pytest.raises(Exception):
        net.capture(0, target=False, x0=60, y0=0)  # window past the right edge
    with pytest.raises(Exception):
        net.capture(1, target=False)  # crop index outside the set
    p0 = net.get_params()
    losses = net.train(steps=2, batch=1, lr=0.0, seed=1)  # lr 0: weights unchanged
    assert np.array_equal(net.get_params(), p0)
    host, _ = net.loss_grad(feat, target, with_grad=False)
    assert losses[0] == pytest.approx(host, rel=1e-6) and losses[1] == pytest.approx(host, rel=1e-6)


def test_pretrained_weights_beat_atrous_at_one_spp(renderer):
    """The shipped weights on a scene outside their training seeds: at 1 spp
    the learned denoiser is closer to a 2048-spp render than the a-trous
    filter (relMSE on the tonemapped image)."""
    from webgpuraytracer_amd import Denoiser
    from webgpuraytracer_amd.unet_train import rel_mse, tonemap
    scene = SceneBuilder.config(1, seed=7)
    renderer.set_scene(scene, isgbuf=True)
    renderer.resize(192, 112)
    renderer.gbuffer()
    renderer.reset()
    renderer.render(spp=2048, base_seed=99)
    ref = tonemap(renderer.read_accum()[..., :3])
    renderer.reset()
    renderer.render(spp=1, base_seed=5)
    net = LearnedDenoiser(renderer, pretrained=True)
    net.execute()
    un = tonemap(net.read_result()[..., :3])
    dn = Denoiser(renderer)
    dn.passes = 5
    dn.execute(0)
    at = dn.read_result()[..., :3]
    e_un, e_at = rel_mse(un, ref), rel_mse(at, ref)
    # fp8 inference, calibrated on this frame
    net.set_precision(True)
    net.execute()
    e_8 = rel_mse(tonemap(net.read_result()[..., :3]), ref)
    print("1 spp relMSE: U-Net %.4f (fp8 %.4f), a-trous %.4f" % (e_un, e_8, e_at))
    assert e_un < 0.25 * e_at and e_un < 0.02  # measured: 0.0087 against 0.0837 at 192x112
    # fp8 calibrated on this small frame: measured 0.0103 (1.18x bf16); the full
    # 1080p frame is held to 1.15x below
    assert e_8 < 1.3 * e_un


def _fp8_frame(renderer, seed, w=96, h=64, spp=1):
    renderer.set_scene(SceneBuilder.config(1, seed=seed), isgbuf=True)
    renderer.resize(w, h)
    renderer.gbuffer()
    renderer.reset()
    renderer.render(spp=spp, base_seed=5)


def test_fp8_quantization_matches_host_restatement(renderer):
    """pt_unet_set_precision(FP8): the weight scales and e4m3 weight bytes are
    exactly what the documented rule gives from the returned input scales and
    the fp32 parameters (torch's float8_e4m3fn conversion as the encoder)."""
    import torch
    net = LearnedDenoiser(renderer, pretrained=True)
    _fp8_frame(renderer, 8)
    with pytest.raises(Exception):
        net.fp8_state(*R.fp8_sizes())  # not in fp8 mode yet
    net.set_precision(True)
    sa, sw, w8 = net.fp8_state(*R.fp8_sizes())
    p = net.get_params()
    assert (sa[10:16] == 1).all() and (sa[:10] > 0).all()  # padded input channels never light up
    ao = so = wo = 0
    for (ci, co, _), (name, ws, _) in zip(R.layer_table(), R.layer_slices()):
        w = torch.from_numpy(p[ws].reshape(9, co, ci)) * torch.from_numpy(sa[ao:ao + ci])
        m = w.abs().amax(dim=(0, 2))
        s = torch.where(m > 0, m / 448.0, torch.ones_like(m))
        assert np.array_equal(s.numpy(), sw[so:so + co]), name
        q = (w / s[None, :, None]).to(torch.float8_e4m3fn).view(torch.uint8).numpy().ravel()
        assert np.array_equal(q, w8[wo:wo + 9 * co * ci]), name
        ao, so, wo = ao + ci, so + co, wo + 9 * co * ci
    net.set_params(p)  # new weights drop the fp8 state
    with pytest.raises(Exception):
        net.fp8_state(*R.fp8_sizes())


def test_fp8_forward_matches_fp8_restatement(renderer):
    """The fp8 convolutions (e4m3 operands on the 16x16x128 MFMA) against
    the same quantization in fp32 PyTorch.  The MFMA does not accumulate
    like sequential fp32: its 128-term dot products of e4m3 values were
    measured up to 1.1e-4 of the largest product's magnitude away from the
    exact sum (sequential fp32: within 572 ulps of the result), which flips
    bf16 roundings and then e4m3 steps (6%) in later layers; so the bound is
    relative: the device output is several times closer to the fp8
    restatement than to the bf16 network (a wrong tap, channel or scale
    mapping lands at the bf16 distance or beyond)."""
    import torch
    net = LearnedDenoiser(renderer, pretrained=True)
    _fp8_frame(renderer, 8)
    net.set_precision(True)  # calibrated on seed 8
    _fp8_frame(renderer, 9, w=112, h=80, spp=2)
    feat = net.frame_features()
    pred8 = net.forward(feat)
    net.set_precision(False)
    pred16 = net.forward(feat)
    net.set_precision(True)  # recalibrated on the seed-9 frame
    sa, sw, w8 = net.fp8_state(*R.fp8_sizes())
    pred8b = net.forward(feat)
    with torch.no_grad():
        emu = R.forward_fp8(R.unflatten(net.get_params()), feat, sa, sw, w8).permute(0, 2, 3, 1).numpy()
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))  # noqa: E731
    e_emu, e_16, e_cal = rel(pred8b, emu), rel(pred8b, pred16), rel(pred8, pred16)
    print("fp8 vs restatement %.2e, fp8 vs bf16 %.2e (other-frame calibration %.2e)" % (e_emu, e_16, e_cal))
    assert e_emu < 0.5 * e_16 and e_emu < 0.025  # measured: 1.5e-2 against 3.7e-2
    assert e_16 < 0.05 and e_cal < 0.05


def test_c5_1080p_unet_one_spp_beats_atrous_four_spp(renderer):
    """The equal-time comparison's frame: 1920x1080, a seeded small-analytic
    scene outside the training seeds, a 4096-spp reference.  The learned
    denoiser on a 1-spp frame is closer to the reference than 5-pass a-trous
    on a 4-spp frame (relMSE of the tonemapped images), and its fp8 variant
    stays within 15% of bf16."""
    from webgpuraytracer_amd import Denoiser
    from webgpuraytracer_amd.unet_train import rel_mse, tonemap
    scene = SceneBuilder.config(1, seed=7)
    renderer.set_scene(scene, isgbuf=True)
    renderer.resize(1920, 1080)
    renderer.gbuffer()
    renderer.reset()
    renderer.render(spp=4096, base_seed=123)
    ref = tonemap(renderer.read_accum()[..., :3])
    dn = Denoiser(renderer)
    dn.passes = 5
    renderer.reset()
    renderer.render(spp=4, base_seed=5)
    dn.execute(0)
    e_at4 = rel_mse(dn.read_result()[..., :3], ref)
    renderer.reset()
    renderer.render(spp=1, base_seed=5)
    net = LearnedDenoiser(renderer, pretrained=True)
    net.execute()
    e_un = rel_mse(tonemap(net.read_result()[..., :3]), ref)
    net.set_precision(True)
    net.execute()
    e_8 = rel_mse(tonemap(net.read_result()[..., :3]), ref)
    net.set_precision(False)
    print("1080p relMSE vs 4096 spp: U-Net 1 spp %.5f (fp8 %.5f), a-trous 4 spp %.5f" % (e_un, e_8, e_at4))
    assert e_un < e_at4
    assert e_8 < 1.15 * e_un
