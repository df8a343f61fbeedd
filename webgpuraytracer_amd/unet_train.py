"""Training data, training and evaluation of the learned denoiser.

    python -m webgpuraytracer_amd.unet_train train --scenes 128 --variants 16 --out unet.npy
    python -m webgpuraytracer_amd.unet_train eval --weights unet.npy --width 1920 --height 1080
    python -m webgpuraytracer_amd.unet_train timing

`train` renders seeded small-scene frames (the 12-primitive analytic scene
family with an emissive quad and a procedural sky) with the GPU tracer:
per scene one high-spp target and `variants` independent low-spp inputs
with the G-buffer, captured into a device-side training set
(pt_unet_dataset_capture), then trains the U-Net with Adam on the device.

`eval` is the equal-time comparison: one 1080p frame of a held-out scene at
1, 2, 4, ... spp per frame, denoised by the à-trous filter (pt_denoise, the
reference's Denoiser with its default sigmas, 5 iterations) and by the U-Net
(pt_unet_denoise), scored against a high-spp render of the same frame by
relMSE and SSIM on the tonemapped image, with the wall-clock time of
render + denoise per frame.

`timing` measures the training step and 1080p inference of the network in
TFLOP/s of algorithmic convolution work.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

from . import _abi as A
from .renderer import Denoiser, LearnedDenoiser, Renderer
from .scene import SceneBuilder

# (cin, cout) of the 19 convolutions with the channels that carry data
# (10 input features, 3 outputs), and the level (1/2^level resolution)
_C = (32, 48, 64, 96, 128)


def conv_table():
    t = []
    cin = 10
    for lvl in range(4):
        t += [(cin, _C[lvl], lvl), (_C[lvl], _C[lvl], lvl)]
        cin = _C[lvl]
    t += [(_C[3], _C[4], 4), (_C[4], _C[4], 4)]
    for lvl in (3, 2, 1, 0):
        t += [(_C[lvl + 1] + _C[lvl], _C[lvl], lvl), (_C[lvl], _C[lvl], lvl)]
    t.append((_C[0], 3, 0))
    return t


def flops_per_pixel(training):
    """Multiply-adds x 2 of the 3x3 convolutions per full-resolution pixel:
    forward only, or forward + weight gradient + data gradient (none for the
    first layer, whose input is data)."""
    f = 0.0
    for i, (ci, co, lvl) in enumerate(conv_table()):
        fwd = 2.0 * 9 * ci * co / 4 ** lvl
        f += fwd
        if training:
            f += fwd + (fwd if i > 0 else 0.0)
    return f


def tonemap(x, exposure=1.0):
    """The display tonemap of the tracer and the à-trous filter."""
    t = np.maximum(x, 0.0) * exposure
    return np.power(t / (t + 0.3), 0.4545)


def rel_mse(x, ref):
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 0.01)))


def ssim(x, ref):
    """Mean SSIM over channels (Gaussian window sigma 1.5, K1 0.01, K2 0.03,
    data range 1)."""
    from scipy.ndimage import gaussian_filter
    c1, c2 = 0.01 ** 2, 0.03 ** 2
    vals = []
    for ch in range(3):
        a, b = x[..., ch].astype(np.float64), ref[..., ch].astype(np.float64)
        mu_a, mu_b = gaussian_filter(a, 1.5), gaussian_filter(b, 1.5)
        saa = gaussian_filter(a * a, 1.5) - mu_a ** 2
        sbb = gaussian_filter(b * b, 1.5) - mu_b ** 2
        sab = gaussian_filter(a * b, 1.5) - mu_a * mu_b
        m = ((2 * mu_a * mu_b + c1) * (2 * sab + c2)) / ((mu_a ** 2 + mu_b ** 2 + c1) * (saa + sbb + c2))
        vals.append(m.mean())
    return float(np.mean(vals))


def _log(msg):
    print(msg, flush=True)


def build_training_set(r, net, scenes, variants, size, frame, target_spp, input_spp, seed0):
    """Renders scenes x variants crops (size x size, at seeded positions in a
    frame[0] x frame[1] render) into the device training set; variant v's
    input has input_spp[v % len(input_spp)] samples per pixel.  Returns the
    render time in seconds."""
    fw, fh = frame
    if fw < size or fh < size:
        raise ValueError("frame %dx%d is smaller than the %d^2 crops" % (fw, fh, size))
    net.dataset_reserve(scenes * variants, size, size)
    rng = np.random.default_rng(seed0)
    t0 = time.perf_counter()
    for s in range(scenes):
        scene = SceneBuilder.config(1, seed=seed0 + s)
        r.set_scene(scene, isgbuf=True)
        if s == 0:
            r.resize(fw, fh)
        r.gbuffer()
        r.reset()
        r.render(spp=target_spp, base_seed=0x7A000000 + s)
        pos = [(int(rng.integers(0, fw - size + 1)), int(rng.integers(0, fh - size + 1))) for _ in range(variants)]
        for v, (x0, y0) in enumerate(pos):
            net.capture(s * variants + v, target=True, x0=x0, y0=y0)
        for v, (x0, y0) in enumerate(pos):
            r.reset()
            r.render(spp=input_spp[v % len(input_spp)], base_seed=(s * variants + v) * 7919 + 13)
            net.capture(s * variants + v, target=False, x0=x0, y0=y0)
        if (s + 1) % 16 == 0 or s + 1 == scenes:
            _log("  scenes %d/%d rendered (%.1f s)" % (s + 1, scenes, time.perf_counter() - t0))
    return time.perf_counter() - t0


def cmd_train(a):
    r = Renderer(0)
    r.set_option(A.PT_OPT_COUNT_STATS, 0)
    net = LearnedDenoiser(r, seed=a.seed)
    _log("rendering %d scenes x %d crops of %d^2 from %dx%d frames (targets %d spp, inputs %s spp)"
         % (a.scenes, a.variants, a.size, a.frame[0], a.frame[1], a.target_spp, a.input_spp))
    t_data = build_training_set(r, net, a.scenes, a.variants, a.size, a.frame, a.target_spp, a.input_spp,
                                a.scene_seed)
    # a run split over several processes: each renders its own scenes, resumes the
    # parameters and Adam state of the previous part and continues its lr schedule
    step0 = 0
    if a.resume:
        with np.load(a.resume, allow_pickle=False) as z:
            st = {k: z[k] for k in ("params", "m", "v", "steps")}
        net.set_state(st)
        step0 = int(st["steps"][0])
        _log("resumed from %s at step %d" % (a.resume, step0))
    total = max(a.total_steps or (step0 + a.steps), 1)
    losses = []
    t0 = time.perf_counter()
    chunk = 200
    done = 0
    while done < a.steps:
        n = min(chunk, a.steps - done)
        frac = (step0 + done) / total
        lr = a.lr * (0.5 * (1.0 + np.cos(np.pi * frac)) * 0.9 + 0.1)  # cosine decay to lr / 10
        losses.extend(net.train(steps=n, batch=a.batch, lr=float(lr),
                                seed=a.seed * 1000003 + step0 + done).tolist())
        done += n
        _log("  step %d/%d loss %.4f (lr %.2e, %.1f s)" % (step0 + done, total, float(np.mean(losses[-n:])), lr,
                                                        time.perf_counter() - t0))
    t_train = time.perf_counter() - t0
    np.save(a.out, net.get_params())
    if a.save_state:
        np.savez(a.save_state, **net.get_state())
    log = {"scenes": a.scenes, "scene_seed": a.scene_seed, "variants": a.variants, "size": a.size, "frame": a.frame,
           "target_spp": a.target_spp, "input_spp": a.input_spp, "steps": a.steps, "step0": step0,
           "total_steps": total, "batch": a.batch, "lr": a.lr,
           "render_s": t_data, "train_s": t_train, "ms_per_step": 1e3 * t_train / a.steps,
           "loss_first100": float(np.mean(losses[:100])), "loss_last100": float(np.mean(losses[-100:]))}
    _log(json.dumps(log))
    if a.log:
        with open(a.log, "w") as f:
            json.dump({"summary": log, "losses": losses[::10]}, f)
    r.close()


def _timed(fn, reps=3):
    fn()
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        t.append(time.perf_counter() - t0)
    return 1e3 * float(np.median(t))


def cmd_eval(a):
    r = Renderer(0)
    r.set_option(A.PT_OPT_COUNT_STATS, 0)
    net = LearnedDenoiser(r)
    net.set_params(np.load(a.weights, allow_pickle=False))
    dn = Denoiser(r)
    dn.passes = a.atrous_passes
    scene = SceneBuilder.config(1, seed=a.scene_seed)
    r.set_scene(scene, isgbuf=True)
    r.resize(a.width, a.height)
    r.gbuffer()
    _log("reference: %d spp at %dx%d" % (a.ref_spp, a.width, a.height))
    t0 = time.perf_counter()
    r.reset()
    left = a.ref_spp
    while left > 0:  # in pieces of 1024 passes, so the call stays short
        r.render(spp=min(1024, left), base_seed=0x51A7)
        left -= min(1024, left)
    ref_lin = r.read_accum()[..., :3].copy()
    ref = tonemap(ref_lin)
    _log("  %.1f s" % (time.perf_counter() - t0))
    rows = []
    for spp in a.spp:
        def render():
            r.reset()
            r.render(spp=spp, base_seed=0xBEEF + spp)
        t_render = _timed(render)
        render()
        noisy = tonemap(r.read_accum()[..., :3])
        # the denoise calls alone (synchronous), without reading the image back
        u = dn.uniforms()
        t_atrous = _timed(lambda: r._check(r.lib.pt_denoise(r.ctx, ctypes.byref(u), 0)))
        at = dn.read_result()[..., :3].copy()
        t_unet = _timed(lambda: r._check(r.lib.pt_unet_denoise(r.ctx, net.exposure)))
        un = tonemap(net.read_result()[..., :3])
        row = {"spp": spp, "ms_render": t_render, "ms_atrous": t_atrous, "ms_unet": t_unet,
               "relmse_noisy": rel_mse(noisy, ref), "relmse_atrous": rel_mse(at, ref),
               "relmse_unet": rel_mse(un, ref), "ssim_noisy": ssim(noisy, ref), "ssim_atrous": ssim(at, ref),
               "ssim_unet": ssim(un, ref)}
        rows.append(row)
        _log(json.dumps(row))
    if a.fp8:
        # fp8 inference, calibrated on a 1-spp frame of a training-family scene (not the evaluated
        # one), then the same seeded frames again
        r.set_scene(SceneBuilder.config(1, seed=a.calib_seed), isgbuf=True)
        r.gbuffer()
        r.reset()
        r.render(spp=1, base_seed=0xCA1B)
        net.set_precision(True)
        r.set_scene(scene, isgbuf=True)
        r.gbuffer()
        for row in rows:
            r.reset()
            r.render(spp=row["spp"], base_seed=0xBEEF + row["spp"])
            row["ms_unet_fp8"] = _timed(lambda: r._check(r.lib.pt_unet_denoise(r.ctx, net.exposure)))
            un8 = tonemap(net.read_result()[..., :3])
            row["relmse_unet_fp8"], row["ssim_unet_fp8"] = rel_mse(un8, ref), ssim(un8, ref)
            _log(json.dumps({k: row[k] for k in ("spp", "ms_unet", "ms_unet_fp8", "relmse_unet", "relmse_unet_fp8")}))
    # equal quality: the fewest U-Net spp whose relMSE is at or below the à-trous relMSE at each spp
    eq = []
    for ra in rows:
        best = [ru for ru in rows if ru["relmse_unet"] <= ra["relmse_atrous"]]
        if best:
            ru = min(best, key=lambda x: x["spp"])
            eq.append({"atrous_spp": ra["spp"], "unet_spp": ru["spp"], "spp_ratio": ra["spp"] / ru["spp"],
                       "ms_atrous_frame": ra["ms_render"] + ra["ms_atrous"],
                       "ms_unet_frame": ru["ms_render"] + ru["ms_unet"],
                       "speedup": (ra["ms_render"] + ra["ms_atrous"]) / (ru["ms_render"] + ru["ms_unet"])})
    out = {"width": a.width, "height": a.height, "ref_spp": a.ref_spp, "scene_seed": a.scene_seed,
           "atrous_passes": a.atrous_passes, "rows": rows, "equal_quality": eq}
    _log(json.dumps({"equal_quality": eq}))
    if a.json:
        with open(a.json, "w") as f:
            json.dump(out, f, indent=1)
    r.close()


def cmd_timing(a):
    r = Renderer(0)
    net = LearnedDenoiser(r, seed=1)
    rng = np.random.default_rng(0)
    n = 2 * a.batch
    feat = rng.random((n, a.size, a.size, 10), dtype=np.float32)
    net.set_dataset(feat, rng.random((n, a.size, a.size, 3), dtype=np.float32))
    net.train(steps=3, batch=a.batch, lr=1e-4)
    t0 = time.perf_counter()
    net.train(steps=a.steps, batch=a.batch, lr=1e-4)
    ms = 1e3 * (time.perf_counter() - t0) / a.steps
    px = a.batch * a.size * a.size
    train = {"workload": "unet train step", "batch": a.batch, "crop": a.size, "ms_per_step": ms,
             "tflops": flops_per_pixel(True) * px / ms * 1e-9}
    _log(json.dumps(train))
    f = rng.random((1, a.height, a.width, 10), dtype=np.float32)
    net.forward(f)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        net.forward(f)
    ms = 1e3 * (time.perf_counter() - t0) / a.steps
    infer = {"workload": "unet forward (host features, incl. copies)", "h": a.height, "w": a.width, "ms": ms,
             "tflops": flops_per_pixel(False) * a.height * a.width / ms * 1e-9}
    _log(json.dumps(infer))
    r.close()


def main(argv=None):
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = p.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--scenes", type=int, default=128)
    t.add_argument("--variants", type=int, default=16)
    t.add_argument("--size", type=int, default=256)
    t.add_argument("--frame", type=lambda s: [int(x) for x in s.split("x")], default=[256, 256])
    t.add_argument("--target-spp", type=int, default=2048)
    t.add_argument("--input-spp", type=lambda s: [int(x) for x in s.split(",")], default=[1, 2, 4, 8])
    t.add_argument("--steps", type=int, default=3000)
    t.add_argument("--batch", type=int, default=16)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--seed", type=int, default=1)
    t.add_argument("--scene-seed", type=int, default=1000)
    t.add_argument("--out", default="unet.npy")
    t.add_argument("--resume", default=None, help="an .npz written by --save-state: continue that run")
    t.add_argument("--save-state", default=None, help="write parameters and Adam state (.npz) to resume from")
    t.add_argument("--total-steps", type=int, default=0,
                   help="steps of the whole (possibly split) run, for the lr schedule (default: this part's)")
    t.add_argument("--log", default=None)
    e = sub.add_parser("eval")
    e.add_argument("--weights", required=True)
    e.add_argument("--width", type=int, default=1920)
    e.add_argument("--height", type=int, default=1080)
    e.add_argument("--ref-spp", type=int, default=16384)
    e.add_argument("--spp", type=lambda s: [int(x) for x in s.split(",")], default=[1, 2, 4, 8, 16, 32])
    e.add_argument("--scene-seed", type=int, default=7)
    e.add_argument("--atrous-passes", type=int, default=5)
    e.add_argument("--json", default=None)
    e.add_argument("--fp8", action="store_true", help="also denoise with fp8 inference (pt_unet_set_precision)")
    e.add_argument("--calib-seed", type=int, default=1000, help="scene seed of the fp8 calibration frame")
    m = sub.add_parser("timing")
    m.add_argument("--batch", type=int, default=16)
    m.add_argument("--size", type=int, default=256)
    m.add_argument("--steps", type=int, default=20)
    m.add_argument("--width", type=int, default=1920)
    m.add_argument("--height", type=int, default=1088)
    a = p.parse_args(argv)
    {"train": cmd_train, "eval": cmd_eval, "timing": cmd_timing}[a.cmd](a)


if __name__ == "__main__":
    sys.exit(main())
