"""Renderer / Denoiser: the tracer.js host classes over the libpt C ABI.

Method names and argument meanings follow tracer.js's Renderer
(tracer.js:1718-2417) and Denoiser (tracer.js:2418-2638); every call goes
through libpt.so, which fails loudly when the HIP library or a GPU is absent.
"""
import ctypes as C
import os

import numpy as np

from . import _abi as A


class Renderer:
    """new Renderer(canvas) + init() (tracer.js:1719-1770)."""

    def __init__(self, device=0):
        self.lib = A.load_pt()
        self.ctx = C.c_void_p()
        rc = self.lib.pt_create(C.byref(self.ctx), device)
        if rc:
            raise A.PTError(rc, "pt_create(device=%d) failed: no usable HIP device" % device)
        self.scene = None
        self.frame = 0
        self.width = self.height = 0
        self.isgbuf = False

    def _check(self, rc):
        if rc:
            raise A.PTError(rc, (self.lib.pt_last_error(self.ctx) or b"").decode())
        return rc

    def close(self):
        if self.ctx:
            self.lib.pt_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_option(self, key, value):
        return self._check(self.lib.pt_set_option(self.ctx, key, value))

    def get_option(self, key):
        v = C.c_int64()
        self._check(self.lib.pt_get_option(self.ctx, key, C.byref(v)))
        return v.value

    # -- setScene / updateObjects / resize (tracer.js:1955-2205) --
    def set_scene(self, scene, isgbuf=False):
        """scene: a finalized SceneBuilder, or an A.SceneDesc."""
        desc = scene.desc() if hasattr(scene, "desc") else scene
        self._check(self.lib.pt_upload_scene(self.ctx, C.byref(desc)))
        self.scene = scene
        self.isgbuf = isgbuf
        if self.width:
            self.resize(self.width, self.height)
        return self

    def update_objects(self, scene):
        """Fast path; falls back to a full set_scene on capacity / flag changes,
        as Renderer.updateObjects does (tracer.js:2140-2163)."""
        desc = scene.desc() if hasattr(scene, "desc") else scene
        rc = self.lib.pt_update_objects(self.ctx, C.byref(desc))
        if rc in (A.PT_ERR_CAPACITY, A.PT_ERR_FLAGS):
            return self.set_scene(scene, self.isgbuf)
        self._check(rc)
        self.scene = scene
        return self

    def resize(self, width, height):
        self._check(self.lib.pt_resize(self.ctx, width, height, int(self.isgbuf)))
        self.width, self.height = width, height
        self.reset()
        return self

    def reset(self):
        self.frame = 0

    # -- render / renderSection (tracer.js:2242-2309) --
    def params(self, sample_number=None, seed=0, section=(0, 0)):
        p = self.scene.params(self.width, self.height, self.frame if sample_number is None else sample_number, seed)
        p.section[0], p.section[1] = section
        return p

    def render_section(self, params, section_size=(0, 0), passes=1, seeds=None):
        st = A.Stats()
        sp = None
        if seeds is not None:
            s = np.ascontiguousarray(seeds, np.uint32)
            assert s.size == passes
            sp = s.ctypes.data_as(C.POINTER(A.u32))
        self._check(self.lib.pt_trace(self.ctx, C.byref(params), section_size[0], section_size[1], passes, sp,
                                      C.byref(st)))
        return st

    def render(self, spp=1, seeds=None, base_seed=0):
        """spp progressive samples over the whole frame, starting at self.frame."""
        if seeds is None:
            seeds = [(base_seed * 0x9E3779B1 + (self.frame + k) * 0x85EBCA77) & 0xFFFFFFFF for k in range(spp)]
        p = self.params(seed=seeds[0])
        st = self.render_section(p, passes=spp, seeds=seeds)
        self.frame += spp
        return st

    def gbuffer(self, params=None):
        self._check(self.lib.pt_gbuffer(self.ctx, C.byref(params or self.params(sample_number=0))))

    def trace_hits(self, params, n_bounces=2, rays=False):
        """Hits of each path's first n_bounces trace_scene calls (one sample per
        pixel); with rays=True also the rays traced ([pixels * n_bounces][6])."""
        out = (A.Hit * (self.width * self.height * n_bounces))()
        if not rays:
            self._check(self.lib.pt_trace_hits(self.ctx, C.byref(params), n_bounces, out))
            return out
        r = np.zeros((self.width * self.height * n_bounces, 6), np.float32)
        self._check(self.lib.pt_trace_hits_rays(self.ctx, C.byref(params), n_bounces, out, r.ctypes.data))
        return out, r

    def bounce_stats(self, n=16):
        """Per-bounce work of the last render (wavefront backend): extension
        rays, shadow rays, and device ms per bounce (with
        PT_OPT_BOUNCE_TIMING set, else zeros), as three arrays of n."""
        ext, sh, ms = np.zeros(n, np.uint64), np.zeros(n, np.uint64), np.zeros(n, np.float64)
        self._check(self.lib.pt_get_bounce_stats(self.ctx, ext.ctypes.data, sh.ctypes.data, ms.ctypes.data, n))
        return ext, sh, ms

    # -- readback --
    def read_accum(self):
        a = np.empty((self.height, self.width, 4), np.float32)
        self._check(self.lib.pt_read_accum(self.ctx, a.ctypes.data, a.nbytes))
        return a

    def write_accum(self, a):
        a = np.ascontiguousarray(a, np.float32)
        self._check(self.lib.pt_write_accum(self.ctx, a.ctypes.data, a.nbytes))

    def read_output(self):
        a = np.empty((self.height, self.width, 4), np.uint8)
        self._check(self.lib.pt_read_output(self.ctx, a.ctypes.data, a.nbytes))
        return a

    def read_gbuffer(self):
        a = np.empty((self.height, self.width, 12), np.float32)
        self._check(self.lib.pt_read_gbuffer(self.ctx, a.ctypes.data, a.nbytes))
        return a

    def write_gbuffer(self, g):
        g = np.ascontiguousarray(g, np.float32)
        self._check(self.lib.pt_write_gbuffer(self.ctx, g.ctypes.data, g.nbytes))

    def read_sky_cdf(self, count):
        a = np.empty(count, np.float32)
        tot = A.f32()
        self._check(self.lib.pt_read_sky_cdf(self.ctx, a.ctypes.data, count, C.byref(tot)))
        return a, tot.value

    # -- ray queries --
    def trace_rays(self, rays):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        out = (A.Hit * r.shape[0])()
        self._check(self.lib.pt_trace_rays(self.ctx, r.ctypes.data, r.shape[0], out))
        return out

    def trace_shadow_rays(self, rays, target_idx, target_dist):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        t = np.ascontiguousarray(target_idx, np.int32)
        d = np.ascontiguousarray(target_dist, np.float32)
        out = np.empty((r.shape[0], 3), np.float32)
        self._check(self.lib.pt_trace_shadow_rays(self.ctx, r.ctypes.data, t.ctypes.data, d.ctypes.data,
                                                  r.shape[0], out.ctypes.data))
        return out


class Denoiser:
    """Denoiser (tracer.js:2418-2637) with the reference's default sigmas."""

    def __init__(self, renderer):
        self.r = renderer
        self.passes = 4
        self.sigma_color = 0.30
        self.sigma_normal = 128.0
        self.sigma_depth = 0.2
        self.sigma_albedo = 0.07
        self.sigma_material = 0.07
        self.glossy_protect = 0.75
        self.exposure = 1.0
        self.use_median = True
        self.use_demodulate = True

    def uniforms(self):
        u = A.DenoiseParams()
        u.width, u.height, u.step_width, u.passes = self.r.width, self.r.height, 1, self.passes
        u.sigma_color, u.sigma_normal, u.sigma_depth = self.sigma_color, self.sigma_normal, self.sigma_depth
        u.sigma_albedo, u.sigma_material, u.glossy_protect = self.sigma_albedo, self.sigma_material, self.glossy_protect
        u.exposure = self.exposure
        u.flags = (1 if self.use_median else 0) | (2 if self.use_demodulate else 0)
        return u

    def execute(self, stride_mode=0):
        """stride_mode 0: steps 1,2,4,...; 1: the reference's constant last step."""
        self.r._check(self.r.lib.pt_denoise(self.r.ctx, C.byref(self.uniforms()), stride_mode))
        return self.r.read_output()

    def read_result(self):
        a = np.empty((self.r.height, self.r.width, 4), np.float32)
        self.r._check(self.r.lib.pt_read_denoised(self.r.ctx, a.ctypes.data, a.nbytes))
        return a


class LearnedDenoiser:
    """The learned alternative to Denoiser on the same frame (pt_unet.h): a
    5-level U-Net over radiance, albedo, normal and depth, trained on the
    device.  Features are fp32 [n][h][w][10], targets [n][h][w][3]."""

    N_FEATURES = 10
    # Weights trained by `python -m webgpuraytracer_amd.unet_train train --scenes 500 --variants 40
    # --frame 960x540 --target-spp 1024 --input-spp 1,2,4,8 --steps 20000` on the small analytic
    # scene family (12 primitives, emissive quad, procedural sky).
    PRETRAINED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "unet_small_scenes.npy")

    def __init__(self, renderer, seed=0, pretrained=False):
        self.r = renderer
        self.exposure = 1.0
        self.r._check(self.r.lib.pt_unet_init(self.r.ctx, seed))
        if pretrained:
            self.set_params(np.load(self.PRETRAINED, allow_pickle=False))

    def param_count(self):
        n = C.c_uint64()
        self.r._check(self.r.lib.pt_unet_param_count(self.r.ctx, C.byref(n)))
        return n.value

    def get_params(self):
        p = np.empty(self.param_count(), np.float32)
        self.r._check(self.r.lib.pt_unet_get_params(self.r.ctx, p.ctypes.data, p.size))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self.r._check(self.r.lib.pt_unet_set_params(self.r.ctx, p.ctypes.data, p.size))

    def get_state(self):
        """Parameters and Adam state: a dict of arrays (params, m, v, steps)
        that set_state restores, so training resumes in another process."""
        n = self.param_count()
        m, v = np.empty(n, np.float32), np.empty(n, np.float32)
        steps = C.c_uint32()
        self.r._check(self.r.lib.pt_unet_get_adam(self.r.ctx, m.ctypes.data, v.ctypes.data, n, C.byref(steps)))
        return {"params": self.get_params(), "m": m, "v": v, "steps": np.array([steps.value], np.uint32)}

    def set_state(self, st):
        self.set_params(st["params"])
        m = np.ascontiguousarray(st["m"], np.float32)
        v = np.ascontiguousarray(st["v"], np.float32)
        self.r._check(self.r.lib.pt_unet_set_adam(self.r.ctx, m.ctypes.data, v.ctypes.data, m.size,
                                                  int(np.asarray(st["steps"]).reshape(-1)[0])))

    @staticmethod
    def _nhwc(a, c):
        a = np.ascontiguousarray(a, np.float32)
        if a.ndim != 4 or a.shape[3] != c:
            raise ValueError("expected [n][h][w][%d], got %s" % (c, a.shape))
        return a

    def forward(self, feat):
        feat = self._nhwc(feat, self.N_FEATURES)
        n, h, w, _ = feat.shape
        pred = np.empty((n, h, w, 3), np.float32)
        self.r._check(self.r.lib.pt_unet_forward(self.r.ctx, feat.ctypes.data, n, h, w, pred.ctypes.data))
        return pred

    def loss_grad(self, feat, target, with_grad=True):
        feat, target = self._nhwc(feat, self.N_FEATURES), self._nhwc(target, 3)
        n, h, w, _ = feat.shape
        if target.shape[:3] != (n, h, w):
            raise ValueError("target shape %s does not match features %s" % (target.shape, feat.shape))
        loss = C.c_float()
        grad = np.empty(self.param_count(), np.float32) if with_grad else None
        self.r._check(self.r.lib.pt_unet_loss_grad(self.r.ctx, feat.ctypes.data, target.ctypes.data, n, h, w,
                                                   C.byref(loss), grad.ctypes.data if with_grad else None))
        return loss.value, grad

    def set_dataset(self, feat, target):
        feat, target = self._nhwc(feat, self.N_FEATURES), self._nhwc(target, 3)
        n, h, w, _ = feat.shape
        self.r._check(self.r.lib.pt_unet_set_dataset(self.r.ctx, feat.ctypes.data, target.ctypes.data, n, h, w))

    def dataset_reserve(self, n, h, w):
        """Device-side training set of n crops of h x w, filled by capture()."""
        self.r._check(self.r.lib.pt_unet_dataset_reserve(self.r.ctx, n, h, w))

    def capture(self, index, target, x0=0, y0=0):
        """Copies the window at (x0, y0) of the renderer's current frame into
        crop `index`: the network input (radiance + G-buffer) or, with
        target=True, the reference radiance."""
        what = A.PT_UNET_TARGET if target else A.PT_UNET_INPUT
        self.r._check(self.r.lib.pt_unet_dataset_capture(self.r.ctx, index, what, x0, y0))

    def train(self, steps, batch, lr=1e-3, seed=0):
        losses = np.empty(steps, np.float32)
        self.r._check(self.r.lib.pt_unet_train(self.r.ctx, steps, batch, lr, seed, losses.ctypes.data))
        return losses

    def execute(self):
        """Denoises the renderer's current frame; returns the rgba8 output."""
        self.r._check(self.r.lib.pt_unet_denoise(self.r.ctx, self.exposure))
        return self.r.read_output()

    def read_result(self):
        a = np.empty((self.r.height, self.r.width, 4), np.float32)
        self.r._check(self.r.lib.pt_read_denoised(self.r.ctx, a.ctypes.data, a.nbytes))
        return a

    def set_precision(self, fp8):
        """fp8=True: calibrate on the renderer's current frame and run
        inference on fp8 products (pt_unet_set_precision); False: bf16."""
        self.r._check(self.r.lib.pt_unet_set_precision(self.r.ctx, A.PT_UNET_FP8 if fp8 else A.PT_UNET_BF16))

    def fp8_state(self, n_act, n_out, n_weights):
        """(input scales, weight scales, e4m3 weight bytes) of the fp8 mode."""
        sa, sw = np.empty(n_act, np.float32), np.empty(n_out, np.float32)
        w8 = np.empty(n_weights, np.uint8)
        self.r._check(self.r.lib.pt_unet_fp8_state(self.r.ctx, sa.ctypes.data, n_act, sw.ctypes.data, n_out,
                                                    w8.ctypes.data, n_weights))
        return sa, sw, w8

    def frame_features(self):
        """The renderer's current frame as network features [1][h][w][10]
        (accumulated radiance, G-buffer albedo, normal and hit distance)."""
        acc = self.r.read_accum()
        g = self.r.read_gbuffer()
        h, w = self.r.height, self.r.width
        f = np.empty((1, h, w, self.N_FEATURES), np.float32)
        f[0, :, :, 0:3] = acc.reshape(h, w, 4)[:, :, :3]
        gv = g.reshape(h, w, 12)
        f[0, :, :, 3:6] = gv[:, :, 4:7]
        f[0, :, :, 6:9] = gv[:, :, 0:3]
        f[0, :, :, 9] = gv[:, :, 11]
        return f


class Batch:
    """Several Renderers (each with its own scene, frame and options) rendered
    one progressive frame at a time as one graph launch (pt_batch.h)."""

    def __init__(self, device=0):
        self.lib = A.load_pt()
        self.h = C.c_void_p()
        rc = self.lib.pt_batch_create(C.byref(self.h), device)
        if rc:
            raise A.PTError(rc, "pt_batch_create failed")
        self.jobs = []

    def _check(self, rc):
        if rc:
            raise A.PTError(rc, (self.lib.pt_batch_last_error(self.h) or b"").decode())
        return rc

    def close(self):
        if self.h:
            self.lib.pt_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, renderer):
        idx = A.u32()
        self._check(self.lib.pt_batch_add(self.h, renderer.ctx, C.byref(idx)))
        self.jobs.append(renderer)
        return idx.value

    def frame(self, spp, base_seeds=None):
        """One frame: job i renders spp[i] passes from its own progressive
        sample count (seeds as Renderer.render derives them from base_seeds[i]).
        Returns the per-job pt_stats."""
        n = len(self.jobs)
        params = (A.SceneParams * n)()
        passes = (A.u32 * n)(*spp)
        seed_arrays = []
        seed_ptrs = (C.c_void_p * n)()
        for i, r in enumerate(self.jobs):
            bs = 0 if base_seeds is None else base_seeds[i]
            s = np.array([(bs * 0x9E3779B1 + (r.frame + k) * 0x85EBCA77) & 0xFFFFFFFF for k in range(spp[i])],
                         np.uint32)
            params[i] = r.params(seed=int(s[0]) if spp[i] else 0)
            seed_arrays.append(s)
            seed_ptrs[i] = s.ctypes.data if spp[i] else None
        stats = (A.Stats * n)()
        self._check(self.lib.pt_batch_frame(self.h, params, passes, seed_ptrs, stats))
        for i, r in enumerate(self.jobs):
            r.frame += spp[i]
        return list(stats)

    def graph_info(self):
        f, i, u = A.u32(), A.u32(), A.u32()
        self._check(self.lib.pt_batch_graph_info(self.h, C.byref(f), C.byref(i), C.byref(u)))
        return f.value, i.value, u.value
