"""ctypes wrapper of oracle/build/librefcpu.so (see refcpu.h).

TEST INFRASTRUCTURE ONLY: the product never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from webgpuraytracer_amd import _abi as A

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "librefcpu.so")
P, vp = C.POINTER, C.c_void_p
FUNCS = {
    "ref_scene_create": (C.c_int, [P(vp), P(A.SceneDesc)]),
    "ref_scene_destroy": (None, [vp]),
    "ref_trace_rays": (C.c_int, [vp, vp, A.u32, vp, P(A.Stats)]),
    "ref_trace_shadow_rays": (C.c_int, [vp, vp, vp, vp, A.u32, vp]),
    "ref_render": (C.c_int, [vp, P(A.SceneParams), A.u32, A.u32, A.u32, vp, vp, vp, P(A.Stats), C.c_int, A.u32]),
    "ref_trace_hits": (C.c_int, [vp, P(A.SceneParams), A.u32, vp, C.c_int]),
    "ref_trace_hits_rays": (C.c_int, [vp, P(A.SceneParams), A.u32, vp, vp, C.c_int]),
    "ref_gbuffer": (C.c_int, [vp, P(A.SceneParams), vp, vp, C.c_int]),
    "ref_denoise": (C.c_int, [P(A.DenoiseParams), C.c_int, vp, vp, vp, vp]),
    "ref_build_sky_cdf": (C.c_int, [vp, A.u32, A.u32, A.f32, vp, P(C.c_double)]),
    "ref64_closest": (C.c_int, [vp, P(C.c_double), P(A.Hit), P(C.c_double)]),
    "ref_env_sample": (C.c_int, [vp, P(A.SceneParams), A.u32, A.u32, vp]),
    "ref_env_pdf": (C.c_int, [vp, P(A.SceneParams), vp, A.u32, vp]),
    "ref_bsdf_sample": (C.c_int, [P(A.Material), P(A.f32), P(A.f32), A.u32, A.u32, vp]),
    "ref_bsdf_eval": (C.c_int, [P(A.Material), P(A.f32), P(A.f32), vp, A.u32, vp, vp]),
    "ref_pcg": (C.c_int, [A.u32, A.u32, vp, vp]),
    "ref_camera_rays": (C.c_int, [P(A.SceneParams), A.u32, vp, C.c_int, vp, vp]),
    "ref_bsdf_terms": (C.c_int, [C.c_int, vp, A.u32, vp]),
    "ref_light_sample": (C.c_int, [vp, A.u32, P(A.f32), A.u32, A.u32, vp]),
    "ref_light_pdf": (C.c_int, [vp, A.u32, P(A.f32), vp, A.u32, vp]),
}
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        for n, (r, a) in FUNCS.items():
            f = getattr(_lib, n)
            f.restype, f.argtypes = r, a
    return _lib


def _ok(rc, what):
    if rc:
        raise RuntimeError("%s failed: %d" % (what, rc))


class RefScene:
    def __init__(self, desc):
        self.L = lib()
        self.h = C.c_void_p()
        _ok(self.L.ref_scene_create(C.byref(self.h), C.byref(desc)), "ref_scene_create")

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ref_scene_destroy(self.h)
            self.h = None

    def trace_rays(self, rays):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        out = (A.Hit * r.shape[0])()
        st = A.Stats()
        _ok(self.L.ref_trace_rays(self.h, r.ctypes.data, r.shape[0], out, C.byref(st)), "ref_trace_rays")
        return out, st

    def trace_shadow_rays(self, rays, target_idx, target_dist):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        t = np.ascontiguousarray(target_idx, np.int32)
        d = np.ascontiguousarray(target_dist, np.float32)
        out = np.empty((r.shape[0], 3), np.float32)
        _ok(self.L.ref_trace_shadow_rays(self.h, r.ctypes.data, t.ctypes.data, d.ctypes.data, r.shape[0],
                                         out.ctypes.data), "ref_trace_shadow_rays")
        return out

    def render(self, params, passes=1, seeds=None, accum=None, section=(0, 0), threads=0, row_stride=1,
               with_output=True):
        w, h = params.width, params.height
        acc = np.zeros((h, w, 4), np.float32) if accum is None else np.ascontiguousarray(accum, np.float32).copy()
        out = np.zeros((h, w, 4), np.uint8) if with_output else None
        s = None if seeds is None else np.ascontiguousarray(seeds, np.uint32)
        st = A.Stats()
        _ok(self.L.ref_render(self.h, C.byref(params), section[0], section[1], passes,
                              None if s is None else s.ctypes.data, acc.ctypes.data,
                              None if out is None else out.ctypes.data, C.byref(st), threads, row_stride),
            "ref_render")
        return acc, out, st

    def trace_hits(self, params, n_bounces=2, threads=0, rays=False):
        """Hits of the first n_bounces trace_scene calls per pixel; with
        rays=True also the rays traced ([pixels * n_bounces][6] float32)."""
        out = (A.Hit * (params.width * params.height * n_bounces))()
        if not rays:
            _ok(self.L.ref_trace_hits(self.h, C.byref(params), n_bounces, out, threads), "ref_trace_hits")
            return out
        r = np.zeros((params.width * params.height * n_bounces, 6), np.float32)
        _ok(self.L.ref_trace_hits_rays(self.h, C.byref(params), n_bounces, out, r.ctypes.data, threads),
            "ref_trace_hits_rays")
        return out, r

    def gbuffer(self, params, threads=0):
        g = np.zeros((params.height, params.width, 12), np.float32)
        o = np.zeros((params.height, params.width, 4), np.uint8)
        _ok(self.L.ref_gbuffer(self.h, C.byref(params), g.ctypes.data, o.ctypes.data, threads), "ref_gbuffer")
        return g, o

    def env_sample(self, params, n, seed=1):
        out = np.zeros((n, 7), np.float32)
        _ok(self.L.ref_env_sample(self.h, C.byref(params), n, seed, out.ctypes.data), "ref_env_sample")
        return out

    def env_pdf(self, params, dirs):
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        out = np.zeros(d.shape[0], np.float32)
        _ok(self.L.ref_env_pdf(self.h, C.byref(params), d.ctypes.data, d.shape[0], out.ctypes.data), "ref_env_pdf")
        return out

    def light_sample(self, light, pos, n, seed=1):
        out = np.zeros((n, 8), np.float32)
        _ok(self.L.ref_light_sample(self.h, light, A.fvec(*pos), n, seed, out.ctypes.data), "ref_light_sample")
        return out

    def light_pdf(self, light, pos, dirs):
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        out = np.zeros(d.shape[0], np.float32)
        _ok(self.L.ref_light_pdf(self.h, light, A.fvec(*pos), d.ctypes.data, d.shape[0], out.ctypes.data),
            "ref_light_pdf")
        return out

    def closest64(self, ray):
        r = (C.c_double * 6)(*[float(x) for x in ray])
        hit = A.Hit()
        t = C.c_double()
        rc = self.L.ref64_closest(self.h, r, C.byref(hit), C.byref(t))
        return rc, hit, t.value


def denoise(u, accum, gbuf, stride_mode=0):
    a = np.ascontiguousarray(accum, np.float32)
    g = np.ascontiguousarray(gbuf, np.float32)
    out = np.zeros(a.shape, np.float32)
    o8 = np.zeros(a.shape, np.uint8)
    _ok(lib().ref_denoise(C.byref(u), stride_mode, a.ctypes.data, g.ctypes.data, out.ctypes.data, o8.ctypes.data),
        "ref_denoise")
    return out, o8


def build_sky_cdf(rgba16f, exposure=1.0):
    t = np.ascontiguousarray(rgba16f, np.uint16)
    h, w = t.shape[:2]
    cdf = np.zeros(h + 1 + h * (w + 1), np.float32)
    tot = C.c_double()
    _ok(lib().ref_build_sky_cdf(t.ctypes.data, w, h, exposure, cdf.ctypes.data, C.byref(tot)), "ref_build_sky_cdf")
    return cdf, tot.value


def bsdf_sample(material, n, v, count, seed=1):
    out = np.zeros((count, 8), np.float32)
    _ok(lib().ref_bsdf_sample(C.byref(material), A.fvec(*n), A.fvec(*v), count, seed, out.ctypes.data),
        "ref_bsdf_sample")
    return out


def bsdf_eval(material, n, v, L):
    L = np.ascontiguousarray(L, np.float32).reshape(-1, 3)
    f = np.zeros((L.shape[0], 3), np.float32)
    pdf = np.zeros(L.shape[0], np.float32)
    _ok(lib().ref_bsdf_eval(C.byref(material), A.fvec(*n), A.fvec(*v), L.ctypes.data, L.shape[0], f.ctypes.data,
                            pdf.ctypes.data), "ref_bsdf_eval")
    return f, pdf


def pcg(state, n):
    """rand_pcg from `state`, n draws: (output words uint32, values float32)."""
    w = np.zeros(n, np.uint32)
    v = np.zeros(n, np.float32)
    _ok(lib().ref_pcg(state & 0xFFFFFFFF, n, w.ctypes.data, v.ctypes.data), "ref_pcg")
    return w, v


def camera_rays(params, pixels, centre=False):
    """Camera rays of `main` (jittered) or gbuf_main (centre) for (x, y)
    pixel pairs: (rays [n][6] float32, jitter [n][2] float32)."""
    px = np.ascontiguousarray(pixels, np.uint32).reshape(-1, 2)
    rays = np.zeros((px.shape[0], 6), np.float32)
    jit = np.zeros((px.shape[0], 2), np.float32)
    _ok(lib().ref_camera_rays(C.byref(params), px.shape[0], px.ctypes.data, int(centre), rays.ctypes.data,
                              jit.ctypes.data), "ref_camera_rays")
    return rays, jit


BSDF_TERMS = {"D_GGX": 0, "fresnel_dielectric": 1, "G_Smith_GGX_div_NV": 2, "G_Smith_GGX_Uncorrelated_div_NV": 3,
              "G1_GGX_div_NV": 4, "mis_weight": 5}


def bsdf_term(name, *args):
    """One of the oracle's BSDF building blocks on broadcast float32 arguments."""
    cols = np.broadcast_arrays(*[np.asarray(a, np.float32) for a in args])
    shape = cols[0].shape
    a = np.zeros((cols[0].size, 3), np.float32)
    for i, c in enumerate(cols):
        a[:, i] = c.ravel()
    out = np.zeros(a.shape[0], np.float32)
    _ok(lib().ref_bsdf_terms(BSDF_TERMS[name], a.ctypes.data, a.shape[0], out.ctypes.data), "ref_bsdf_terms")
    return out.reshape(shape)
