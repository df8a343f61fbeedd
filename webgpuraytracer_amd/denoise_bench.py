"""Time of the a-trous denoiser (Denoiser.execute, tracer.js:2583-2637) and
of the learned denoiser on a 1920x1080 frame:

    python -m webgpuraytracer_amd.denoise_bench [--size 1920x1080] [--passes 5] [--calls 20] [--unet]

The frame is a 1-spp render of a seeded small-analytic scene with its
G-buffer; each pt_denoise call runs the guide, the prepass, `passes` a-trous
passes with strides 1, 2, 4, ... and the composite.  Prints one JSON line per
weight evaluation (fused, exact): milliseconds per pt_denoise call from the
host clock around synchronous calls (no readback), after one untimed call.
With --unet, also one line per precision of pt_unet_denoise (bf16, fp8
calibrated on the same frame) with the shipped weights.
"""
import argparse
import ctypes as C
import json
import sys
import time

from . import _abi as A
from .renderer import Denoiser, LearnedDenoiser, Renderer
from .scene import SceneBuilder


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--size", default="1920x1080")
    ap.add_argument("--passes", type=int, default=5)
    ap.add_argument("--calls", type=int, default=20)
    ap.add_argument("--unet", action="store_true", help="also time the learned denoiser (bf16 and fp8)")
    ap.add_argument("--no-atrous", action="store_true")
    a = ap.parse_args(argv)
    w, h = (int(x) for x in a.size.split("x"))
    r = Renderer(0)
    scene = SceneBuilder.config(1, seed=5)
    r.set_scene(scene, isgbuf=True)
    r.resize(w, h)
    p = scene.params(w, h, 0, 1)
    r.gbuffer(p)
    r.render_section(p, passes=1)
    dn = Denoiser(r)
    dn.passes = a.passes
    for name, math in () if a.no_atrous else (("fused", A.PT_DENOISE_MATH_FUSED), ("exact", A.PT_DENOISE_MATH_EXACT)):
        r.set_option(A.PT_OPT_DENOISE_MATH, math)
        u = dn.uniforms()
        r._check(r.lib.pt_denoise(r.ctx, C.byref(u), 0))
        t0 = time.perf_counter()
        for _ in range(a.calls):
            r._check(r.lib.pt_denoise(r.ctx, C.byref(u), 0))
        ms = (time.perf_counter() - t0) * 1e3 / a.calls
        print(json.dumps({"math": name, "size": a.size, "passes": a.passes, "calls": a.calls,
                          "ms_per_call": round(ms, 3)}), flush=True)
    r.set_option(A.PT_OPT_DENOISE_MATH, A.PT_DENOISE_MATH_FUSED)
    if a.unet:
        net = LearnedDenoiser(r, pretrained=True)
        for name, fp8 in (("unet_bf16", False), ("unet_fp8", True)):
            net.set_precision(fp8)
            r._check(r.lib.pt_unet_denoise(r.ctx, net.exposure))
            r.lib.pt_synchronize(r.ctx)
            t0 = time.perf_counter()
            for _ in range(a.calls):
                r._check(r.lib.pt_unet_denoise(r.ctx, net.exposure))
            ms = (time.perf_counter() - t0) * 1e3 / a.calls
            print(json.dumps({"denoiser": name, "size": a.size, "calls": a.calls, "ms_per_call": round(ms, 3)}),
                  flush=True)
    r.close()


if __name__ == "__main__":
    sys.exit(main())
