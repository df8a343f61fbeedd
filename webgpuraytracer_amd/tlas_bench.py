"""The top-level tree built on the device (PT_TLAS_BUILD_DEVICE) against the
host-built one, at growing object counts:

    python -m webgpuraytracer_amd.tlas_bench [--counts 1000,10000,100000] [--size 512] [--spp 4]

Scene per count N: N spheres (radius U[0.02, 0.1] * (10^5 / N)^(1/3), seeded
positions in a cube of side 8), a ground plane and the procedural sky; so
the objects fill a similar fraction of space at every N.  Measured per N:
  host_build_s   SceneBuilder.finalize (the tracer.js-style median-split TLAS
                 and packing): what a caller pays per update without the device build
  update_ref_s   pt_update_objects with that TLAS (host conversion + upload)
  update_dev_s   pt_update_objects without a TLAS, the device building it
  mrays_*        trace rate of a spp-pass frame at size^2 with either tree
Prints one JSON line.
"""
import argparse
import ctypes as C
import json
import sys
import time

import numpy as np

from . import _abi as A
from .renderer import Renderer
from .scene import SceneBuilder


def sphere_scene(n, seed=5):
    rng = np.random.default_rng(seed)
    b = SceneBuilder()
    mats = [b.material(color=tuple(rng.uniform(0.2, 0.9, 3)), roughness=float(rng.uniform(0.1, 0.9)))
            for _ in range(8)]
    b.plane(mats[0], (0.0, 1.0, 0.0), -4.5)
    scale = (1e5 / n) ** (1.0 / 3.0)
    pos = rng.uniform(-4.0, 4.0, (n, 3))
    rad = rng.uniform(0.02, 0.1, n) * scale
    for i in range(n):
        b.sphere(mats[i % 8], tuple(float(x) for x in pos[i]), float(rad[i]))
    b.procedural_sky(256, 128, seed=seed)
    b.camera(eye=(0.0, 0.0, 12.0), target=(0.0, 0.0, 0.0), fov=45.0)
    b.bounces(4)
    t0 = time.perf_counter()
    b.finalize()
    return b, time.perf_counter() - t0


def _rate(r, scene, size, spp):
    r.reset()
    r.render(spp=1, base_seed=3)
    r.reset()
    t0 = time.perf_counter()
    st = r.render(spp=spp, base_seed=7)
    dt = time.perf_counter() - t0
    return (st.extension_rays + st.shadow_rays) / dt * 1e-6


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--counts", type=lambda s: [int(x) for x in s.split(",")], default=[1000, 10000, 100000])
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--spp", type=int, default=4)
    a = ap.parse_args(argv)
    rows = []
    for n in a.counts:
        scene, host_s = sphere_scene(n)
        desc = scene.desc()
        bare = scene.desc()
        bare.tlas.count = 0
        row = {"objects": n, "tlas_nodes": int(desc.tlas.count), "host_build_s": host_s}
        r = Renderer(0)
        r.set_option(A.PT_OPT_COUNT_STATS, 0)
        for build, d, key in ((A.PT_TLAS_BUILD_REFERENCE, desc, "ref"), (A.PT_TLAS_BUILD_DEVICE, bare, "dev")):
            r.set_option(A.PT_OPT_TLAS_BUILD, build)
            r.set_scene(d)
            r.scene = scene  # camera parameters come from the builder
            r.resize(a.size, a.size)
            times = []
            for _ in range(3):
                t0 = time.perf_counter()
                r._check(r.lib.pt_update_objects(r.ctx, C.byref(d)))
                times.append(time.perf_counter() - t0)
            row["update_%s_s" % key] = float(np.median(times))
            row["mrays_%s" % key] = _rate(r, scene, a.size, a.spp)
        r.close()
        print(json.dumps(row), file=sys.stderr, flush=True)
        rows.append(row)
    print(json.dumps({"workload": "device-built TLAS vs host-built, spheres", "size": a.size, "spp": a.spp,
                      "rows": rows}))


if __name__ == "__main__":
    sys.exit(main())
