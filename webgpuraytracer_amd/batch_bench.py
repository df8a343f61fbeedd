"""Many render jobs per graph launch, on the batched-jobs workload:

    python -m webgpuraytracer_amd.batch_bench [--jobs 32] [--frames 4] [--seed 1] [--tune]

Job j (seeded): a square frame whose pixel count is log-uniform in
[128^2, 1024^2]; spp per frame from {1, 4, 16, 64}; bounce limit uniform in
2..12; the small analytic scene (12 primitives, emissive quad, ground plane,
512x256 procedural sky) plus instanced meshes: unique triangles
log-uniform in [1k, 1M] (four displaced geodesic spheres of equal
frequency), instances log-uniform in 1..128, placed among the primitives.

Measured: the same frames rendered job after job with pt_trace (one host
wait per job), then as batch frames (pt_batch_frame: one graph launch and one
host wait per frame), optionally after per-job option tuning (autotune.py).
Prints one JSON line: aggregate and per-job rays, Mrays/s of both ways.
"""
import argparse
import json
import math
import sys
import time

import numpy as np

from . import _abi as A
from .autotune import TuneCache, describe
from .renderer import Batch, Renderer
from .scene import SceneBuilder


def job_specs(n, seed):
    rng = np.random.default_rng(seed)
    specs = []
    for j in range(n):
        px = math.exp(rng.uniform(math.log(128 * 128), math.log(1024 * 1024)))
        side = max(16, int(round(math.sqrt(px) / 8.0)) * 8)
        spp = int(rng.choice([1, 4, 16, 64]))
        bounces = int(rng.integers(2, 13))
        tris = math.exp(rng.uniform(math.log(1e3), math.log(1e6)))
        freq = max(2, int(round(math.sqrt(tris / 80.0))))  # 4 meshes x 20 freq^2 triangles
        inst = int(round(math.exp(rng.uniform(0.0, math.log(128.0)))))
        specs.append({"job": j, "width": side, "height": side, "spp": spp, "bounces": bounces,
                      "triangles": 80 * freq * freq, "freq": freq, "instances": max(1, inst),
                      "scene_seed": int(seed * 1000 + j)})
    return specs


def build(spec):
    scene = SceneBuilder.config(3, seed=spec["scene_seed"], mesh_freq=[spec["freq"]] * 4,
                                n_instances=spec["instances"], bounces=spec["bounces"])
    r = Renderer(0)
    r.set_option(A.PT_OPT_COUNT_STATS, 0)
    r.set_option(A.PT_OPT_WAVEFRONT_POOL, 1 << 22)
    r.set_scene(scene)
    r.resize(spec["width"], spec["height"])
    return r, scene


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--jobs", type=int, default=32)
    ap.add_argument("--frames", type=int, default=4)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--tune", action="store_true")
    ap.add_argument("--mode", choices=["both", "batched", "job-after-job"], default="both",
                    help="which way to render (a profile of one way takes one)")
    a = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    specs = job_specs(a.jobs, a.seed)
    t0 = time.perf_counter()
    jobs = []
    for s in specs:
        jobs.append(build(s))
    log("built %d jobs in %.1f s" % (len(jobs), time.perf_counter() - t0))
    rs = [r for r, _ in jobs]
    spp = [s["spp"] for s in specs]
    tuned = []
    if a.tune:
        cache = TuneCache()
        for s, (r, scene) in zip(specs, jobs):
            log("  job %d: %dx%d, %d spp, %d bounces, %d tris x %d" % (s["job"], s["width"], s["height"], s["spp"],
                                                                     s["bounces"], s["triangles"], s["instances"]))
            opts = cache.get(r, scene.desc().flags, s["triangles"], s["spp"], log=log)
            tuned.append(describe(opts))
    # job after job: pt_trace per job, one host wait each
    seq_t, seq_rays = [], 0
    per_job, seq_job_rays = [0.0] * len(rs), [0] * len(rs)
    for r, s in zip(rs, specs):
        r.reset()
        if a.mode != "batched":
            r.render(spp=s["spp"], base_seed=11)
    for f in range(a.frames if a.mode != "batched" else 0):
        t = time.perf_counter()
        for i, (r, s) in enumerate(zip(rs, specs)):
            tj = time.perf_counter()
            st = r.render(spp=s["spp"], base_seed=11)
            per_job[i] += time.perf_counter() - tj
            seq_job_rays[i] += st.extension_rays + st.shadow_rays
            seq_rays += st.extension_rays + st.shadow_rays
        seq_t.append(time.perf_counter() - t)
    # batched: one graph launch per frame
    b = Batch(0)
    for r in rs:
        r.reset()
        b.add(r)
    if a.mode != "job-after-job":
        b.frame(spp, base_seeds=[11] * len(rs))  # records and instantiates the graph
    bat_t, bat_rays, job_rays = [], 0, [0] * len(rs)
    for f in range(a.frames if a.mode != "job-after-job" else 0):
        t = time.perf_counter()
        st = b.frame(spp, base_seeds=[11] * len(rs))
        bat_t.append(time.perf_counter() - t)
        for i, x in enumerate(st):
            job_rays[i] += x.extension_rays + x.shadow_rays
        bat_rays += sum(x.extension_rays + x.shadow_rays for x in st)
    frames, inst, upd = b.graph_info()
    rate = lambda rays, t: rays / sum(t) * 1e-6 if t else None  # noqa: E731
    per_frame = lambda t: 1e3 * sum(t) / len(t) if t else None  # noqa: E731
    out = {
        "metric": "aggregate Mrays/s, batched jobs", "jobs": len(rs), "frames": a.frames,
        "mrays_per_s_batched": rate(bat_rays, bat_t),
        "mrays_per_s_job_after_job": rate(seq_rays, seq_t),
        "ms_per_frame_batched": per_frame(bat_t),
        "ms_per_frame_job_after_job": per_frame(seq_t),
        "graph": {"frames": frames, "instantiations": inst, "updates": upd},
        "tuned": a.tune,
        "per_job": [{"job": s["job"], "res": s["width"], "spp": s["spp"], "bounces": s["bounces"],
                     "triangles": s["triangles"], "instances": s["instances"],
                     "mrays_per_frame": (job_rays[i] or seq_job_rays[i]) / max(a.frames, 1) * 1e-6,
                     "mrays_per_s_alone": seq_job_rays[i] / per_job[i] * 1e-6 if per_job[i] else None,
                     "options": tuned[i] if tuned else None} for i, s in enumerate(specs)],
    }
    print(json.dumps(out))
    b.close()
    for r in rs:
        r.close()


if __name__ == "__main__":
    sys.exit(main())
