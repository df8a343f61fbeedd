#!/usr/bin/env python3
"""Throughput of the tracer hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): 1920x1080,
64 spp, BOUNCE_LIMIT 8, 64 instances of 4 noise-displaced geodesic spheres
(808,500 triangles, two-level TLAS/BLAS), 2048x1024 HDR environment with
importance-sampled NEE.  One step = one 64-spp progressive frame (one
pt_trace call with 64 passes) with the scene resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Prints one JSON line.  Mrays/s = (extension rays + shadow rays) / time, rays
counted on the device exactly as the reference issues them (one trace_scene
call = one extension ray, one trace_scene_shadow call = one shadow ray).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s @1080p 8-bounce; equal-time relMSE vs 16k-spp ref (learned vs à-trous)"
W, H, SPP = 1920, 1080, 64
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(st, sky_w, sky_h):
    """Bytes the traversal and shading must touch, counted on the reference-order traversal:
    32 per node pop, +48 per inner pop (two child boxes), 96 per TLAS-leaf
    object, 36 per triangle test, and per shading event 144 (material) + 96
    (object) + 80 (winning triangle's normals/uvs) + 2 sky lookups x 4 texels
    x 8 B + CDF binary-search reads; plus 36 B of accum/output per path."""
    cdf_reads = (math.ceil(math.log2(max(sky_h, 2))) + math.ceil(math.log2(max(sky_w, 2)))) * 4
    shade = 144 + 96 + 80 + 2 * 4 * 8 + cdf_reads
    return (32 * st.node_pops + 48 * st.inner_pops + 96 * st.object_visits + 36 * st.tri_tests +
            shade * st.shade_events + 36 * st.paths)


def cpu_model():
    """The host CPU's model name (lscpu's "Model name", read from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.lower().startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(scene, params, budget_s):
    """The oracle (a CPU restatement of shader.wgsl `main`) on the same scene,
    1 spp on every 8th row, repeated until budget_s of work; all host threads."""
    from oracle.refcpu import RefScene
    ref = RefScene(scene.desc())
    threads = os.cpu_count() or 1
    rays = 0
    t0 = time.perf_counter()
    passes = 0
    while True:
        params.sample_number = passes
        params.seed = 1000 + passes
        _, _, st = ref.render(params, passes=1, threads=threads, row_stride=8, with_output=False)
        rays += st.extension_rays + st.shadow_rays
        passes += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": rays / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port", "cpu_model": cpu_model(),
            "sample": "oracle FP32 restatement, same scene and camera, 1 spp x %d passes over every 8th row "
                      "of 1920x1080 (%d rays in %.1f s)" % (passes, rays, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-count", action="store_true",
                    help="skip the extra counted step that measures node/triangle work (no roofline)")
    ap.add_argument("--no-bounces", action="store_true",
                    help="skip the extra step that times each bounce (wavefront backend)")
    ap.add_argument("--backend", choices=["megakernel", "wavefront"], default="wavefront")
    ap.add_argument("--schedule", choices=["regen", "sync"], default="regen",
                    help="megakernel pass schedule (same bits either way)")
    ap.add_argument("--traversal", choices=["wide", "quant", "reference"], default="wide",
                    help="BVH traversal: 4-wide nodes, 4-wide nodes with 8-bit quantised boxes, or the "
                         "reference's child-pair order")
    ap.add_argument("--blas-build", choices=["reference", "device"], default="reference",
                    help="bottom-level trees of the wide traversal: the uploaded BVH, or rebuilt on the GPU")
    ap.add_argument("--tlas-build", choices=["reference", "device"], default="reference",
                    help="top-level tree: the uploaded TLAS, or built on the GPU from the objects "
                         "(counted work is then of that tree)")
    ap.add_argument("--lds-stack", type=int, default=None,
                    help="traversal-stack entries per lane kept in LDS (PT_OPT_LDS_STACK; default the library's)")
    ap.add_argument("--refill", type=int, default=None,
                    help="idle lanes that make a traversal wave fetch new rays (PT_OPT_WAVEFRONT_REFILL; "
                         "default the library's)")
    ap.add_argument("--compare", action="store_true",
                    help="also time the other backend/schedule variants (reported under alt)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (a frame would shard by sample index "
              "across GPUs, each keeping a running sum and count, merged once at the end); run with --gpus 1",
              file=sys.stderr)
        return 2

    from webgpuraytracer_amd import Renderer, SceneBuilder
    from webgpuraytracer_amd import _abi as A

    t0 = time.perf_counter()
    scene = SceneBuilder.config(2, seed=args.seed)
    build_s = time.perf_counter() - t0
    r = Renderer(0)
    if args.blas_build == "device":
        r.set_option(A.PT_OPT_BLAS_BUILD, A.PT_BLAS_BUILD_DEVICE)
    if args.tlas_build == "device":
        r.set_option(A.PT_OPT_TLAS_BUILD, A.PT_TLAS_BUILD_DEVICE)
    t0 = time.perf_counter()
    r.set_scene(scene)
    r.resize(W, H)
    upload_s = time.perf_counter() - t0
    # Timed steps run the production kernel; node/triangle counters (which
    # cost registers) are switched on only for one extra step afterwards.
    r.set_option(A.PT_OPT_COUNT_STATS, 0)
    if args.lds_stack is not None:
        r.set_option(A.PT_OPT_LDS_STACK, args.lds_stack)
    if args.refill is not None:
        r.set_option(A.PT_OPT_WAVEFRONT_REFILL, args.refill)
    d = scene.desc()
    mk = (A.PT_OPT_BACKEND, A.PT_BACKEND_MEGAKERNEL)
    regen, sync = (A.PT_OPT_SCHEDULE, A.PT_SCHEDULE_REGEN), (A.PT_OPT_SCHEDULE, A.PT_SCHEDULE_SYNC)
    trav = {"wide": (A.PT_OPT_TRAVERSAL, A.PT_TRAVERSAL_WIDE),
            "quant": (A.PT_OPT_TRAVERSAL, A.PT_TRAVERSAL_WIDE_QUANTIZED),
            "reference": (A.PT_OPT_TRAVERSAL, A.PT_TRAVERSAL_REFERENCE)}
    wf = (A.PT_OPT_BACKEND, A.PT_BACKEND_WAVEFRONT)
    variants = {"wavefront/wide": [wf, trav["wide"]],
                "wavefront/quant": [wf, trav["quant"]],
                "wavefront/reference": [wf, trav["reference"]],
                "megakernel/regen/wide": [mk, regen, trav["wide"]],
                "megakernel/regen/quant": [mk, regen, trav["quant"]],
                "megakernel/regen/reference": [mk, regen, trav["reference"]],
                "megakernel/sync/wide": [mk, sync, trav["wide"]]}
    if args.backend == "wavefront":
        chosen = "wavefront/" + args.traversal
    else:
        chosen = "megakernel/%s/%s" % (args.schedule, args.traversal)
    if chosen not in variants:
        variants[chosen] = [mk, sync if args.schedule == "sync" else regen, trav[args.traversal]]

    def step(k):
        seeds = [(args.seed * 0x9E3779B1 + (k * SPP + i) * 0x85EBCA77) & 0xFFFFFFFF for i in range(SPP)]
        p = scene.params(W, H, sample_number=0, seed=seeds[0])
        return r.render_section(p, passes=SPP, seeds=seeds)

    def timed(variant):
        for key, val in variants[variant]:
            r.set_option(key, val)
        for k in range(args.warmup):
            step(1000 + k)
        r.lib.pt_synchronize(r.ctx)
        st = []
        t0 = time.perf_counter()
        for k in range(args.steps):
            st.append(step(k))
            print("step %d/%d done" % (k + 1, args.steps), file=sys.stderr, flush=True)
        r.lib.pt_synchronize(r.ctx)
        return st, time.perf_counter() - t0

    stats, wall = timed(chosen)

    rays = sum(s.extension_rays + s.shadow_rays for s in stats)
    kern_ms = sum(s.kernel_ms for s in stats)
    value = rays / wall / 1e6
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall * 1000 / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plain tracer, large scene: 1920x1080, 64 spp/step, 8 bounces, 64 instances of 4 "
                               "displaced geodesic spheres (%d triangles), 2048x1024 env, %s"
                               % (d.triangles.count, chosen + ("" if args.blas_build == "reference"
                                                               else ", BLAS built on the device")
                                  + ("" if args.tlas_build == "reference" else ", TLAS built on the device")),
                   "extension_rays": int(sum(s.extension_rays for s in stats)),
                   "shadow_rays": int(sum(s.shadow_rays for s in stats)),
                   "paths": int(sum(s.paths for s in stats)),
                   "kernel_mrays_s": round(rays / (kern_ms / 1000) / 1e6, 2) if kern_ms > 0 else None,
                   "scene_build_s": round(build_s, 3), "upload_s": round(upload_s, 3)},
    }
    if args.backend == "wavefront" and not args.no_bounces:
        # The same launch as timed step 0 with an event before each bounce:
        # rays and device time per bounce (bounce 0: the coherent camera rays
        # and their NEE; later bounces: incoherent secondary rays).
        r.set_option(A.PT_OPT_BOUNCE_TIMING, 1)
        step(0)
        r.set_option(A.PT_OPT_BOUNCE_TIMING, 0)
        ext, sh, ms = r.bounce_stats(16)
        rows = []
        for b in range(16):
            if ext[b] == 0 and sh[b] == 0:
                continue
            n = int(ext[b]) + int(sh[b])
            rows.append({"bounce": b, "extension_rays": int(ext[b]), "shadow_rays": int(sh[b]),
                         "ms": round(float(ms[b]), 3), "mrays_s": round(n / (ms[b] / 1000) / 1e6, 1) if ms[b] else None})
        later = [x for x in rows if x["bounce"] > 0]
        n_later = sum(x["extension_rays"] + x["shadow_rays"] for x in later)
        ms_later = sum(x["ms"] for x in later)
        out["config"]["per_bounce"] = rows
        out["config"]["bounce0_mrays_s"] = rows[0]["mrays_s"] if rows else None
        out["config"]["secondary_mrays_s"] = round(n_later / (ms_later / 1000) / 1e6, 1) if ms_later else None
    if not args.no_count:
        # The same launch as timed step 0 (same seeds) with work counters on.
        # Counting runs the reference-order traversal, so the node and
        # triangle counts are the reference's work for these rays; the ray
        # counts agree with the timed launch up to paths that a within-
        # rounding tie sent elsewhere.
        r.set_option(A.PT_OPT_COUNT_STATS, 1)
        counted = step(0)
        r.set_option(A.PT_OPT_COUNT_STATS, 0)
        for k in ("extension_rays", "shadow_rays"):
            a, b = getattr(counted, k), getattr(stats[0], k)
            assert abs(a - b) <= 1e-4 * b, (k, a, b)
        avg_bytes = algorithmic_bytes(counted, d.sky_width, d.sky_height)
        avg_ms = kern_ms / len(stats)
        achieved = avg_bytes / (avg_ms / 1000) / 1e9
        traffic = None
        tf = os.path.join(ROOT, "profiles", "%s_c2_traffic.json" % args.backend)
        if os.path.exists(tf):
            with open(tf) as f:
                traffic = json.load(f).get("hbm_bytes_per_launch")
        # achieved counts every node, triangle and shading fetch as if it came
        # from HBM (cache-oblivious), so frac can pass 1 once the caches serve
        # most of them; hbm_gbs is the measured traffic over the same time
        out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                           "algorithmic_bytes_per_launch": int(avg_bytes), "kernel_ms_per_launch": round(avg_ms, 3)}
        if traffic:
            hbm = traffic / (avg_ms / 1000) / 1e9
            out["roofline"]["hbm_gbs"] = round(hbm, 1)
            out["roofline"]["hbm_frac"] = round(hbm / HBM_PEAK_GBS, 4)
    if args.compare:
        out["alt"] = []
        for other in variants:
            if other == chosen:
                continue
            st2, wall2 = timed(other)
            rays2 = sum(x.extension_rays + x.shadow_rays for x in st2)
            ms2 = sum(x.kernel_ms for x in st2)
            out["alt"].append({"variant": other, "value": round(rays2 / wall2 / 1e6, 2), "unit": "Mrays/s",
                               "ms_per_step": round(wall2 * 1000 / args.steps, 3),
                               "kernel_ms_per_step": round(ms2 / args.steps, 3)})
    if not args.no_cpu:
        p = scene.params(W, H, 0, 1)
        out["cpu_baseline"] = cpu_baseline(scene, p, args.cpu_seconds)
    print(json.dumps(out), flush=True)
    r.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
