"""Many jobs per graph launch (pt_batch.h): every job's frame is bit-identical
to pt_trace on the same context state, frame after frame, with eager renders
and synchronisations between the graph launches."""
import numpy as np
import pytest

from helpers import small_mesh_scene
from webgpuraytracer_amd import Batch, Renderer, SceneBuilder
from webgpuraytracer_amd import _abi as A

pytestmark = pytest.mark.gpu

# (scene factory, width, height, options)
JOBS = [
    (lambda: SceneBuilder.config(1, seed=3), 64, 48, {}),
    (lambda: small_mesh_scene(seed=5), 40, 32, {}),
    (lambda: SceneBuilder.config(1, seed=9, bounces=6), 33, 17, {}),  # ragged tiles
    (lambda: small_mesh_scene(seed=7, n_instances=5), 24, 40, {A.PT_OPT_BACKEND: A.PT_BACKEND_MEGAKERNEL}),
    (lambda: SceneBuilder.config(1, seed=4), 48, 48, {A.PT_OPT_WAVEFRONT_POOL: 4096}),  # several batches per frame
]


def _make(job):
    make, w, h, opts = job
    r = Renderer(0)
    for k, v in opts.items():
        r.set_option(k, v)
    r.set_scene(make())
    r.resize(w, h)
    return r


@pytest.fixture(scope="module")
def pairs():
    batched = [_make(j) for j in JOBS]
    eager = [_make(j) for j in JOBS]
    yield batched, eager
    for r in batched + eager:
        r.close()


def test_batch_frames_match_pt_trace_bitwise(pairs):
    batched, eager = pairs
    b = Batch(0)
    for r in batched:
        b.add(r)
    frames = [[2, 1, 3, 1, 4], [2, 2, 2, 2, 2], [2, 2, 2, 2, 2], [1, 0, 2, 3, 1], [1, 0, 2, 3, 1]]
    for f, spp in enumerate(frames):
        st = b.frame(spp, base_seeds=list(range(7, 7 + len(JOBS))))
        for i, (rb, re) in enumerate(zip(batched, eager)):
            if spp[i]:
                se = re.render(spp=spp[i], base_seed=7 + i)  # eager work between the graph launches
                assert st[i].extension_rays == se.extension_rays and st[i].shadow_rays == se.shadow_rays, (f, i)
                assert st[i].paths == se.paths == spp[i] * rb.width * rb.height
            else:
                assert st[i].extension_rays == 0
            assert np.array_equal(rb.read_accum(), re.read_accum()), (f, i)
            assert np.array_equal(rb.read_output(), re.read_output()), (f, i)
    n_frames, inst, upd = b.graph_info()
    assert n_frames == len(frames)
    assert upd >= 2 and inst + upd == len(frames)  # unchanged topologies update the graph in place
    b.close()


def test_batch_argument_errors(pairs):
    batched, _ = pairs
    b = Batch(0)
    b.add(batched[0])
    with pytest.raises(A.PTError):
        b.add(batched[0])  # one context cannot be two jobs
    r = Renderer(0)  # no scene, no frame
    b.add(r)
    import ctypes as C
    params = (A.SceneParams * 2)(batched[0].params(), batched[0].params())
    passes = (A.u32 * 2)(1, 1)
    rc = b.lib.pt_batch_frame(b.h, params, passes, None, None)
    assert rc == A.PT_ERR_STATE and b"job 1" in b.lib.pt_batch_last_error(b.h)
    passes = (A.u32 * 2)(1, 0)  # the job without a scene skipped: fine
    assert b.lib.pt_batch_frame(b.h, params, passes, None, None) == A.PT_OK
    b.close()
    r.close()
    del C


def test_c3_workload_one_batch_frame():
    """The batched-jobs workload itself: 32 seeded jobs (resolutions
    log-uniform in 128^2..1024^2, 1-64 spp, 2-12 bounces, 1k-1M triangles,
    1-128 instances) rendered as one graph launch.  Every job's frame equals
    pt_trace on the same context with the same seeds, bit for bit; the
    smallest job also agrees with the oracle."""
    from oracle.refcpu import RefScene
    from webgpuraytracer_amd.batch_bench import build, job_specs
    specs = job_specs(32, 1)
    jobs = [build(s) for s in specs]
    rs = [r for r, _ in jobs]
    spp = [s["spp"] for s in specs]
    b = Batch(0)
    try:
        for r in rs:
            b.add(r)
        st = b.frame(spp, base_seeds=list(range(31, 31 + len(rs))))
        batched = [r.read_accum() for r in rs]
        total = 0
        for i, (r, s) in enumerate(zip(rs, specs)):
            r.resize(s["width"], s["height"])  # cleared frame, sample count 0
            se = r.render(spp=spp[i], base_seed=31 + i)
            assert st[i].paths == se.paths == spp[i] * s["width"] * s["height"], i
            assert st[i].extension_rays == se.extension_rays and st[i].shadow_rays == se.shadow_rays, i
            assert np.array_equal(batched[i].view(np.uint32), r.read_accum().view(np.uint32)), i
            total += st[i].paths
        print("c3: 32 jobs, %d paths in one batch frame" % total)
        # the oracle on the job with the least work
        k = min(range(len(specs)), key=lambda i: spp[i] * specs[i]["width"] * specs[i]["height"])
        r, scene = jobs[k]
        w, h = specs[k]["width"], specs[k]["height"]
        seeds = [((31 + k) * 0x9E3779B1 + j * 0x85EBCA77) & 0xFFFFFFFF for j in range(spp[k])]
        p = scene.params(w, h, sample_number=0, seed=seeds[0])
        ra, _, _ = RefScene(scene.desc()).render(p, passes=spp[k], seeds=seeds, with_output=False)
        rel = np.abs(batched[k][..., :3] - ra[..., :3]) / np.maximum(np.abs(ra[..., :3]), 1e-3)
        frac = float(np.mean(np.all(rel <= 1e-4, axis=-1)))
        print("c3 job %d (%dx%d, %d spp, %d triangles) against the oracle: %.5f" % (
            k, w, h, spp[k], specs[k]["triangles"], frac))
        assert frac >= 0.999
    finally:
        b.close()
        for r in rs:
            r.close()


def test_autotune_keeps_bits_and_caches(monkeypatch):
    """The tuner only moves options that never change results: a frame with
    the tuned options equals the frame with the defaults bit for bit; a second
    job of the same shape reuses the cached options without timing trials."""
    from webgpuraytracer_amd import autotune
    scene = small_mesh_scene(seed=8, n_instances=16)
    r = Renderer(0)
    try:
        r.set_scene(scene)
        r.resize(96, 64)
        opts, t = autotune.tune(r, 4, reps=1)
        assert t > 0 and set(opts) == set(autotune.DEFAULTS)
        r.reset()
        r.render(spp=4, base_seed=5)
        tuned = r.read_accum()
        for k, v in autotune.DEFAULTS.items():
            r.set_option(k, v)
        r.resize(96, 64)
        r.render(spp=4, base_seed=5)
        assert np.array_equal(tuned.view(np.uint32), r.read_accum().view(np.uint32))
        cache = autotune.TuneCache()
        calls = []
        real = autotune.tune
        monkeypatch.setattr(autotune, "tune", lambda *a, **k: calls.append(1) or real(*a, **k))
        first = cache.get(r, scene.desc().flags, 20000, 4)
        second = cache.get(r, scene.desc().flags, 30000, 4)  # same triangle decade, pixels, spp
        assert len(calls) == 1 and second == first
        for k, v in first.items():
            assert r.get_option(k) == v  # the cached options are applied
        cache.get(r, scene.desc().flags, 200000, 4)  # another decade: tuned anew
        assert len(calls) == 2
    finally:
        r.close()
