"""HIP path vs the oracle, through the C ABI (libpt.so), on seeded inputs.

Tolerances (written here, per the measurement plan):
  * hit identities (object, material, triangle, plane) exact; t within 1e-5
    relative; a disagreement passes only when the FP64 brute-force arbiter
    shows the two candidates tied;
  * 1-spp images: per-pixel agreement within 1e-4 relative on >= 99.9% of
    pixels (>= 99.5% within 1e-3 and >= 98% within 1e-4 with material
    textures); the rest are paths that a last-ulp difference in
    sin/cos/pow/exp/log between the device and host math libraries sent
    elsewhere, and every ray of those paths is checked hit for hit;
  * converged images: relMSE <= 1e-3 per channel;
  * env CDF built on the device: within 1 f32 ulp of the oracle's;
  * G-buffer: exact identities, values within 1e-5; denoiser: 1e-5 relative.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import compare_hits, hits_to_np, random_rays, rel_mse, small_mesh_scene
from oracle.refcpu import RefScene, build_sky_cdf as ref_cdf, denoise as ref_denoise
from webgpuraytracer_amd import Denoiser, SceneBuilder
from webgpuraytracer_amd import _abi as A

pytestmark = pytest.mark.gpu


def _setup(renderer, scene, w, h, gbuf=False):
    renderer.set_scene(scene, isgbuf=gbuf)
    renderer.resize(w, h)
    return RefScene(scene.desc())


def _agreement(ga, ra, tol):
    """Fraction of pixels whose three channels agree within tol relative."""
    rel = np.abs(ga[..., :3] - ra[..., :3]) / np.maximum(np.abs(ra[..., :3]), 1e-3)
    return float(np.mean(np.all(rel <= tol, axis=-1)))


def _assert_1spp(ga, ra, label, textured=False):
    """1-spp images.  Untextured: >= 99.9% of pixels within 1e-4.  With
    material textures a last-ulp difference in the hit's texture coordinate
    (atan2/acos/pow between the device and host math libraries) is amplified
    by bilinear filtering of 512x512 texels and the 2.2 power: >= 99.5% of
    pixels within 1e-3 and >= 98% within 1e-4, while every ray of every path
    is checked hit for hit by _assert_device_paths."""
    f4, f3 = _agreement(ga, ra, 1e-4), _agreement(ga, ra, 1e-3)
    print(label, "1spp pixel agreement %.5f within 1e-4, %.5f within 1e-3" % (f4, f3))
    if textured:
        assert f3 >= 0.995 and f4 >= 0.98
    else:
        assert f4 >= 0.999


def _assert_device_paths(renderer, ref, scene, p, n_bounces=6):
    """Every ray the device traced along each 1-spp path (the first
    n_bounces trace_scene calls), traced again by the oracle: the same hit,
    each mismatch adjudicated by the FP64 arbiter."""
    g, grays = renderer.trace_hits(p, n_bounces, rays=True)
    traced = np.any(grays[:, 3:] != 0.0, axis=1)
    rr, _ = ref.trace_rays(grays[traced])
    s = compare_hits(np.ctypeslib.as_array(g)[traced], rr, grays[traced], ref, torus=_torus_ids(scene))
    print("device path rays", s)
    assert s["within_t"] + s["adjudicated"] == s["n"] and s["n"] > 1000


@pytest.fixture
def traversal(renderer, request):
    """Runs a test under one BVH traversal and restores the default."""
    renderer.set_option(A.PT_OPT_TRAVERSAL, request.param)
    yield request.param
    renderer.set_option(A.PT_OPT_TRAVERSAL, A.PT_TRAVERSAL_WIDE)


TRAVERSALS = [A.PT_TRAVERSAL_REFERENCE, A.PT_TRAVERSAL_WIDE, A.PT_TRAVERSAL_WIDE_QUANTIZED]
TRAV_IDS = ["reference", "wide", "quant"]


@pytest.mark.parametrize("traversal", TRAVERSALS, indirect=True, ids=TRAV_IDS)
@pytest.mark.parametrize("which", ["c1", "mesh"])
def test_trace_rays_match_oracle(renderer, which, traversal):
    scene = SceneBuilder.config(1, seed=11) if which == "c1" else small_mesh_scene(seed=5)
    ref = _setup(renderer, scene, 16, 16)
    rays = random_rays(20000, seed=42, box=4.0 if which == "c1" else 9.0, spread=3.0 if which == "c1" else 8.0)
    g = renderer.trace_rays(rays)
    r, _ = ref.trace_rays(rays)
    s = compare_hits(g, r, rays, ref)
    print(which, s)
    assert s["exact_t"] >= 0.999 * s["n"]
    if traversal == A.PT_TRAVERSAL_REFERENCE:  # the reference's visit order: its node counts too
        assert np.array_equal(hits_to_np(g)["count"], hits_to_np(r)["count"])


def _hit_scene(which):
    if which == "c1":
        return SceneBuilder.config(1, seed=21), (96, 64)
    if which == "mesh":
        return small_mesh_scene(seed=9, n_instances=40), (128, 96)
    return SceneBuilder.config(2, seed=1), (192, 108)  # the 808k-triangle scene itself


def _c2_scene():
    global _C2
    if "_C2" not in globals():
        _C2 = SceneBuilder.config(2, seed=1)
    return _C2


def _hit_fields(a):
    return {k: a[k][...] for k in ("t", "o", "m", "prim", "plane", "count")}


@pytest.mark.parametrize("which", ["c1", "mesh", "c2"])
def test_primary_secondary_hits_match_oracle(renderer, which):
    """The deterministic hit pass with the per-pixel PCG seeds: primary rays
    and hits bit-identical to the oracle's; every secondary ray the device
    traced, traced again by the oracle, finds the same primitive, triangle
    and plane with t within 1e-5 relative, each mismatch adjudicated by the
    FP64 arbiter (reference order and the wide traversal).  Where both sides
    produced the same secondary ray, the hits are bit-identical; where the
    rays differ, they differ by the last-ulp rounding of the BSDF sampler's
    sin/cos/sqrt/pow between the device and host math libraries."""
    scene, (w, h) = (_c2_scene(), (192, 108)) if which == "c2" else _hit_scene(which)
    ref = _setup(renderer, scene, w, h)
    p = scene.params(w, h, sample_number=3, seed=0xC0FFEE)
    g, grays = renderer.trace_hits(p, 2, rays=True)
    r, rrays = ref.trace_hits(p, 2, rays=True)
    ga, ra = hits_to_np(g), hits_to_np(r)
    prim, sec = slice(0, None, 2), slice(1, None, 2)
    assert np.array_equal(grays[prim], rrays[prim])  # camera rays, bit for bit
    for k in ("o", "m", "prim", "plane", "count"):
        assert np.array_equal(ga[k][prim], ra[k][prim]), k
    assert np.array_equal(ga["t"][prim], ra["t"][prim])
    gsec, rsec = grays[sec], rrays[sec]
    traced = np.any(gsec[:, 3:] != 0.0, axis=1)
    assert traced.sum() >= 300
    gs = np.ctypeslib.as_array(g)[sec][traced]
    # the device's own secondary rays on the oracle: reference order ...
    r_same, _ = ref.trace_rays(gsec[traced])
    s = compare_hits(gs, r_same, gsec[traced], ref)
    print(which, "secondary rays", s)
    assert s["within_t"] + s["adjudicated"] == s["n"]
    assert np.array_equal(hits_to_np(gs)["count"], hits_to_np(r_same)["count"])
    # ... and the wide traversal the renderer uses by default
    s_w = compare_hits(renderer.trace_rays(gsec[traced]), r_same, gsec[traced], ref)
    assert s_w["within_t"] + s_w["adjudicated"] == s_w["n"]
    # identical secondary rays give identical hits
    same_ray = traced & np.all(gsec == rsec, axis=1)
    frac_same = float(same_ray.sum()) / float(traced.sum())
    gsa, rsa = ga, ra
    for k in ("t", "o", "m", "prim", "plane", "count"):
        assert np.array_equal(gsa[k][sec][same_ray], rsa[k][sec][same_ray]), k
    # rays that differ start at the same point and point within rounding of each other
    both = traced & np.any(rsec[:, 3:] != 0.0, axis=1)
    diff = both & ~same_ray
    assert np.array_equal(traced, both)  # a path continues on both sides or on neither
    if diff.any():
        assert np.array_equal(gsec[diff, :3], rsec[diff, :3])
        ddir = np.abs(gsec[diff, 3:] - rsec[diff, 3:]).max()
    else:
        ddir = 0.0
    print(which, "identical secondary rays %.4f of %d, max direction difference elsewhere %.3g"
          % (frac_same, int(traced.sum()), ddir))
    assert frac_same >= 0.5  # the rest differ by the sampler's last-ulp rounding
    assert ddir <= 5e-6


def test_c2_scene_rays_match_oracle(renderer):
    """The configuration-2 scene (64 instances of 4 displaced geodesic
    spheres, 808,500 triangles): 20,000 rays against the oracle, both
    traversals, with the reference order's per-ray node counts."""
    scene = _c2_scene()
    ref = _setup(renderer, scene, 16, 16)
    rays = random_rays(20000, seed=46, box=9.0, spread=8.0)
    r, _ = ref.trace_rays(rays)
    assert (hits_to_np(r)["prim"] >= 0).mean() > 0.3  # triangles are hit
    for trav in TRAVERSALS:
        renderer.set_option(A.PT_OPT_TRAVERSAL, trav)
        try:
            g = renderer.trace_rays(rays)
        finally:
            renderer.set_option(A.PT_OPT_TRAVERSAL, A.PT_TRAVERSAL_WIDE)
        s = compare_hits(g, r, rays, ref)
        print("c2 traversal", trav, s)
        assert s["within_t"] + s["adjudicated"] == s["n"] and s["exact_t"] >= 0.999 * s["n"]
        if trav == A.PT_TRAVERSAL_REFERENCE:
            assert np.array_equal(hits_to_np(g)["count"], hits_to_np(r)["count"])


def test_c2_section_1spp_matches_oracle(renderer):
    """An 8-bounce 1-spp render of a 240x135 section of the 1920x1080
    configuration-2 frame against the oracle."""
    scene = _c2_scene()
    w, h = 1920, 1080
    ref = _setup(renderer, scene, w, h)
    p = scene.params(w, h, sample_number=0, seed=4321)
    x0, y0 = 690, 200  # a part of the frame the instances fill
    p.section[0], p.section[1] = x0, y0
    st = renderer.render_section(p, section_size=(240, 135), passes=1)
    ga = renderer.read_accum()[y0:y0 + 135, x0:x0 + 240]
    ra, _, rst = ref.render(p, passes=1, section=(240, 135), with_output=False)
    ra = ra[y0:y0 + 135, x0:x0 + 240]
    print("c2 section rays gpu", st.extension_rays, st.shadow_rays, "ref", rst.extension_rays, rst.shadow_rays)
    assert st.paths == 240 * 135 and rst.tri_tests > 0 and rst.shadow_rays > 0.5 * st.paths
    _assert_1spp(ga, ra, "c2 section")
    assert abs(int(st.extension_rays) - int(rst.extension_rays)) <= 0.005 * rst.extension_rays
    assert abs(int(st.shadow_rays) - int(rst.shadow_rays)) <= 0.005 * rst.shadow_rays


@pytest.mark.parametrize("traversal", TRAVERSALS, indirect=True, ids=TRAV_IDS)
def test_shadow_rays_match_oracle(renderer, traversal):
    scene = SceneBuilder.config(1, seed=13)
    ref = _setup(renderer, scene, 16, 16)
    rays = random_rays(20000, seed=7, box=4.0, spread=3.0)
    rng = np.random.default_rng(3)
    tgt = rng.integers(-1, 12, size=len(rays)).astype(np.int32)
    dist = rng.uniform(0.5, 12.0, size=len(rays)).astype(np.float32)
    g = renderer.trace_shadow_rays(rays, tgt, dist)
    r = ref.trace_shadow_rays(rays, tgt, dist)
    assert np.array_equal(g, r)
    assert 0.05 < g[:, 0].mean() < 0.95  # both outcomes occur


def test_render_1spp_matches_oracle(renderer):
    scene = SceneBuilder.config(1, seed=31)
    w, h = 128, 96
    ref = _setup(renderer, scene, w, h)
    assert renderer.get_option(A.PT_OPT_SPECIALISED) == 1
    assert renderer.get_option(A.PT_OPT_KERNEL_FLAGS) == scene.desc().flags
    p = scene.params(w, h, sample_number=0, seed=777)
    st = renderer.render_section(p, passes=1)
    ga = renderer.read_accum()
    ra, ro, rst = ref.render(p, passes=1)
    print("rays gpu", st.extension_rays, st.shadow_rays, "ref", rst.extension_rays, rst.shadow_rays)
    _assert_1spp(ga, ra, "c1")
    assert st.paths == w * h
    assert abs(int(st.extension_rays) - int(rst.extension_rays)) <= 0.002 * rst.extension_rays
    assert abs(int(st.shadow_rays) - int(rst.shadow_rays)) <= 0.002 * rst.shadow_rays
    out = renderer.read_output()
    assert np.mean(np.abs(out.astype(int) - ro.astype(int)) <= 1) >= 0.999


def test_render_1spp_mixed_scene_matches_oracle(renderer):
    """Analytic primitives and instanced meshes in one scene: the kernels'
    run-time-flags variant (the batched-jobs scenes)."""
    scene = SceneBuilder.config(3, seed=12, mesh_freq=[6, 7, 8, 9], n_instances=9, bounces=5)
    w, h = 96, 64
    ref = _setup(renderer, scene, w, h)
    p = scene.params(w, h, sample_number=0, seed=4242)
    st = renderer.render_section(p, passes=1)
    ga = renderer.read_accum()
    ra, _, rst = ref.render(p, passes=1)
    print("rays gpu", st.extension_rays, "ref", rst.extension_rays)
    _assert_1spp(ga, ra, "mixed scene")
    assert abs(int(st.extension_rays) - int(rst.extension_rays)) <= 0.002 * rst.extension_rays
    assert rst.tri_tests > 0 and st.shadow_rays > 0  # meshes are hit and lights are sampled


def _batch_means(render_batch, batches):
    return np.stack([render_batch(k)[..., :3].astype(np.float64) for k in range(batches)])


def test_render_converges_to_oracle_with_independent_seeds(renderer):
    """Converged images from independent seed sets: the device at 16,384 spp
    and the oracle at 4,096 spp, on disjoint per-pass seeds, agree at relMSE
    <= 1e-3 per channel, and pixel by pixel within their batch error bars
    (a z-test: a bias in any shading term shows many standard errors wide)."""
    scene = SceneBuilder.config(1, seed=41)
    w, h = 32, 24
    ref = _setup(renderer, scene, w, h)
    p = scene.params(w, h, sample_number=0, seed=1)

    def seeds(base, k, n):
        return ((np.arange(n, dtype=np.uint64) + k * n) * 2654435761 + base) % (1 << 32)

    def gpu_batch(k):
        renderer.resize(w, h)  # clears the accumulation buffer
        renderer.render_section(p, passes=1024, seeds=seeds(0x1234567, k, 1024).astype(np.uint32))
        return renderer.read_accum()

    def ref_batch(k):
        return ref.render(p, passes=256, seeds=seeds(0x89ABCDE, k, 256).astype(np.uint32), with_output=False)[0]

    g = _batch_means(gpu_batch, 16)
    r = _batch_means(ref_batch, 16)
    mg, mr = g.mean(0), r.mean(0)
    se = np.sqrt(g.var(0, ddof=1) / 16 + r.var(0, ddof=1) / 16)
    z = (mg - mr) / np.maximum(se, 1e-12)
    for c in range(3):
        e = rel_mse(mg[..., c], mr[..., c])
        print("channel", c, "relMSE", e)
        assert e <= 1e-3
    print("mean z^2 %.3f max |z| %.2f" % (np.mean(z * z), np.abs(z).max()))
    assert np.mean(z * z) < 1.5 and np.abs(z).max() < 6.5


def test_sky_cdf_built_on_device(renderer):
    scene = SceneBuilder.config(1, seed=51, sky=(300, 150))
    d = scene.desc()
    d.sky_cdf = None
    renderer.set_scene(d)
    n = 151 + 150 * 301
    g, tot = renderer.read_sky_cdf(n)
    arr = scene.arrays()
    r, rtot = ref_cdf(arr["sky"])
    ulp = np.abs(g.view(np.int32).astype(np.int64) - r.view(np.int32).astype(np.int64))
    print("max ulp", ulp.max(), "differing", int((ulp > 0).sum()))
    assert ulp.max() <= 1
    assert abs(tot - np.float32(rtot)) <= 1e-6 * rtot


def test_gbuffer_matches_oracle(renderer):
    scene = SceneBuilder.config(1, seed=61)
    w, h = 80, 60
    ref = _setup(renderer, scene, w, h, gbuf=True)
    p = scene.params(w, h, 0, 5)
    renderer.gbuffer(p)
    g = renderer.read_gbuffer()
    r, ro = ref.gbuffer(p)
    assert np.allclose(g, r, rtol=1e-5, atol=1e-6)
    assert np.mean(np.abs(renderer.read_output().astype(int) - ro.astype(int)) <= 1) >= 0.999


@pytest.fixture(params=[A.PT_DENOISE_MATH_EXACT, A.PT_DENOISE_MATH_FUSED], ids=["exact", "fused"])
def denoise_math(renderer, request):
    renderer.set_option(A.PT_OPT_DENOISE_MATH, request.param)
    yield request.param
    renderer.set_option(A.PT_OPT_DENOISE_MATH, A.PT_DENOISE_MATH_FUSED)


@pytest.mark.parametrize("passes", [4, 7])
@pytest.mark.parametrize("stride_mode", [0, 1])
def test_denoise_matches_oracle(renderer, stride_mode, passes, denoise_math):
    """Prepass, a-trous passes and composite against the oracle on a frame
    whose size is no multiple of the tiles; 7 passes reach strides of 64,
    wider than the frame's sub-lattice tiles.  EXACT evaluates each weight as
    the oracle does; FUSED is held to the same 1e-5 relative."""
    scene = SceneBuilder.config(1, seed=71)
    w, h = 72, 56
    _setup(renderer, scene, w, h, gbuf=True)
    p = scene.params(w, h, 0, 9)
    renderer.gbuffer(p)
    renderer.render_section(p, passes=4, seeds=[1, 2, 3, 4])
    acc = renderer.read_accum()
    acc[3, 5, 0] = np.nan  # sanitised by the prepass
    acc[7, 9, 1] = np.inf
    renderer.write_accum(acc)
    gb = renderer.read_gbuffer()
    dn = Denoiser(renderer)
    dn.passes = passes
    out8 = dn.execute(stride_mode)
    res = dn.read_result()
    u = dn.uniforms()
    rres, r8 = ref_denoise(u, acc, gb, stride_mode)
    err = np.abs(res - rres) / np.maximum(np.abs(rres), 1e-3)
    print("denoise max rel err", float(err.max()))
    assert np.all(np.isfinite(res))
    assert np.allclose(res, rres, rtol=1e-5, atol=1e-6)
    assert np.mean(np.abs(out8.astype(int) - r8.astype(int)) <= 1) >= 0.999


def test_denoise_1080p_frame_matches_oracle(renderer):
    """The configuration-5 frame: 1920x1080, a 4-spp render of a seeded
    small-analytic scene with its G-buffer, the reference's default sigmas
    and 5 a-trous passes (strides 1..16), against the oracle over the whole
    frame; the fused weights within 5e-5 of the exact ones at worst (their
    mean difference is ~1e-7)."""
    scene = SceneBuilder.config(1, seed=5)
    w, h = 1920, 1080
    _setup(renderer, scene, w, h, gbuf=True)
    p = scene.params(w, h, 0, 1)
    renderer.gbuffer(p)
    renderer.render_section(p, passes=4, seeds=[11, 12, 13, 14])
    acc, gb = renderer.read_accum(), renderer.read_gbuffer()
    dn = Denoiser(renderer)
    dn.passes = 5
    out = {}
    try:
        for m in (A.PT_DENOISE_MATH_EXACT, A.PT_DENOISE_MATH_FUSED):
            renderer.set_option(A.PT_OPT_DENOISE_MATH, m)
            out8 = dn.execute(0)
            out[m] = (dn.read_result(), out8)
    finally:
        renderer.set_option(A.PT_OPT_DENOISE_MATH, A.PT_DENOISE_MATH_FUSED)
    rres, r8 = ref_denoise(dn.uniforms(), acc, gb, 0)
    for m, (res, out8) in out.items():
        err = np.abs(res - rres) / np.maximum(np.abs(rres), 1e-3)
        print("1080p denoise math", m, "max rel err", float(err.max()), "mean", float(err.mean()))
        assert np.allclose(res, rres, rtol=2e-5, atol=2e-6)
        assert np.mean(np.abs(out8.astype(int) - r8.astype(int)) <= 1) >= 0.9999
    d = np.abs(out[0][0] - out[1][0]) / np.maximum(np.abs(out[1][0]), 1e-3)
    print("fused vs exact max rel", float(d.max()), "mean", float(d.mean()))
    assert d.max() <= 5e-5 and d.mean() <= 1e-6


@pytest.mark.parametrize("name", ["c1_64x48", "mesh_48x32"])
def test_gpu_matches_golden_fixture(renderer, name):
    """GPU primary hits and 1-spp image against the committed oracle fixture."""
    import os
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.join(here, "golden"))
    from make_golden import CASES, scene_for
    g = np.load(os.path.join(here, "golden", name + ".npz"))
    config, seed, w, h, sn, sseed = CASES[name]
    s = scene_for(config, seed)
    renderer.set_scene(s)
    renderer.resize(w, h)
    p = s.params(w, h, sn, sseed)
    hits = hits_to_np(renderer.trace_hits(p, 2))
    prim = slice(0, None, 2)
    for k in ("o", "m", "prim", "plane"):
        assert np.array_equal(hits[k][prim].astype(np.int32), g[k][prim]), k
    assert np.array_equal(hits["t"][prim], g["t"][prim])
    renderer.render_section(p, passes=1)
    acc = renderer.read_accum()
    _assert_1spp(acc, g["accum"], "golden " + name)


def _restore_defaults(renderer):
    renderer.set_option(A.PT_OPT_BACKEND, A.PT_BACKEND_WAVEFRONT)
    renderer.set_option(A.PT_OPT_SCHEDULE, A.PT_SCHEDULE_REGEN)
    renderer.set_option(A.PT_OPT_WAVEFRONT_POOL, 1 << 27)
    renderer.set_option(A.PT_OPT_WAVEFRONT_REFILL, 56)


@pytest.mark.parametrize("section", [None, (16, 8, 40, 24), (5, 3, 37, 21)], ids=["frame", "section", "ragged"])
@pytest.mark.parametrize("which", ["c1", "mesh"])
def test_backends_and_schedules_bitwise(renderer, which, section):
    """Both megakernel pass schedules and the wavefront backend (several
    batch sizes and refill thresholds) run the same shading code with the same
    per-pixel RNG order and add NEE contributions in the same order:
    identical bits and identical work counts, for a whole frame and for a
    section of it."""
    scene = SceneBuilder.config(1, seed=81) if which == "c1" else small_mesh_scene(seed=13)
    w, h = 80, 48
    renderer.set_scene(scene)
    seeds = [11, 22, 33, 44, 55]
    p = scene.params(w, h, sample_number=0, seed=seeds[0])
    size = (0, 0)
    if section:
        p.section[0], p.section[1] = section[0], section[1]
        size = section[2:]
    npix = (size[0] * size[1]) if section else w * h
    mk = (A.PT_OPT_BACKEND, A.PT_BACKEND_MEGAKERNEL)
    wf = (A.PT_OPT_BACKEND, A.PT_BACKEND_WAVEFRONT)
    variants = {"regen": [mk, (A.PT_OPT_SCHEDULE, A.PT_SCHEDULE_REGEN)],
                "sync": [mk, (A.PT_OPT_SCHEDULE, A.PT_SCHEDULE_SYNC)],
                # 2 passes per batch: 3 batches
                "wavefront_batches": [wf, (A.PT_OPT_WAVEFRONT_POOL, 2 * npix)],
                "wavefront_refill1": [wf, (A.PT_OPT_WAVEFRONT_POOL, 1 << 27), (A.PT_OPT_WAVEFRONT_REFILL, 1)],
                "wavefront": [wf, (A.PT_OPT_WAVEFRONT_REFILL, 56)]}
    runs = {}
    try:
        for name, opts in variants.items():
            for k, v in opts:
                renderer.set_option(k, v)
            renderer.resize(w, h)
            st = renderer.render_section(p, section_size=size, passes=5, seeds=seeds)
            runs[name] = (st, renderer.read_accum(), renderer.read_output())
    finally:
        _restore_defaults(renderer)
    st_m, acc_m, out_m = runs["regen"]
    if section:  # pixels outside the section stay untouched
        x0, y0, sw, sh = section
        mask = np.ones((h, w), bool)
        mask[y0:y0 + sh, x0:x0 + sw] = False
        assert np.all(acc_m[mask] == 0.0)
        assert np.any(acc_m[~mask] != 0.0)
    for name in variants:
        st_o, acc_o, out_o = runs[name]
        assert np.array_equal(acc_m.view(np.uint32), acc_o.view(np.uint32)), name
        assert np.array_equal(out_m, out_o), name
        for k in ("extension_rays", "shadow_rays", "paths", "node_pops", "inner_pops", "tri_tests", "shade_events"):
            assert getattr(st_m, k) == getattr(st_o, k), (name, k)


def test_bounce_stats_and_zero_bounces(renderer):
    """Per-bounce ray counts add up to the call's totals (camera rays at
    bounce 0, fewer rays at every later bounce), bounce times are reported
    only with PT_OPT_BOUNCE_TIMING and do not change the image; a scene with
    a bounce limit of 0 adds zero radiance on both backends, also after a
    render that left other values in the wavefront's path slots."""
    scene = SceneBuilder.config(1, seed=81)
    w, h = 64, 48
    renderer.set_scene(scene)
    renderer.resize(w, h)
    p = scene.params(w, h, sample_number=0, seed=5)
    seeds = [5, 6, 7]
    try:
        renderer.set_option(A.PT_OPT_BACKEND, A.PT_BACKEND_WAVEFRONT)
        st = renderer.render_section(p, passes=3, seeds=seeds)
        plain = renderer.read_accum()
        ext, sh, ms = renderer.bounce_stats(16)
        assert not ms.any()
        renderer.resize(w, h)
        renderer.set_option(A.PT_OPT_BOUNCE_TIMING, 1)
        st2 = renderer.render_section(p, passes=3, seeds=seeds)
        ext2, sh2, ms2 = renderer.bounce_stats(16)
        assert np.array_equal(renderer.read_accum().view(np.uint32), plain.view(np.uint32))
    finally:
        renderer.set_option(A.PT_OPT_BOUNCE_TIMING, 0)
    bl = scene.desc().bounce_limit
    assert 0 < bl <= 16
    for e, s, t in ((ext, sh, st), (ext2, sh2, st2)):
        assert int(e.sum()) == t.extension_rays and int(s.sum()) == t.shadow_rays
        assert e[0] == 3 * w * h
        assert np.all(np.diff(e[:bl].astype(np.int64)) <= 0) and not e[bl:].any() and not s[bl:].any()
    assert np.array_equal(ext, ext2) and np.array_equal(sh, sh2)
    assert np.all(ms2[:bl] > 0) and not ms2[bl:].any()
    # bounce limit 0: no bounce runs, every path adds zero
    d = scene.desc()
    d.bounce_limit = 0
    try:
        for backend in (A.PT_BACKEND_WAVEFRONT, A.PT_BACKEND_MEGAKERNEL):
            renderer.set_option(A.PT_OPT_BACKEND, backend)
            renderer.set_scene(scene)
            renderer.resize(w, h)
            renderer.render_section(p, passes=3, seeds=seeds)  # leaves radiance in the path slots
            renderer.set_scene(d)
            renderer.write_accum(np.full((h, w, 4), 2.0, np.float32))
            st0 = renderer.render_section(p, passes=3, seeds=seeds)
            acc = renderer.read_accum()
            assert st0.extension_rays == 0 and st0.shadow_rays == 0
            # accum lerps toward zero: 2 * (1 - 1/1) after pass 0, then stays 0
            assert not acc[..., :3].any(), backend
    finally:
        _restore_defaults(renderer)
        renderer.set_scene(scene)


@pytest.mark.parametrize("which", ["c1", "mesh"])
def test_wide_traversal_renders_like_reference_order(renderer, which):
    """The wide traversals (exact and quantised boxes) pick a different
    closest hit only on ties within rounding; on these scenes a 5-pass render
    is pixel-identical to the reference-order traversal's on at least 99.9%
    of pixels."""
    scene = SceneBuilder.config(1, seed=82) if which == "c1" else small_mesh_scene(seed=14)
    w, h = 96, 64
    renderer.set_scene(scene)
    seeds = [5, 6, 7, 8, 9]
    p = scene.params(w, h, sample_number=0, seed=seeds[0])
    acc = {}
    try:
        for t in TRAVERSALS:
            renderer.set_option(A.PT_OPT_TRAVERSAL, t)
            renderer.resize(w, h)
            renderer.render_section(p, passes=5, seeds=seeds)
            acc[t] = renderer.read_accum()
    finally:
        renderer.set_option(A.PT_OPT_TRAVERSAL, A.PT_TRAVERSAL_WIDE)
    for t in TRAVERSALS[1:]:
        same = np.all(acc[0].view(np.uint32) == acc[t].view(np.uint32), axis=-1)
        print(which, "traversal", t, "identical pixels", same.mean())
        assert same.mean() >= 0.999


@pytest.mark.parametrize("traversal", TRAVERSALS, indirect=True, ids=TRAV_IDS)
def test_stack_spill_bitwise(renderer, traversal):
    """Traversal-stack entries beyond the LDS share spill to device memory:
    a 2-entry LDS share (most pushes spill) renders the same bits as a
    stack held entirely in LDS."""
    scene = small_mesh_scene(seed=15)
    w, h = 64, 48
    renderer.set_scene(scene)
    seeds = [3, 4, 5]
    p = scene.params(w, h, sample_number=0, seed=seeds[0])
    acc = []
    try:
        for entries in (0, 2):
            renderer.set_option(A.PT_OPT_LDS_STACK, entries)
            renderer.resize(w, h)
            renderer.render_section(p, passes=3, seeds=seeds)
            acc.append(renderer.read_accum())
    finally:
        renderer.set_option(A.PT_OPT_LDS_STACK, 16)
    assert np.array_equal(acc[0].view(np.uint32), acc[1].view(np.uint32))


@pytest.fixture
def device_blas(renderer):
    """Bottom-level trees rebuilt on the device at the next upload."""
    renderer.set_option(A.PT_OPT_BLAS_BUILD, A.PT_BLAS_BUILD_DEVICE)
    yield
    renderer.set_option(A.PT_OPT_BLAS_BUILD, A.PT_BLAS_BUILD_REFERENCE)


@pytest.mark.parametrize("which", ["mesh", "mixed"])
def test_device_built_blas_finds_the_same_hits(renderer, device_blas, which):
    """A linear BVH built on the device from the triangles: the same closest
    hits (primitive ids of the uploaded order, distances) as the oracle on the
    uploaded BVH; the reference-order traversal still reads the uploaded BVH."""
    if which == "mesh":
        scene = small_mesh_scene(seed=5, freq=(9, 12, 15, 20), n_instances=20)
        box, spread = 9.0, 8.0
    else:
        scene = SceneBuilder.config(3, seed=21, mesh_freq=[10, 12, 14, 16], n_instances=12, bounces=4)
        box, spread = 4.0, 3.0
    ref = _setup(renderer, scene, 16, 16)
    rays = random_rays(20000, seed=43, box=box, spread=spread)
    r, _ = ref.trace_rays(rays)
    g = renderer.trace_rays(rays)
    s = compare_hits(g, r, rays, ref)
    print(which, "device BLAS", s)
    assert s["exact_t"] >= 0.999 * s["n"]
    assert hits_to_np(r)["prim"].max() > 0  # triangles are hit
    renderer.set_option(A.PT_OPT_TRAVERSAL, A.PT_TRAVERSAL_REFERENCE)
    try:
        g2 = renderer.trace_rays(rays)
    finally:
        renderer.set_option(A.PT_OPT_TRAVERSAL, A.PT_TRAVERSAL_WIDE)
    assert np.array_equal(hits_to_np(g2)["count"], hits_to_np(r)["count"])


def test_device_built_blas_renders_like_uploaded_blas(renderer):
    scene = small_mesh_scene(seed=14, freq=(9, 12, 15, 20), n_instances=20)
    w, h = 96, 64
    seeds = [5, 6, 7, 8, 9]
    p = scene.params(w, h, sample_number=0, seed=seeds[0])
    acc = {}
    try:
        for build in (A.PT_BLAS_BUILD_REFERENCE, A.PT_BLAS_BUILD_DEVICE):
            renderer.set_option(A.PT_OPT_BLAS_BUILD, build)
            renderer.set_scene(scene)
            renderer.resize(w, h)
            renderer.render_section(p, passes=len(seeds), seeds=seeds)
            acc[build] = renderer.read_accum()
    finally:
        renderer.set_option(A.PT_OPT_BLAS_BUILD, A.PT_BLAS_BUILD_REFERENCE)
    same = np.all(acc[0] == acc[1], axis=-1).mean()
    print("device vs uploaded BLAS: identical pixels", same)
    assert same >= 0.999


# ---- top-level tree built on the device (PT_TLAS_BUILD_DEVICE) ----

def _uv_sphere(nu=12, nv=8):
    """A small closed triangle mesh: positions, normals, indices."""
    th = np.linspace(0.0, np.pi, nv + 1)
    ph = np.linspace(0.0, 2.0 * np.pi, nu + 1)
    p = np.array([[np.sin(a) * np.cos(b), np.cos(a), np.sin(a) * np.sin(b)] for a in th for b in ph], np.float32)
    idx = []
    for i in range(nv):
        for j in range(nu):
            a, b = i * (nu + 1) + j, i * (nu + 1) + j + 1
            c, d = a + nu + 1, b + nu + 1
            idx += [(a, c, b), (b, c, d)]
    return p, p.copy(), np.array(idx, np.uint32)


def _moving_scene(t):
    """The same objects (every primitive type and mesh instances, all with
    rotations) at time t: each moves along its own seeded velocity."""
    rng = np.random.default_rng(77)
    b = SceneBuilder()
    mats = [b.material(color=tuple(rng.uniform(0.2, 0.9, 3)), roughness=float(rng.uniform(0.1, 0.9)))
            for _ in range(4)]
    mesh = b.model(*_uv_sphere())
    b.plane(mats[0], (0.0, 1.0, 0.0), -2.0)

    def pose():
        pos = rng.uniform(-3.0, 3.0, 3) + t * rng.uniform(-1.0, 1.0, 3)
        q = rng.normal(size=4)
        return tuple(float(x) for x in pos), tuple(float(x) for x in q / np.linalg.norm(q))  # unit quaternion

    for k in range(24):
        pos, _ = pose()
        b.sphere(mats[k % 4], pos, float(rng.uniform(0.1, 0.5)))
    for k in range(12):
        pos, rot = pose()
        b.cube(mats[k % 4], pos, tuple(float(x) for x in rng.uniform(0.05, 0.4, 3)), rotation=rot)
    for k in range(8):
        pos, rot = pose()
        b.cylinder(mats[k % 4], pos, float(rng.uniform(0.1, 0.4)), float(rng.uniform(0.2, 0.8)), rotation=rot,
                   top_ratio=float(rng.uniform(0.3, 1.5)))
    for k in range(8):
        pos, rot = pose()
        b.torus(mats[k % 4], pos, float(rng.uniform(0.3, 0.6)), float(rng.uniform(0.05, 0.2)), rotation=rot)
    for k in range(20):
        pos, rot = pose()
        b.instance(mesh, mats[k % 4], pos, rotation=rot, scale=tuple(float(x) for x in rng.uniform(0.2, 0.6, 3)))
    b.sky()
    b.camera()
    b.bounces(4)
    return b.finalize()


def _without_tlas(scene):
    d = scene.desc()
    d.tlas.count = 0  # the device builds the TLAS; the array is not read
    return d


@pytest.fixture
def device_tlas(renderer):
    renderer.set_option(A.PT_OPT_TLAS_BUILD, A.PT_TLAS_BUILD_DEVICE)
    yield
    renderer.set_option(A.PT_OPT_TLAS_BUILD, A.PT_TLAS_BUILD_REFERENCE)


def _gpu_hits(renderer, scene, build, rays):
    """Hits of both traversals with the TLAS built one way (the device build
    gets the descriptor with its TLAS emptied)."""
    renderer.set_option(A.PT_OPT_TLAS_BUILD, build)
    renderer.set_scene(_without_tlas(scene) if build == A.PT_TLAS_BUILD_DEVICE else scene)
    renderer.resize(16, 16)
    out = {}
    for trav in TRAVERSALS:
        renderer.set_option(A.PT_OPT_TRAVERSAL, trav)
        try:
            out[trav] = renderer.trace_rays(rays)
        finally:
            renderer.set_option(A.PT_OPT_TRAVERSAL, A.PT_TRAVERSAL_WIDE)
    return out


def _check_against_oracle(hits, ref, rays, label, torus=None):
    """The oracle's closest hits: identities equal and t within 1e-5 (nearly
    all bit for bit).  Rays whose closest hit is a torus (object ids in
    `torus`) are held to equal identities and t within 1e-3 only: the torus
    quartic (resolvent cubic, square and cube roots) turns last-bit
    differences between the device's and the host's built-ins into t
    differences, 7e-7 relative on most such rays and near sqrt(eps) at a
    grazing hit (a double root)."""
    r, _ = ref.trace_rays(rays)
    R = hits_to_np(r)
    assert (R["o"] >= 0).mean() > 0.2  # objects are hit
    tor = np.isin(R["o"], sorted(torus)) & (R["plane"] < 0) if torus else np.zeros(len(rays), bool)
    keep = np.nonzero(~tor)[0]
    ra = np.ctypeslib.as_array(r)
    for trav, g in hits.items():
        ga = np.ctypeslib.as_array(g)
        s = compare_hits(ga[keep], ra[keep], rays[keep], ref)
        print(label, "traversal", trav, s, "torus rays", int(tor.sum()))
        assert s["exact_t"] >= 0.995 * s["n"]
        G = hits_to_np(g)
        for k in ("o", "m", "prim", "plane"):
            assert np.array_equal(G[k][tor], R[k][tor]), k
        assert np.all(np.abs(G["t"][tor] - R["t"][tor]) <= 1e-3 * np.abs(R["t"][tor]))


def _torus_ids(scene):
    return set(np.nonzero(scene.arrays()["objects"]["object_type"] == A.PT_OBJ_TORUS)[0].tolist())


def _same_bits(a, b):
    A_, B_ = hits_to_np(a), hits_to_np(b)
    return all(np.array_equal(A_[k], B_[k]) for k in ("t", "o", "m", "prim", "plane"))


@pytest.mark.parametrize("which", ["c1", "mixed", "moving"])
def test_device_built_tlas_finds_the_same_hits(renderer, which):
    """Both forms of the top-level tree built on the device from the objects
    (the descriptor's TLAS left empty): the oracle's closest hits, and the
    very same hits (every field, bit for bit) as the uploaded TLAS gives on
    the device, through the wide and the child-pair traversal."""
    if which == "c1":
        scene, box, spread = SceneBuilder.config(1, seed=11), 4.0, 3.0
    elif which == "mixed":
        scene = SceneBuilder.config(3, seed=21, mesh_freq=[6, 8, 10, 12], n_instances=40, bounces=4)
        box, spread = 4.0, 3.0
    else:
        scene, box, spread = _moving_scene(0.0), 5.0, 4.0
    rays = random_rays(20000, seed=44, box=box, spread=spread)
    try:
        up = _gpu_hits(renderer, scene, A.PT_TLAS_BUILD_REFERENCE, rays)
        dev = _gpu_hits(renderer, scene, A.PT_TLAS_BUILD_DEVICE, rays)
    finally:
        renderer.set_option(A.PT_OPT_TLAS_BUILD, A.PT_TLAS_BUILD_REFERENCE)
    _check_against_oracle(dev, RefScene(scene.desc()), rays, which, _torus_ids(scene))
    for trav in TRAVERSALS:
        assert _same_bits(dev[trav], up[trav]), trav


def test_update_objects_rebuilds_the_device_tlas(renderer, device_tlas):
    """pt_update_objects with moved objects and no TLAS: the device rebuilds
    the tree, and the hits equal the oracle's on the moved scene (traversing
    the TLAS the host builder made for it)."""
    a, b = _moving_scene(0.0), _moving_scene(1.5)
    rays = random_rays(20000, seed=45, box=5.0, spread=4.0)
    hits_a = _gpu_hits(renderer, a, A.PT_TLAS_BUILD_DEVICE, rays)
    _check_against_oracle(hits_a, RefScene(a.desc()), rays, "t=0", _torus_ids(a))
    rc = renderer.lib.pt_update_objects(renderer.ctx, C.byref(_without_tlas(b)))
    assert rc == 0, renderer.lib.pt_last_error(renderer.ctx)
    hits_b = {}
    for trav in TRAVERSALS:
        renderer.set_option(A.PT_OPT_TRAVERSAL, trav)
        try:
            hits_b[trav] = renderer.trace_rays(rays)
        finally:
            renderer.set_option(A.PT_OPT_TRAVERSAL, A.PT_TRAVERSAL_WIDE)
    _check_against_oracle(hits_b, RefScene(b.desc()), rays, "t=1.5 (update)", _torus_ids(b))
    fresh = _gpu_hits(renderer, b, A.PT_TLAS_BUILD_DEVICE, rays)  # a full upload of the moved scene
    for trav in TRAVERSALS:
        assert _same_bits(hits_b[trav], fresh[trav]), trav


def test_device_built_tlas_renders_like_uploaded_tlas(renderer):
    scene = SceneBuilder.config(1, seed=14)
    w, h = 96, 64
    seeds = [5, 6, 7, 8, 9]
    p = scene.params(w, h, sample_number=0, seed=seeds[0])
    acc = {}
    try:
        for build in (A.PT_TLAS_BUILD_REFERENCE, A.PT_TLAS_BUILD_DEVICE):
            renderer.set_option(A.PT_OPT_TLAS_BUILD, build)
            renderer.set_scene(_without_tlas(scene) if build else scene)
            renderer.resize(w, h)
            renderer.render_section(p, passes=len(seeds), seeds=seeds)
            acc[build] = renderer.read_accum()
    finally:
        renderer.set_option(A.PT_OPT_TLAS_BUILD, A.PT_TLAS_BUILD_REFERENCE)
    same = np.all(acc[0] == acc[1], axis=-1).mean()
    print("device vs uploaded TLAS: identical pixels", same)
    assert same >= 0.999


def test_tlas_option_and_empty_tlas_errors(renderer):
    with pytest.raises(A.PTError):
        renderer.set_option(A.PT_OPT_TLAS_BUILD, 7)
    with pytest.raises(A.PTError):  # the host conversion needs the TLAS
        renderer.set_scene(_without_tlas(SceneBuilder.config(1, seed=3)))


# ---- scenes from ingested assets (RGBE sky, OBJ mesh) ----

def _octahedron_obj():
    """An OBJ text: a subdivided octahedron with per-vertex normals, quads
    on one half (fan-triangulated by the reader) and v/vt/vn corners."""
    lines = ["# octahedron", "o oct"]
    verts = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    for v in verts:
        lines.append("v %.6f %.6f %.6f" % v)
    for v in verts:
        lines.append("vn %.6f %.6f %.6f" % v)
    lines += ["vt 0 0", "vt 1 0", "vt 0.5 1"]
    faces = [(0, 2, 4), (2, 1, 4), (1, 3, 4), (3, 0, 4), (2, 0, 5), (1, 2, 5), (3, 1, 5), (0, 3, 5)]
    for f in faces:
        lines.append("f " + " ".join("%d/%d/%d" % (i + 1, k + 1, i + 1) for k, i in enumerate(f)))
    return "\n".join(lines) + "\n"


def test_render_scene_from_ingested_assets_matches_oracle(renderer):
    """A scene whose sky is a Radiance RGBE file and whose meshes come from an
    OBJ text, both read by the scene library: the device's 1-spp render
    agrees with the oracle's on the same packed scene."""
    from helpers import hdr_bytes
    rng = np.random.default_rng(9)
    sky = rng.integers(0, 256, (32, 64, 4)).astype(np.uint8)
    sky[..., 3] = rng.integers(124, 131, (32, 64))
    sky[6:9, 40:44, 3] = 138  # a small bright sun
    b = SceneBuilder()
    b.sky_from_hdr(hdr_bytes(sky))
    mats = [b.material(color=(0.8, 0.6, 0.4), roughness=0.4), b.material(color=(0.3, 0.7, 0.9), metallic=1.0,
                                                                       roughness=0.2)]
    mesh = b.model_from_obj(_octahedron_obj())
    for k in range(6):
        b.instance(mesh, mats[k % 2], (float(k - 2.5) * 0.9, 0.4, float(k % 3) - 1.0), scale=(0.4, 0.4, 0.4))
    b.plane(mats[0], (0.0, 1.0, 0.0), -0.2)
    b.camera()
    b.bounces(4)
    scene = b.finalize()
    d = scene.desc()
    assert d.sky_width == 64 and d.sky_height == 32 and d.triangles.count == 8
    w, h = 96, 64
    ref = _setup(renderer, scene, w, h)
    p = scene.params(w, h, sample_number=0, seed=99)
    st = renderer.render_section(p, passes=1)
    ga = renderer.read_accum()
    ra, _, rst = ref.render(p, passes=1)
    _assert_1spp(ga, ra, "ingested assets")
    assert rst.tri_tests > 0 and abs(int(st.extension_rays) - int(rst.extension_rays)) <= 0.002 * rst.extension_rays


# ---- material textures ----

def _tex_layer(kind, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:512, 0:512]
    t = np.empty((512, 512, 4), np.uint8)
    if kind == "checker":
        c = ((xx // 16 + yy // 16) % 2).astype(np.uint8)
        t[..., 0], t[..., 1], t[..., 2] = 40 + 200 * c, 90 + 120 * c, 200 - 150 * c
    elif kind == "ramp":
        t[..., 0] = (xx // 2).astype(np.uint8)
        t[..., 1] = (yy // 2).astype(np.uint8)
        t[..., 2] = ((xx + yy) // 4).astype(np.uint8)
    else:
        t[..., :3] = rng.integers(0, 256, (512, 512, 3)).astype(np.uint8)
    t[..., 3] = 255
    return t


def _textured_scene():
    """Albedo, roughness, metallic and emissive maps on every primitive type
    (the light's emission is textured too), with uv scales."""
    b = SceneBuilder()
    b.procedural_sky(128, 64, seed=5)
    checker, ramp, noise = b.texture(_tex_layer("checker", 1)), b.texture(_tex_layer("ramp", 2)), \
        b.texture(_tex_layer("noise", 3))
    m_plane = b.material(color=(0.9, 0.9, 0.9), roughness=0.8)
    b.material_maps(m_plane, albedo=checker, roughness=noise, uv_scale=(0.25, 0.25))
    m_sph = b.material(color=(1.0, 0.9, 0.8), roughness=0.3, metallic=0.8)
    b.material_maps(m_sph, albedo=ramp, metallic=noise, roughness=checker, uv_scale=(2.0, 1.0))
    m_box = b.material(color=(0.7, 0.8, 0.9), roughness=0.5)
    b.material_maps(m_box, albedo=noise)
    m_light = b.material(color=(1.0, 1.0, 1.0), emittance=(1.0, 0.9, 0.8), emission_intensity=12.0)
    b.material_maps(m_light, emissive=ramp)
    m_plain = b.material(color=(0.5, 0.6, 0.3), roughness=0.4)
    m_mesh = b.material(color=(0.8, 0.5, 0.5), roughness=0.3)
    b.material_maps(m_mesh, albedo=checker, metallic=ramp)
    b.plane(m_plane, (0.0, 1.0, 0.0), -0.5)
    for k in range(4):
        b.sphere(m_sph, (float(k) - 1.5, 0.2, -0.5), 0.35)
    b.cube(m_box, (1.2, 0.1, 0.8), (0.3, 0.3, 0.3))
    b.cube(m_light, (0.0, 2.2, 0.0), (0.8, 0.01, 0.8), enable_nee=True)
    b.cylinder(m_plain, (-1.3, -0.5, 0.9), 0.3, 0.8)
    b.torus(m_plain, (0.3, 0.0, 1.2), 0.4, 0.12)
    mesh = b.model_from_obj(_octahedron_obj())
    b.instance(mesh, m_mesh, (-0.4, 0.3, 1.0), scale=(0.35, 0.35, 0.35))
    b.camera()
    b.bounces(5)
    return b.finalize()


def test_textured_scene_matches_oracle(renderer):
    """Textured materials (runs on the kernels' run-time-flags variant): the
    1-spp render and the G-buffer agree with the oracle, both backends give
    the same bits, and the textures change the image."""
    scene = _textured_scene()
    d = scene.desc()
    assert d.texture_layers == 3
    w, h = 96, 64
    ref = _setup(renderer, scene, w, h, gbuf=True)
    p = scene.params(w, h, sample_number=0, seed=2024)
    st = renderer.render_section(p, passes=1)
    ga = renderer.read_accum()
    ra, _, rst = ref.render(p, passes=1)
    assert renderer.get_option(A.PT_OPT_SPECIALISED) == 0  # tori and meshes: the run-time-flags variant
    _assert_1spp(ga, ra, "textured", textured=True)
    _assert_device_paths(renderer, ref, scene, p)
    assert abs(int(st.extension_rays) - int(rst.extension_rays)) <= 0.002 * rst.extension_rays
    # G-buffer (textured albedo, roughness, metallic)
    renderer.gbuffer(scene.params(w, h))
    gg = renderer.read_gbuffer()
    rg, _ = ref.gbuffer(scene.params(w, h))
    gerr = np.abs(gg - rg) / np.maximum(np.abs(rg), 1e-3)
    print("gbuffer max rel err", float(gerr.max()))
    assert float(np.mean(np.all(gerr <= 1e-4, axis=-1))) >= 0.99
    # both backends, same bits
    renderer.reset()
    accs = []
    for be in (A.PT_BACKEND_WAVEFRONT, A.PT_BACKEND_MEGAKERNEL):
        renderer.set_option(A.PT_OPT_BACKEND, be)
        renderer.resize(w, h)
        renderer.render_section(p, passes=2, seeds=[11, 12])
        accs.append(renderer.read_accum())
    renderer.set_option(A.PT_OPT_BACKEND, A.PT_BACKEND_WAVEFRONT)
    assert np.array_equal(accs[0], accs[1])
    # without the maps the image differs
    plain = scene.desc()
    mats = (A.Material * plain.materials.count).from_address(plain.materials.ptr)
    saved = [(m.albedo_idx, m.roughness_idx, m.metallic_idx, m.emissive_idx) for m in mats]
    try:
        for m in mats:
            m.albedo_idx = m.roughness_idx = m.metallic_idx = m.emissive_idx = -1
        renderer.set_scene(plain, isgbuf=True)
        renderer.resize(w, h)
        renderer.render_section(p, passes=1)
        assert np.mean(np.abs(renderer.read_accum() - ga)) > 0.02 * np.mean(np.abs(ga))
    finally:
        for m, s in zip(mats, saved):
            m.albedo_idx, m.roughness_idx, m.metallic_idx, m.emissive_idx = s


def _bump_layers(seed):
    """A smooth height field and its tangent-space normal map (rgb = n * 0.5 + 0.5)."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:512, 0:512].astype(np.float64) * (2 * np.pi / 512)
    hgt = np.zeros((512, 512))
    for _ in range(4):
        fx, fy = rng.integers(1, 6, 2)
        ph = rng.uniform(0, 2 * np.pi)
        hgt += np.sin(fx * xx + fy * yy + ph) / 4
    hgt = (hgt - hgt.min()) / (hgt.max() - hgt.min())
    gy, gx = np.gradient(hgt * 40.0)
    n = np.stack([-gx, -gy, np.ones_like(gx)], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    nt = np.empty((512, 512, 4), np.uint8)
    nt[..., :3] = np.round((n * 0.5 + 0.5) * 255)
    nt[..., 3] = 255
    ht = np.empty((512, 512, 4), np.uint8)
    ht[..., :3] = np.round(hgt * 255)[..., None]
    ht[..., 3] = 255
    return nt, ht


def _pom_scene():
    """Normal and height maps (parallax occlusion, its shadow march in NEE
    and after each BSDF sample) on planes, spheres and cubes, beside plain
    cylinders and tori, a light box and an importance-sampled sky."""
    b = SceneBuilder()
    b.procedural_sky(128, 64, seed=9)
    nt, ht = _bump_layers(4)
    nrm, hgt, checker = b.texture(nt), b.texture(ht), b.texture(_tex_layer("checker", 1))
    m_plane = b.material(color=(0.9, 0.9, 0.9), roughness=0.6)
    b.material_maps(m_plane, albedo=checker, normal=nrm, height=hgt, uv_scale=(0.5, 0.5))
    b.height_params(m_plane, normal_multiplier=1.0, height_multiplier=0.08, layers=24.0)
    m_n = b.material(color=(0.8, 0.6, 0.4), roughness=0.25, metallic=0.5)
    b.material_maps(m_n, normal=nrm)
    b.height_params(m_n, normal_multiplier=2.0)
    m_h = b.material(color=(0.5, 0.7, 0.9), roughness=0.5)
    b.material_maps(m_h, height=hgt, albedo=checker, uv_scale=(2.0, 1.0))
    b.height_params(m_h, height_multiplier=-0.06, layers=16.0, offset=0.3)
    m_nh = b.material(color=(0.9, 0.9, 0.9), roughness=0.15, transmission=0.7, ior=1.4)
    b.material_maps(m_nh, normal=nrm, height=hgt)
    b.height_params(m_nh, normal_multiplier=0.7, height_multiplier=0.05, layers=8.0)
    m_light = b.material(color=(1.0, 1.0, 1.0), emittance=(1.0, 0.9, 0.8), emission_intensity=12.0)
    m_plain = b.material(color=(0.5, 0.6, 0.3), roughness=0.4)
    # a dim emitter with a height map: a BSDF sample the height field blocks
    # adds its emission after the NEE contributions, in both backends
    m_he = b.material(color=(0.8, 0.7, 0.5), roughness=0.5, emittance=(1.0, 0.6, 0.3), emission_intensity=0.5)
    b.material_maps(m_he, height=hgt, uv_scale=(3.0, 2.0))
    b.height_params(m_he, height_multiplier=0.08, layers=16.0)
    b.plane(m_plane, (0.0, 1.0, 0.0), -0.5)
    b.sphere(m_he, (-0.4, 0.9, 0.6), 0.3)
    b.sphere(m_n, (-1.2, 0.2, -0.4), 0.4)
    b.sphere(m_h, (0.0, 0.2, -0.6), 0.45)
    b.sphere(m_nh, (1.1, 0.15, -0.3), 0.35)
    b.cube(m_h, (1.0, 0.0, 0.9), (0.5, 0.5, 0.5))
    b.cube(m_n, (-1.0, 0.0, 0.9), (0.4, 0.6, 0.4))
    b.cube(m_light, (0.0, 2.2, 0.0), (0.8, 0.01, 0.8), enable_nee=True)
    b.cylinder(m_plain, (-0.2, -0.5, 1.1), 0.25, 0.6)
    b.torus(m_plain, (0.4, 0.0, 1.4), 0.3, 0.1)
    b.camera()
    b.bounces(5)
    return b.finalize()


def test_pom_scene_matches_oracle(renderer):
    """Normal maps and parallax occlusion mapping: the 1-spp render and the
    G-buffer agree with the oracle, both backends give the same bits, and
    each kind of map changes the image."""
    scene = _pom_scene()
    d = scene.desc()
    assert d.flags & A.HAS_HEIGHTMAPS
    w, h = 96, 64
    ref = _setup(renderer, scene, w, h, gbuf=True)
    p = scene.params(w, h, sample_number=0, seed=4242)
    st = renderer.render_section(p, passes=1)
    ga = renderer.read_accum()
    ra, _, rst = ref.render(p, passes=1)
    _assert_1spp(ga, ra, "pom", textured=True)
    _assert_device_paths(renderer, ref, scene, p)
    assert abs(int(st.extension_rays) - int(rst.extension_rays)) <= 0.002 * rst.extension_rays
    assert abs(int(st.shadow_rays) - int(rst.shadow_rays)) <= 0.002 * rst.shadow_rays
    renderer.gbuffer(scene.params(w, h))
    gg = renderer.read_gbuffer()
    rg, _ = ref.gbuffer(scene.params(w, h))
    gerr = np.abs(gg - rg) / np.maximum(np.abs(rg), 1e-3)
    print("pom gbuffer max rel err", float(gerr.max()))
    assert float(np.mean(np.all(gerr <= 1e-4, axis=-1))) >= 0.99
    renderer.reset()
    accs = []
    for be in (A.PT_BACKEND_WAVEFRONT, A.PT_BACKEND_MEGAKERNEL):
        renderer.set_option(A.PT_OPT_BACKEND, be)
        renderer.resize(w, h)
        renderer.render_section(p, passes=2, seeds=[21, 22])
        accs.append(renderer.read_accum())
    renderer.set_option(A.PT_OPT_BACKEND, A.PT_BACKEND_WAVEFRONT)
    assert np.array_equal(accs[0], accs[1])
    # each map matters: the image without it differs (the oracle agrees on that image too)
    mats = (A.Material * d.materials.count).from_address(d.materials.ptr)
    for field in ("normal_idx", "height_idx"):
        saved = [getattr(m, field) for m in mats]
        try:
            for m in mats:
                setattr(m, field, -1)
            renderer.set_scene(d, isgbuf=True)
            renderer.resize(w, h)
            renderer.render_section(p, passes=1)
            gb = renderer.read_accum()
            assert np.mean(np.abs(gb - ga)) > 0.01 * np.mean(np.abs(ga)), field
        finally:
            for m, v in zip(mats, saved):
                setattr(m, field, v)


def _alpha_layer(seed):
    """rgb noise with an alpha pattern holding opaque (255), clear (0) and
    partial regions."""
    t = _tex_layer("noise", seed)
    yy, xx = np.mgrid[0:512, 0:512]
    t[..., 3] = np.clip(128 + 220 * np.sin(xx / 37.0 + seed) * np.cos(yy / 29.0), 0, 255).astype(np.uint8)
    return t


def _translucent_scene():
    """Translucent albedo textures (cut-outs with partial alpha) on a plane,
    spheres, cubes, a cylinder, a torus and a mesh, beside opaque objects, a
    light box and an importance-sampled sky."""
    b = SceneBuilder()
    b.procedural_sky(128, 64, seed=3)
    leaf, lace = b.texture(_alpha_layer(1)), b.texture(_alpha_layer(2))
    m_a = b.material(color=(0.9, 0.8, 0.7), roughness=0.5)
    b.material_maps(m_a, albedo=leaf, uv_scale=(2.0, 1.0))
    m_b = b.material(color=(0.6, 0.8, 0.9), roughness=0.3, metallic=0.3)
    b.material_maps(m_b, albedo=lace)
    m_floor = b.material(color=(0.8, 0.8, 0.8), roughness=0.7)
    b.material_maps(m_floor, albedo=lace, uv_scale=(0.5, 0.5))
    m_plain = b.material(color=(0.5, 0.6, 0.3), roughness=0.4)
    m_light = b.material(color=(1.0, 1.0, 1.0), emittance=(1.0, 0.9, 0.8), emission_intensity=12.0)
    b.plane(m_floor, (0.0, 1.0, 0.0), -0.5)
    b.plane(m_plain, (0.0, 0.0, 1.0), -3.0)
    b.sphere(m_a, (-1.0, 0.3, 0.0), 0.5)
    b.sphere(m_plain, (1.4, 0.0, -1.0), 0.4)
    b.cube(m_b, (0.9, 0.2, 0.3), (0.5, 0.5, 0.5))
    b.cylinder(m_a, (-0.2, -0.5, 0.9), 0.35, 0.9)
    b.torus(m_b, (0.3, 0.6, -0.6), 0.45, 0.15)
    b.instance(b.model_from_obj(_octahedron_obj()), m_a, (-0.4, 0.9, 0.6), scale=(0.35, 0.35, 0.35))
    b.cube(m_light, (0.0, 2.2, 0.0), (0.8, 0.01, 0.8), enable_nee=True)
    b.camera()
    b.bounces(5)
    return b.finalize()


def test_translucent_shadows_match_oracle(renderer):
    """Albedo textures with alpha below 1: shadow rays attenuate by (1 - alpha)
    at every surface they cross (exact transmittance per ray against the
    oracle), and the 1-spp render, the G-buffer (seen-through surfaces) and
    both backends agree."""
    scene = _translucent_scene()
    w, h = 96, 64
    ref = _setup(renderer, scene, w, h, gbuf=True)
    rays = random_rays(20000, seed=11, box=2.5, spread=2.0)
    rng = np.random.default_rng(5)
    tgt = rng.integers(-1, 8, size=len(rays)).astype(np.int32)
    dist = rng.uniform(0.5, 8.0, size=len(rays)).astype(np.float32)
    g = renderer.trace_shadow_rays(rays, tgt, dist)
    r = ref.trace_shadow_rays(rays, tgt, dist)
    close = np.abs(g - r) <= 1e-5 + 1e-4 * np.abs(r)
    partial = (r[:, 0] > 0.0) & (r[:, 0] < 1.0)
    print("transmittance agreement", float(close.all(-1).mean()), "partial", float(partial.mean()),
          "blocked", float((r[:, 0] == 0).mean()))
    assert close.all(-1).mean() >= 0.999
    assert partial.mean() > 0.05 and (r[:, 0] == 0).mean() > 0.05
    p = scene.params(w, h, sample_number=0, seed=99)
    st = renderer.render_section(p, passes=1)
    ga = renderer.read_accum()
    ra, _, rst = ref.render(p, passes=1)
    _assert_1spp(ga, ra, "translucent", textured=True)
    _assert_device_paths(renderer, ref, scene, p)
    assert abs(int(st.extension_rays) - int(rst.extension_rays)) <= 0.002 * rst.extension_rays
    renderer.gbuffer(scene.params(w, h))
    gg = renderer.read_gbuffer()
    rg, _ = ref.gbuffer(scene.params(w, h))
    gerr = np.abs(gg - rg) / np.maximum(np.abs(rg), 1e-3)
    print("translucent gbuffer max rel err", float(gerr.max()))
    assert float(np.mean(np.all(gerr <= 1e-4, axis=-1))) >= 0.99
    renderer.reset()
    accs = []
    for be in (A.PT_BACKEND_WAVEFRONT, A.PT_BACKEND_MEGAKERNEL):
        renderer.set_option(A.PT_OPT_BACKEND, be)
        renderer.resize(w, h)
        renderer.render_section(p, passes=2, seeds=[31, 32])
        accs.append(renderer.read_accum())
    renderer.set_option(A.PT_OPT_BACKEND, A.PT_BACKEND_WAVEFRONT)
    assert np.array_equal(accs[0], accs[1])


@pytest.mark.parametrize("bad", ["normal", "height", "index", "layers"])
def test_texture_validation(renderer, bad):
    """Rejected: normal and height maps on hits without the reference's own
    tangent frame (cylinders, tori, meshes), texture indices beyond the
    layers and height maps with no parallax layers."""
    b = SceneBuilder()
    layer = _tex_layer("noise", 7)
    t = b.texture(layer)
    m = b.material(color=(0.5, 0.5, 0.5))
    if bad == "normal":
        b.material_maps(m, normal=t)
        b.cylinder(m, (0.0, 0.0, 0.0), 0.3, 0.8)
    elif bad in ("height", "layers"):
        b.material_maps(m, height=t)
        if bad == "height":
            b.torus(m, (0.0, 0.0, 0.0), 0.4, 0.1)
    else:
        b.material_maps(m, albedo=t)
    b.sphere(m, (0.0, 0.0, 0.0), 0.5)
    b.sky()
    b.camera()
    scene = b.finalize()
    d = scene.desc()
    mats = (A.Material * d.materials.count).from_address(d.materials.ptr)
    if bad == "index":
        mats[m].albedo_idx = 5
    if bad == "layers":
        mats[m].height_params[2] = 0.0
    with pytest.raises(A.PTError) as e:
        renderer.set_scene(d)
    want = {"index": A.PT_ERR_LAYOUT, "layers": A.PT_ERR_ARG}.get(bad, A.PT_ERR_UNSUPPORTED)
    assert e.value.code == want, e.value


SPECIALISED_STAGES = {  # stage: (height maps, translucent albedo)
    "albedo": (False, False), "normal": (False, False), "pom": (True, False), "translucent": (True, True),
    "alpha": (False, True)}


@pytest.mark.parametrize("stage", sorted(SPECIALISED_STAGES))
def test_specialised_textured_kernels_match_oracle(renderer, stage):
    """Configuration-1 scenes with textured materials run kernels specialised
    for their flag word (textures; with height maps; either with translucent
    albedo): the scene selects the variant it should, its 1-spp render agrees
    with the oracle and both backends give the same bits."""
    from webgpuraytracer_amd.texture_bench import scene as tex_scene
    scene = tex_scene(stage, 5)
    d = scene.desc()
    height, alpha = SPECIALISED_STAGES[stage]
    assert d.flags == (0x70F | (A.HAS_HEIGHTMAPS if height else 0))
    w, h = 96, 64
    ref = _setup(renderer, scene, w, h)
    want = 0x70F | A.KERNEL_TEXTURES_BIT | (A.HAS_HEIGHTMAPS if height else 0) | (A.KERNEL_TRANSLUCENT_BIT if alpha else 0)
    assert renderer.get_option(A.PT_OPT_KERNEL_FLAGS) == want
    assert renderer.get_option(A.PT_OPT_SPECIALISED) == 1
    p = scene.params(w, h, sample_number=0, seed=7)
    renderer.render_section(p, passes=1)
    ga = renderer.read_accum()
    ra, _, _ = ref.render(p, passes=1)
    _assert_1spp(ga, ra, stage + " specialised", textured=True)
    _assert_device_paths(renderer, ref, scene, p)
    renderer.reset()
    accs = []
    for be in (A.PT_BACKEND_WAVEFRONT, A.PT_BACKEND_MEGAKERNEL):
        renderer.set_option(A.PT_OPT_BACKEND, be)
        renderer.resize(w, h)
        renderer.render_section(p, passes=2, seeds=[3, 4])
        accs.append(renderer.read_accum())
    renderer.set_option(A.PT_OPT_BACKEND, A.PT_BACKEND_WAVEFRONT)
    assert np.array_equal(accs[0], accs[1])
