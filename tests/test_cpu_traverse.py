"""The kernels' traversal code compiled for the CPU (tests/native/
host_traverse.cpp over dev/traverse.h and pt_layout.cpp), checked against
the oracle without a GPU:

* the reference-order traversal over the child-pair nodes must reproduce the
  oracle's hits bit for bit, node-pop counts included;
* the wide traversal, over exact or quantised boxes, must find the same
  closest hit, except where the FP64 arbiter finds two candidates tied, and
  the same shadow answers;
* every quantised box must contain its exact box and stay within two
  quantisation steps of it;
* the wide stack bound computed at upload must never drop a push.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from helpers import compare_hits, hits_to_np, random_rays, small_mesh_scene
from oracle.refcpu import RefScene
from webgpuraytracer_amd import SceneBuilder
from webgpuraytracer_amd import _abi as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "webgpuraytracer_amd", "csrc")


@pytest.fixture(scope="module")
def ht(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("host_traverse") / "libhosttraverse.so")
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
           "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
           os.path.join(ROOT, "tests", "native", "host_traverse.cpp"), os.path.join(CSRC, "pt_layout.cpp"), "-o", out]
    subprocess.run(cmd, check=True, capture_output=True)
    L = C.CDLL(out)
    L.ht_layout_info.argtypes = [C.c_void_p, C.POINTER(C.c_uint32)]
    L.ht_wide_nodes.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.ht_wide_nodes.restype = C.c_long
    L.ht_quant_nodes.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    L.ht_quant_nodes.restype = C.c_long
    L.ht_quantize_node.argtypes = [C.c_void_p, C.c_void_p]
    L.ht_trace_rays.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_uint32, C.c_void_p]
    L.ht_trace_shadow_rays.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_int,
                                       C.c_void_p]
    L.ht_shadow_transmittance.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_int,
                                          C.c_void_p]
    return L


def _trace(ht, desc, rays, wide, cap=0):
    r = np.ascontiguousarray(rays, np.float32)
    out = (A.Hit * r.shape[0])()
    assert ht.ht_trace_rays(C.byref(desc), r.ctypes.data, r.shape[0], int(wide), cap, out) == 0
    return out


def _info(ht, desc):
    info = (C.c_uint32 * 8)()
    assert ht.ht_layout_info(C.byref(desc), info) == 0
    keys = ("tlas_cnodes", "blas_cnodes", "tlas_wnodes", "blas_wnodes", "stack_cap", "wstack_cap", "tlas_depth",
            "blas_depth")
    return dict(zip(keys, list(info)))


def _camera_rays(scene, w, h):
    """Pixel-centre primary rays of the scene camera (shader.wgsl:2284-2287)."""
    p = scene.params(w, h, 0, 1)
    r00, r10, r01, r11 = (np.array(list(getattr(p, k)), np.float64) for k in ("ray00", "ray10", "ray01", "ray11"))
    eye = np.array(list(p.eye), np.float32)
    rays = []
    for y in range(h):
        for x in range(w):
            u, v = (x + 0.5) / w, (y + 0.5) / h
            top, bot = r00 + (r10 - r00) * u, r01 + (r11 - r01) * u
            d = top + (bot - top) * (1 - v)
            d = d / np.linalg.norm(d)
            rays.append(list(eye) + list(d.astype(np.float32)))
    return np.array(rays, np.float32)


SCENES = {
    "c1": lambda: SceneBuilder.config(1, seed=11),
    "mesh": lambda: small_mesh_scene(seed=5),
}


@pytest.mark.parametrize("which", ["c1", "mesh"])
def test_reference_order_matches_oracle_bitwise(ht, which):
    scene = SCENES[which]()
    desc = scene.desc()
    ref = RefScene(desc)
    rays = np.concatenate([random_rays(6000, seed=42, box=4.0 if which == "c1" else 9.0,
                                       spread=3.0 if which == "c1" else 8.0), _camera_rays(scene, 48, 32)])
    g = hits_to_np(_trace(ht, desc, rays, wide=False))
    r = hits_to_np(ref.trace_rays(rays)[0])
    for k in ("o", "m", "prim", "plane", "count"):
        assert np.array_equal(g[k], r[k]), k
    assert np.array_equal(g["t"].view(np.uint32), r["t"].view(np.uint32))


WIDE = {"wide": A.PT_TRAVERSAL_WIDE, "quant": A.PT_TRAVERSAL_WIDE_QUANTIZED}


@pytest.mark.parametrize("kind", sorted(WIDE))
@pytest.mark.parametrize("which", ["c1", "mesh"])
def test_wide_matches_oracle(ht, which, kind):
    scene = SCENES[which]()
    desc = scene.desc()
    ref = RefScene(desc)
    rays = np.concatenate([random_rays(6000, seed=7, box=4.0 if which == "c1" else 9.0,
                                       spread=3.0 if which == "c1" else 8.0), _camera_rays(scene, 48, 32)])
    g = _trace(ht, desc, rays, wide=WIDE[kind])
    r, _ = ref.trace_rays(rays)
    s = compare_hits(g, r, rays, ref)
    print(which, kind, s)
    assert s["exact_t"] >= 0.999 * s["n"]


def test_wide_matches_oracle_large_mesh_scene(ht):
    """The benchmark scene (808,500 triangles, 64 instances), 3000 rays, both
    wide traversals."""
    scene = SceneBuilder.config(2, seed=7, sky=(64, 32))
    desc = scene.desc()
    ref = RefScene(desc)
    rays = random_rays(3000, seed=3, box=9.0, spread=8.0)
    r, _ = ref.trace_rays(rays)
    info = _info(ht, desc)
    for kind in sorted(WIDE):
        g = _trace(ht, desc, rays, wide=WIDE[kind])
        s = compare_hits(g, r, rays, ref)
        print(kind, s, info)
        assert s["exact_t"] >= 0.999 * s["n"]
    # two levels per wide node: about half the depth, at most 3 pushes a level
    assert info["wstack_cap"] <= 3 * (info["tlas_depth"] + info["blas_depth"]) + 2


@pytest.mark.parametrize("which", ["c1", "mesh"])
def test_wide_shadow_equals_reference(ht, which):
    scene = SCENES[which]()
    desc = scene.desc()
    ref = RefScene(desc)
    rays = random_rays(4000, seed=9, box=4.0 if which == "c1" else 9.0, spread=3.0 if which == "c1" else 8.0)
    rng = np.random.default_rng(1)
    n_obj = desc.objects.count
    target = rng.integers(-1, max(n_obj, 1), size=len(rays)).astype(np.int32)
    dist = rng.uniform(0.5, 20.0, size=len(rays)).astype(np.float32)
    occ = {}
    for wide in (0, 1, 2):
        o = np.zeros(len(rays), np.uint8)
        assert ht.ht_trace_shadow_rays(C.byref(desc), rays.ctypes.data, target.ctypes.data, dist.ctypes.data,
                                       len(rays), wide, o.ctypes.data) == 0
        occ[wide] = o
    oracle = ref.trace_shadow_rays(rays, target, dist)[:, 0] == 0.0
    assert np.array_equal(occ[0].astype(bool), oracle)
    assert np.array_equal(occ[1], occ[0])
    assert np.array_equal(occ[2], occ[0])
    assert 0.05 < oracle.mean() < 0.95


@pytest.mark.parametrize("kind", sorted(WIDE))
@pytest.mark.parametrize("which", ["c1", "mesh"])
def test_wide_stack_bound_drops_nothing(ht, which, kind):
    """Results with the computed stack capacity equal those with a 128-entry
    stack (a dropped push would lose a subtree)."""
    scene = SCENES[which]()
    desc = scene.desc()
    info = _info(ht, desc)
    rays = random_rays(6000, seed=21, box=4.0 if which == "c1" else 9.0, spread=3.0 if which == "c1" else 8.0)
    a = hits_to_np(_trace(ht, desc, rays, wide=WIDE[kind]))
    b = hits_to_np(_trace(ht, desc, rays, wide=WIDE[kind], cap=128))
    for k in ("o", "m", "prim", "plane", "t"):
        assert np.array_equal(a[k], b[k]), k
    assert 2 <= info["wstack_cap"] <= 128


def test_wide_nodes_regroup_the_reference_tree(ht):
    """Every wide node has 2-4 children, unused slots hold the empty +inf
    box, and the wide trees reach exactly the reference trees' leaves."""
    scene = small_mesh_scene(seed=5)
    desc = scene.desc()
    info = _info(ht, desc)
    wt = np.zeros((info["tlas_wnodes"], 32), np.uint32)
    wb = np.zeros((info["blas_wnodes"], 32), np.uint32)
    n = ht.ht_wide_nodes(C.byref(desc), wt.ctypes.data if len(wt) else None, wb.ctypes.data if len(wb) else None)
    assert n == len(wt) + len(wb)

    def leaves(nodes):
        cnt = nodes[:, 28]
        assert np.all((cnt >= 2) & (cnt <= 4))
        boxes = nodes[:, :24].view(np.float32).reshape(-1, 6, 4)
        refs = nodes[:, 24:28]
        total = 0
        for i in range(len(nodes)):
            c = int(cnt[i])
            assert np.all(np.isinf(boxes[i, :, c:])), i
            total += int(np.sum(refs[i, :c] >> 31))
        return total

    tlas = np.ctypeslib.as_array((C.c_uint8 * (desc.tlas.count * 32)).from_address(desc.tlas.ptr)).view(np.uint32)
    tlas = tlas.reshape(-1, 8)
    bvh = np.ctypeslib.as_array((C.c_uint8 * (desc.bvh.count * 32)).from_address(desc.bvh.ptr)).view(np.uint32)
    bvh = bvh.reshape(-1, 8)
    # node layout (32 B): aabb_min vec3, count u32, aabb_max vec3, next u32
    tlas_leaves = int(np.sum(tlas[:, 3] > 0))
    bvh_leaves = int(np.sum(bvh[:, 3] > 0))
    assert leaves(wt) == tlas_leaves
    assert leaves(wb) == bvh_leaves  # each distinct BLAS converted once


def _decode_quant(q):
    """Child boxes of QNodes (n x 16 uint32) as the traversal decodes them:
    org + float(q) * 2^e per axis in float32.  Returns (lo, hi, count, full)
    with lo, hi of shape (n, 3, 4)."""
    org = q[:, 0:3].view(np.float32)
    meta = q[:, 3]
    full = (meta & 0x80000000) != 0
    cnt = (meta >> 24) & 7
    words_lo, words_hi = q[:, [4, 5, 6]], q[:, [7, 8, 9]]
    lo = np.zeros((len(q), 3, 4), np.float32)
    hi = np.zeros((len(q), 3, 4), np.float32)
    for a in range(3):
        step = np.ldexp(np.float32(1.0), ((meta >> (8 * a)) & 255).astype(np.int32) - 128).astype(np.float32)
        for i in range(4):
            ql = ((words_lo[:, a] >> (8 * i)) & 255).astype(np.float32)
            qh = ((words_hi[:, a] >> (8 * i)) & 255).astype(np.float32)
            lo[:, a, i] = org[:, a] + ql * step
            hi[:, a, i] = org[:, a] + qh * step
    return lo, hi, cnt, full, step


def _check_quant(w, q):
    """Every decoded box contains the WNode's box and lies within two steps
    of it; counts and refs carry over; nothing is flagged in a finite tree."""
    lo, hi, cnt, full, _ = _decode_quant(q)
    assert not full.any()
    assert np.array_equal(cnt, w[:, 28])
    assert np.array_equal(q[:, 12:16], w[:, 24:28])
    boxes = w[:, :24].view(np.float32).reshape(-1, 6, 4)
    org = q[:, 0:3].view(np.float32)
    meta = q[:, 3]
    for a in range(3):
        e = ((meta >> (8 * a)) & 255).astype(np.int64) - 128
        step = np.ldexp(1.0, e)
        for i in range(4):
            used = cnt > i
            wl, wh = boxes[used, a, i], boxes[used, 3 + a, i]
            assert np.all(lo[used, a, i] <= wl) and np.all(hi[used, a, i] >= wh)
            # one step, plus the rounding of org + q * step to float32
            assert np.all(wl - lo[used, a, i] <= step[used] + np.spacing(np.abs(wl)))
            assert np.all(hi[used, a, i] - wh <= step[used] + np.spacing(np.abs(wh)))
        # the step is the smallest power of two giving 253 steps over the extent
        ext = boxes[:, 3 + a, :].astype(np.float64).copy()
        lo_ax = boxes[:, a, :].astype(np.float64).copy()
        for i in range(4):
            ext[cnt <= i, i] = -np.inf
            lo_ax[cnt <= i, i] = np.inf
        span = ext.max(axis=1) - lo_ax.min(axis=1)
        assert np.array_equal(org[:, a].astype(np.float64), lo_ax.min(axis=1))
        pos = span > 0
        assert np.all(span[pos] <= 253 * step[pos])
        assert np.all((span[pos] >= 253 * step[pos] / 2) | (e[pos] == -100))


@pytest.mark.parametrize("which", ["c1", "mesh"])
def test_quantised_nodes_contain_exact_boxes(ht, which):
    scene = SCENES[which]()
    desc = scene.desc()
    info = _info(ht, desc)
    wt = np.zeros((info["tlas_wnodes"], 32), np.uint32)
    wb = np.zeros((info["blas_wnodes"], 32), np.uint32)
    qt = np.zeros((info["tlas_wnodes"], 16), np.uint32)
    qb = np.zeros((info["blas_wnodes"], 16), np.uint32)
    ht.ht_wide_nodes(C.byref(desc), wt.ctypes.data if len(wt) else None, wb.ctypes.data if len(wb) else None)
    n = ht.ht_quant_nodes(C.byref(desc), qt.ctypes.data if len(qt) else None, qb.ctypes.data if len(qb) else None)
    assert n == len(qt) + len(qb)
    for w, q in ((wt, qt), (wb, qb)):
        if len(w):
            _check_quant(w, q)


def test_quantiser_edge_cases(ht):
    """Degenerate, far-from-origin, tiny and huge child boxes; a non-finite
    bound or an extent beyond 2^120 flags the node for the full box."""
    inf = np.float32(np.inf)

    def node(boxes):
        w = np.zeros(32, np.uint32)
        f = w[:24].view(np.float32).reshape(6, 4)
        f[:3, :] = inf
        f[3:, :] = -inf
        for i, (lo, hi) in enumerate(boxes):
            f[:3, i] = lo
            f[3:, i] = hi
        w[24:28] = [0x80000000 | (7 + i) for i in range(4)]
        w[28] = len(boxes)
        return w

    def quant(w):
        q = np.zeros(16, np.uint32)
        ht.ht_quantize_node(w.ctypes.data, q.ctypes.data)
        return q

    cases = [
        [((0, 0, 0), (0, 0, 0)), ((0, 0, 0), (0, 0, 0))],                        # points: zero extent
        [((1e6, 1e6, 1e6), (1e6 + 0.125, 1e6, 1e6 + 3)), ((1e6 - 1, 1e6, 1e6), (1e6, 1e6 + 1, 1e6))],
        [((1e-30, 0, 0), (2e-30, 1, 1)), ((3e-30, 0, 0), (4e-30, 1, 1)), ((-1, -1, -1), (-1, -1, -1))],
        [((-3e37, -1, -1), (3e37, 1, 1)), ((0, 0, 0), (1, 1, 1)), ((5, 5, 5), (6, 6, 6)),
         ((-7, 2, 2), (-6.5, 2.25, 9))],
        [((0.1, 0.2, 0.3), (0.1000001, 0.2000001, 0.3000001)), ((0.1, 0.2, 0.3), (0.1, 0.2, 0.3))],
        [((1e-30, 0, 0), (2e-30, 0, 0)), ((3e-30, 0, 0), (4e-30, 0, 0))],       # step clamped at 2^-100
    ]
    for boxes in cases:
        w = node([(np.float32(lo), np.float32(hi)) for lo, hi in
                  [(np.array(l, np.float32), np.array(h, np.float32)) for l, h in boxes]])
        _check_quant(w[None, :], quant(w)[None, :])
    for bad in ([((0, 0, 0), (inf, 1, 1)), ((0, 0, 0), (1, 1, 1))],
                [((-3e38, 0, 0), (3e38, 1, 1)), ((0, 0, 0), (1, 1, 1))]):
        w = node([(np.array(l, np.float32), np.array(h, np.float32)) for l, h in bad])
        q = quant(w)
        assert q[3] & 0x80000000 and (q[3] >> 24) & 7 == 2
        assert np.array_equal(q[12:16], w[24:28])


_OCTA_OBJ = b"""v 1 0 0
v -1 0 0
v 0 1 0
v 0 -1 0
v 0 0 1
v 0 0 -1
vt 0 0
vt 1 0
vt 0.5 1
f 1/1 3/2 5/3
f 3/1 2/2 5/3
f 2/1 4/2 5/3
f 4/1 1/2 5/3
f 3/1 1/2 6/3
f 2/1 3/2 6/3
f 4/1 2/2 6/3
f 1/1 4/2 6/3
"""


def _alpha_texture(seed):
    rng = np.random.default_rng(seed)
    t = rng.integers(0, 256, (512, 512, 4)).astype(np.uint8)
    yy, xx = np.mgrid[0:512, 0:512]
    t[..., 3] = np.clip(128 + 220 * np.sin(xx / 31.0 + seed) * np.cos(yy / 23.0), 0, 255).astype(np.uint8)
    return t


def test_translucent_shadows_equal_oracle(ht):
    """Shadow queries through translucent albedo textures on every primitive
    type (the mesh both as a TLAS object and as a list mesh): both traversal
    orders give the oracle's transmittance, up to the order of the product."""
    b = SceneBuilder()
    lace = b.texture(_alpha_texture(1))
    m = b.material(color=(0.8, 0.8, 0.8))
    b.material_maps(m, albedo=lace, uv_scale=(1.5, 1.0))
    m_opaque = b.material(color=(0.5, 0.5, 0.5))
    b.plane(m, (0.0, 1.0, 0.0), -1.5)
    for k in range(6):
        b.sphere(m, (float(k) - 2.5, 0.0, -1.0), 0.45)
        b.cube(m, (float(k) - 2.5, 0.0, 1.0), (0.35, 0.35, 0.35))
    b.cylinder(m, (-1.0, -1.0, 0.0), 0.4, 1.2)
    b.torus(m, (1.0, 0.0, 0.0), 0.5, 0.15)
    b.sphere(m_opaque, (0.0, 1.5, 0.0), 0.3)
    tri = b.model_from_obj(_OCTA_OBJ)
    b.instance(tri, m, (0.0, 0.8, 0.5), scale=(0.5, 0.5, 0.5))
    b.instance(tri, m, (-2.0, 0.8, 0.0), scale=(0.4, 0.4, 0.4), in_scene_bvh=False)
    b.sky()
    b.camera()
    scene = b.finalize()
    desc = scene.desc()
    assert desc.flags & A.HAS_LIST_MESHES
    rays = random_rays(6000, seed=4, box=3.0, spread=2.5)
    rng = np.random.default_rng(8)
    target = rng.integers(-1, desc.objects.count, size=len(rays)).astype(np.int32)
    dist = rng.uniform(0.5, 8.0, size=len(rays)).astype(np.float32)
    want = RefScene(desc).trace_shadow_rays(rays, target, dist)[:, 0]
    for wide in (0, 1, 2):
        got = np.zeros(len(rays), np.float32)
        assert ht.ht_shadow_transmittance(C.byref(desc), rays.ctypes.data, target.ctypes.data, dist.ctypes.data,
                                          len(rays), wide, got.ctypes.data) == 0
        close = np.abs(got - want) <= 1e-6 + 1e-4 * want
        print("wide", wide, "agreement", float(close.mean()))
        assert close.mean() >= 0.999
    partial = (want > 0.0) & (want < 1.0)
    assert partial.mean() > 0.1 and (want == 0.0).mean() > 0.02
