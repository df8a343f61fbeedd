// geometry.h — scene layout in HBM and ray/primitive intersection.
//
// Device layout built at upload (pt_api.cpp) from the tracer.js arrays:
//   * BVH (TLAS and every BLAS): 64-byte child-pair nodes.  One CNode per
//     inner node of the reference tree holds both children's boxes and
//     their references, so an inner visit is one 64 B fetch instead of the
//     reference's three 32 B node reads.  Visit order, culling and the
//     per-pop counter are the reference's (shader.wgsl:473-542, 694-752).
//   * Triangles: 48 B (v0, v1-v0, v2-v0) per triangle for the test; the
//     original 128 B records are read only for the winning triangle.
//   * Objects, materials, planes, lights, list meshes: the tracer.js records.
#pragma once
#include "vec.h"
#include "pt_layouts.h"

namespace pt {

// Child reference encoding:
//   bit31 = 0           inner node: bits 0..30 index a CNode
//   bit31 = 1, bit30 = 0 leaf: bits 25..29 = count-1, bits 0..24 = first item
//                        (triangle index relative to the mesh's tri_offset, or
//                        absolute object index for the TLAS)
//   bit31 = 1, bit30 = 1 leaf too large to pack: bits 0..29 index the original
//                        32 B node, whose count/next give the range
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kIndirectBit = 0x40000000u;

struct alignas(16) CNode {
    float4 a;  // left.min.xyz, left.max.x
    float4 b;  // left.max.yz, right.min.xy
    float4 c;  // right.min.z, right.max.xyz
    uint4 d;   // left ref, right ref, unused, unused
};
static_assert(sizeof(CNode) == 64, "CNode is 64 B");

// 4-wide node for the wide traversal (PT_TRAVERSAL_WIDE): two levels of
// the reference tree collapsed into one node, so a ray takes about half the
// dependent node fetches.  Every child box is a box of the reference tree;
// child refs use the encoding above, inner refs indexing WNodes.  Child i
// sits in component i of each field; 128 B, one L2 line.
struct alignas(16) WNode {
    float4 lox, loy, loz, hix, hiy, hiz;
    uint4 ref;
    uint4 meta;  // x: number of children (2..4)
};
static_assert(sizeof(WNode) == 128, "WNode is 128 B");

// The same 4-wide node with its child boxes quantised to 8 bits per bound,
// 64 B: half a WNode, so twice the nodes per cache line and per byte of HBM.
// Bounds are offsets from the node's origin (the children's per-axis
// minimum) in steps of a power of two per axis: the decoded value is
// org + q * 2^e in float (one exact product, one rounding), and the builder
// rounds every lower bound down and every upper bound up in that same
// arithmetic, so each decoded box contains the child's exact box.  The slab
// test on a containing box accepts every ray the exact box accepts (the
// test is monotone in the bounds), so no hit is lost; only entry distances,
// and with them the visit order of ties, move.  QNode i is WNode i (same
// refs); a node whose box is not finite is flagged and read from the WNode.
struct alignas(16) QNode {
    float4 org;  // xyz: origin; w (bits): byte k = exponent of axis k + 128 (k < 3), byte 3 = child count
                 // | kQFullBit
    uint4 q0;    // x, y, z: the children's lower bounds on x, y, z (child i in byte i); w: upper bounds on x
    uint4 q1;    // x, y: upper bounds on y, z; z, w: 0
    uint4 ref;
};
static_assert(sizeof(QNode) == 64, "QNode is 64 B");
constexpr uint32_t kQFullBit = 0x80000000u;

// QNode i from WNode i (the device builder, and the CPU tests).  Per axis
// the origin is the children's minimum and the step 2^e the smallest with 253
// steps covering the children's extent (two steps to spare for the outward
// corrections); every bound is rounded outward in the decoder's own
// arithmetic, org + float(q) * 2^e, so the decoded box contains the exact
// one.  A node with a bound that is not finite, or an extent beyond 2^120,
// is flagged for the full node.
__host__ __device__ inline float qnode_bound(float org, uint32_t q, float s) { return org + float(q) * s; }

__host__ __device__ inline QNode quantize_node(const WNode& w) {
    const uint32_t cnt = w.meta.x < 4u ? w.meta.x : 4u;
    const float lo[3][4] = {{w.lox.x, w.lox.y, w.lox.z, w.lox.w}, {w.loy.x, w.loy.y, w.loy.z, w.loy.w},
                            {w.loz.x, w.loz.y, w.loz.z, w.loz.w}};
    const float hi[3][4] = {{w.hix.x, w.hix.y, w.hix.z, w.hix.w}, {w.hiy.x, w.hiy.y, w.hiy.z, w.hiy.w},
                            {w.hiz.x, w.hiz.y, w.hiz.z, w.hiz.w}};
    bool ok = cnt >= 1u;
    float org[3] = {0.0f, 0.0f, 0.0f};
    int ex[3] = {0, 0, 0};
    uint32_t qlo[3] = {0u, 0u, 0u}, qhi[3] = {0u, 0u, 0u};
    for (int a = 0; a < 3 && ok; ++a) {
        float mn = __builtin_inff(), mxv = -__builtin_inff();
        for (uint32_t c = 0; c < cnt; ++c) {
            ok = ok && __builtin_isfinite(lo[a][c]) && __builtin_isfinite(hi[a][c]) && lo[a][c] <= hi[a][c];
            mn = lo[a][c] < mn ? lo[a][c] : mn;
            mxv = hi[a][c] > mxv ? hi[a][c] : mxv;
        }
        if (!ok) break;
        org[a] = mn;
        const double ext = double(mxv) - double(mn);
        int e = -100;
        if (ext > 0.0) {
            int k;
            (void)frexp(ext / 253.0, &k);  // ext / 253 < 2^k
            e = k > -100 ? k : -100;
        }
        if (e > 120) { ok = false; break; }
        ex[a] = e;
        const float s = ldexpf(1.0f, e);
        for (uint32_t c = 0; c < cnt; ++c) {
            double fl = floor((double(lo[a][c]) - double(mn)) / double(s));
            double ce = ceil((double(hi[a][c]) - double(mn)) / double(s));
            uint32_t ql = fl < 0.0 ? 0u : (fl > 255.0 ? 255u : uint32_t(fl));
            uint32_t qh = ce < 0.0 ? 0u : (ce > 255.0 ? 255u : uint32_t(ce));
            while (ql > 0u && qnode_bound(mn, ql, s) > lo[a][c]) --ql;
            while (qh < 255u && qnode_bound(mn, qh, s) < hi[a][c]) ++qh;
            if (qnode_bound(mn, ql, s) > lo[a][c] || qnode_bound(mn, qh, s) < hi[a][c]) ok = false;
            qlo[a] |= ql << (8u * c);
            qhi[a] |= qh << (8u * c);
        }
    }
    QNode q;
    uint32_t meta = (cnt & 7u) << 24;
    if (ok) {
        for (int a = 0; a < 3; ++a) meta |= uint32_t(ex[a] + 128) << (8 * a);
    } else {
        meta |= kQFullBit;
        org[0] = org[1] = org[2] = 0.0f;
        for (int a = 0; a < 3; ++a) qlo[a] = qhi[a] = 0u;
    }
    uint32_t mbits;
    __builtin_memcpy(&mbits, &meta, 4);
    float mf;
    __builtin_memcpy(&mf, &mbits, 4);
    q.org = make_float4(org[0], org[1], org[2], mf);
    q.q0 = make_uint4(qlo[0], qlo[1], qlo[2], qhi[0]);
    q.q1 = make_uint4(qhi[1], qhi[2], 0u, 0u);
    q.ref = w.ref;
    return q;
}

struct alignas(16) Root {
    float mn[3];
    uint32_t ref;   // CNode tree
    float mx[3];
    uint32_t wref;  // WNode tree
};
static_assert(sizeof(Root) == 32, "Root is 32 B");

struct DevScene {
    const pt_material* mats;
    const pt_object* objs;
    const uint32_t* obj_root;      // per object: Root slot of its BLAS (meshes)
    const CNode* tnodes;
    const CNode* bnodes;
    const WNode* wtnodes;
    const WNode* wbnodes;
    const QNode* qtnodes;          // the wide trees with quantised boxes (PT_TRAVERSAL_WIDE_QUANTIZED)
    const QNode* qbnodes;
    const Root* roots;             // slot 0: TLAS; then one per distinct BLAS root
    const pt_bvh_node* tlas_raw;
    const pt_bvh_node* bvh_raw;
    const float4* tri_geo;         // 3 per triangle, in the uploaded order (child-pair trees)
    const float4* tri_geo_w;       // the same records in the wide trees' leaf order
    const pt_triangle* tris;
    const pt_plane* planes;
    const pt_mesh_instance* meshes;
    const uint32_t* mesh_root;     // per list mesh: Root slot
    const pt_light* lights;
    const uint16_t* sky;           // rgba16f
    const float* cdf;
    const uint32_t* tex;           // material texture array: layers of 512 x 512 rgba8
    uint32_t tex_layers;
    uint32_t n_planes, n_meshes, n_lights, n_objs;
    uint32_t sky_w, sky_h;
    uint32_t flags;
    int32_t bounce_limit;
    uint32_t stack_cap;            // entries per lane, reference-order traversal
    uint32_t wstack_cap;           // entries per lane, wide traversal
};

// Compile-time flag policy: F == kDynamicFlags reads the uploaded flags at
// run time (wave-uniform branches); any other value is a specialisation.
constexpr uint32_t kDynamicFlags = 0xffffffffu;
// Flag sets compiled as constants: the small analytic scene family, the
// instanced-mesh family, and the two together.
constexpr uint32_t kFlagsC1 = PT_HAS_SPHERES | PT_HAS_CUBES | PT_HAS_PLANES | PT_HAS_CYLINDERS | PT_HAS_LIGHTS |
                              PT_HAS_SKYBOX | PT_MIS_SKYBOX;
constexpr uint32_t kFlagsC2 = PT_HAS_MESHES | PT_HAS_SKYBOX | PT_MIS_SKYBOX;
constexpr uint32_t kFlagsC3 = kFlagsC1 | PT_HAS_MESHES;
template <uint32_t F>
struct Flags {
    uint32_t rt;
    __device__ __forceinline__ bool has(uint32_t f) const { return F == kDynamicFlags ? (rt & f) != 0 : (F & f) != 0; }
};
constexpr uint32_t kAnyObject = PT_HAS_SPHERES | PT_HAS_CUBES | PT_HAS_CYLINDERS | PT_HAS_TORI | PT_HAS_MESHES;
// Not a scene flag of the ABI: set in DevScene::flags when a material uses a
// texture, so the kernels specialised for the configuration flag sets (which
// lack it) compile the texture code out.
constexpr uint32_t kTexturesBit = 1u << 20;
// Also internal: set when an albedo texture has an alpha below 1, so that
// shadow queries attenuate through such surfaces instead of stopping at the
// first one.
constexpr uint32_t kTranslucentBit = 1u << 21;
// Configuration-1 scenes with textured materials (and with height maps, and
// either with translucent albedo): specialised like the configuration flag
// sets, with the texture code kept.
constexpr uint32_t kFlagsC1T = kFlagsC1 | kTexturesBit;
constexpr uint32_t kFlagsC1H = kFlagsC1 | kTexturesBit | PT_HAS_HEIGHTMAPS;
constexpr uint32_t kFlagsC1TA = kFlagsC1T | kTranslucentBit;
constexpr uint32_t kFlagsC1HA = kFlagsC1H | kTranslucentBit;

// sample_texture (shader.wgsl:170-173): textureSampleLevel on the material
// texture array, level 0 of a 512 x 512 rgba8unorm layer, linear filtering,
// repeat addressing (tracer.js:1965-1968), restated with fp32 weights
// (texel centres at uv * 512 - 0.5).
constexpr int kTexSize = 512;
__device__ __forceinline__ float4 sample_texture(const DevScene& S, int idx, v2 uv) {
    const float x = uv.x * float(kTexSize) - 0.5f, y = uv.y * float(kTexSize) - 0.5f;
    const float x0 = floorf(x), y0 = floorf(y);
    const float fx = x - x0, fy = y - y0;
    const int ix = int(x0), iy = int(y0);
    const uint32_t* L = S.tex + size_t(idx) * kTexSize * kTexSize;
    const uint32_t xa = uint32_t(ix) & (kTexSize - 1), xb = uint32_t(ix + 1) & (kTexSize - 1);
    const uint32_t ya = uint32_t(iy) & (kTexSize - 1), yb = uint32_t(iy + 1) & (kTexSize - 1);
    const uint32_t q00 = L[ya * kTexSize + xa], q10 = L[ya * kTexSize + xb];
    const uint32_t q01 = L[yb * kTexSize + xa], q11 = L[yb * kTexSize + xb];
    float o[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float t00 = float((q00 >> (8 * c)) & 255u) / 255.0f, t10 = float((q10 >> (8 * c)) & 255u) / 255.0f;
        const float t01 = float((q01 >> (8 * c)) & 255u) / 255.0f, t11 = float((q11 >> (8 * c)) & 255u) / 255.0f;
        const float a = t00 + (t10 - t00) * fx, b = t01 + (t11 - t01) * fx;
        o[c] = a + (b - a) * fy;
    }
    return make_float4(o[0], o[1], o[2], o[3]);
}


// ---------------------------------------------------------------- stack ----
// The stack's LDS part is addressed through an LDS-typed pointer in the device
// compilation, so that its accesses are ds_read / ds_write: through a generic
// pointer (and a select between it and the spill column) they compile to flat
// accesses, each of which waits for every outstanding global load.  The
// traversal code is also compiled for the CPU (tests), where it is a plain
// pointer.
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(3))) uint32_t lds_u32;
#else
typedef uint32_t lds_u32;
#endif
// Per-lane traversal stack.  The first `cap` entries live in LDS,
// interleaved across the block's lanes (entry e of lane l at e*blockDim + l)
// so that a wave's accesses fall in distinct banks; entries cap..total-1, if
// any, spill to a per-lane column in device memory (`spill`, entries
// `sstride` apart), so a kernel can keep a short LDS stack for occupancy and
// still hold the deepest traversal.  Pushes beyond `total` are dropped (the
// host sizes `total` from the tree depths so that none is).
struct LdsStack {
    lds_u32* base;
    uint32_t stride;
    uint32_t cap;
    int sp;
    uint32_t* spill;
    uint32_t sstride;
    uint32_t total;
    __device__ __forceinline__ void push(uint32_t v) {
        const uint32_t e = uint32_t(sp);
        if (e < cap) {
            base[e * stride] = v;
        } else if (e < total) {
            spill[(e - cap) * sstride] = v;
        } else {
            return;
        }
        ++sp;
    }
    __device__ __forceinline__ uint32_t pop() {
        --sp;
        const uint32_t e = uint32_t(sp);
        if (e < cap) return base[e * stride];
        return spill[(e - cap) * sstride];
    }
};
// The LDS view of lane column `p` of a kernel's shared stack array.
__device__ __forceinline__ lds_u32* lds_column(uint32_t* p) { return (lds_u32*)p; }
// A stack held entirely in LDS.
__device__ __forceinline__ LdsStack lds_stack_of(uint32_t* base, uint32_t stride, uint32_t cap) {
    return LdsStack{lds_column(base), stride, cap, 0, nullptr, 0u, cap};
}

struct Counters {
    uint32_t ext, shadow, paths, pops, inner, objs, tris, shade;
    __device__ void clear() { ext = shadow = paths = pops = inner = objs = tris = shade = 0; }
};

// -------------------------------------------------------------- boxes -------
// intersect_aabb, shader.wgsl:256-265
__device__ __forceinline__ float slab(v3 o, v3 inv, v3 bmin, v3 bmax) {
    const v3 t0 = (bmin - o) * inv;
    const v3 t1 = (bmax - o) * inv;
    const float tn = hmax(vmin(t0, t1));
    const float tf = hmin(vmax(t0, t1));
    if (tn > tf || tf < 0.0f) return -1.0f;
    return tn < 0.0f ? 0.0f : tn;
}
__device__ __forceinline__ void cnode_boxes(const CNode& n, v3& lmin, v3& lmax, v3& rmin, v3& rmax) {
    lmin = mk(n.a.x, n.a.y, n.a.z);
    lmax = mk(n.a.w, n.b.x, n.b.y);
    rmin = mk(n.b.z, n.b.w, n.c.x);
    rmax = mk(n.c.y, n.c.z, n.c.w);
}
__device__ __forceinline__ CNode load_cnode(const CNode* p) {
    CNode n;
    const float4* f = reinterpret_cast<const float4*>(p);
    n.a = f[0];
    n.b = f[1];
    n.c = f[2];
    n.d = reinterpret_cast<const uint4*>(p)[3];
    return n;
}
__device__ __forceinline__ void leaf_range(uint32_t ref, const pt_bvh_node* raw, uint32_t& start, uint32_t& count) {
    if (ref & kIndirectBit) {
        const pt_bvh_node& n = raw[ref & 0x3fffffffu];
        start = n.next;
        count = n.count;
    } else {
        start = ref & 0x01ffffffu;
        count = ((ref >> 25) & 31u) + 1u;
    }
}

// ------------------------------------------------------------ triangle ------
// intersect_triangle, shader.wgsl:267-288, with edges precomputed on the
// host (v1-v0 and v2-v0 round identically there).  Returns t > 1e-6 or -1;
// u, v are the barycentrics of v1 and v2.
__device__ __forceinline__ float tri_hit_v(v3 o, v3 d, float4 A, float4 B, float4 C, float& u, float& v) {
    const v3 v0 = mk(A.x, A.y, A.z), e1 = mk(B.x, B.y, B.z), e2 = mk(C.x, C.y, C.z);
    const v3 h = cross(d, e2);
    const float a = dot(e1, h);
    if (fabsf(a) < 0.000001f) return -1.0f;
    const float f = 1.0f / a;
    const v3 s = o - v0;
    u = f * dot(s, h);
    if (u < 0.0f || u > 1.0f) return -1.0f;
    const v3 q = cross(s, e1);
    v = f * dot(d, q);
    if (v < 0.0f || u + v > 1.0f) return -1.0f;
    const float t = f * dot(e2, q);
    return t > 0.000001f ? t : -1.0f;
}
__device__ __forceinline__ float tri_hit(v3 o, v3 d, const float4* g, float& u, float& v) {
    return tri_hit_v(o, d, g[0], g[1], g[2], u, v);
}

// ------------------------------------------------------------ analytic ------
// hit_unit_sphere, shader.wgsl:290-300
__device__ __forceinline__ float unit_sphere(v3 o, v3 d) {
    const float a = dot(d, d);
    const float hb = dot(o, d);
    const float c = dot(o, o) - 1.0f;
    const float disc = hb * hb - a * c;
    if (disc < 0.0f) return -1.0f;
    const float sq = sqrtf(disc);
    float root = (-hb - sq) / a;
    if (root < 0.001f) root = (-hb + sq) / a;
    return root >= 0.001f ? root : -1.0f;
}
// hit_unit_cube, shader.wgsl:302-312
__device__ __forceinline__ float unit_cube(v3 o, v3 d) {
    const v3 inv = splat(1.0f) / d;
    const v3 t0 = (splat(-1.0f) - o) * inv;
    const v3 t1 = (splat(1.0f) - o) * inv;
    const float tn = hmax(vmin(t0, t1)), tf = hmin(vmax(t0, t1));
    if (tn < tf && tf > 0.0f) return tn > 0.001f ? tn : tf;
    return -1.0f;
}
// hit_plane, shader.wgsl:314-321
__device__ __forceinline__ float plane_t(v3 n, float dd, v3 o, v3 d) {
    const float denom = dot(n, d);
    if (fabsf(denom) > 1e-6f) {
        const float t = (dd - dot(n, o)) / denom;
        if (t > 0.001f) return t;
    }
    return -1.0f;
}
// hit_cylinder, shader.wgsl:323-379.  which: 0 side, 1 bottom cap, 2 top cap.
__device__ __forceinline__ float cone_cylinder(v3 o, v3 d, float r0, float r1, int& which, v3& p_side) {
    const float k = (r1 - r0) / 1.0f;
    const float oer = r0 + k * o.y;
    const float A = d.x * d.x + d.z * d.z - k * k * d.y * d.y;
    const float B = 2.0f * (o.x * d.x + o.z * d.z - oer * k * d.y);
    const float C = o.x * o.x + o.z * o.z - oer * oer;
    float tmin = 1e10f;
    which = -1;
    const float disc = B * B - 4.0f * A * C;
    if (disc > 0.0f) {
        const float sq = sqrtf(disc);
        const float ts0 = (-B - sq) / (2.0f * A);
        const float ts1 = (-B + sq) / (2.0f * A);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const float t = i == 0 ? ts0 : ts1;
            const float y = o.y + t * d.y;
            if (t > 0.001f && t < tmin && y >= 0.0f && y <= 1.0f) { tmin = t; which = 0; p_side = o + t * d; }
        }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const float py = i == 0 ? 0.0f : 1.0f;
        const float rr = i == 0 ? r0 : r1;
        const float t = (py - o.y) / d.y;
        if (t > 0.001f && t < tmin) {
            const v3 p = o + t * d;
            if (p.x * p.x + p.z * p.z <= rr * rr) { tmin = t; which = 1 + i; }
        }
    }
    return which >= 0 ? tmin : -1.0f;
}
// hit_torus, shader.wgsl:381-471 (major radius Ra, minor ra, unit direction)
static __device__ __noinline__ float torus_t(v3 o, v3 d, float Ra, float ra) {
    float po = 1.0f;
    const float Ra2 = Ra * Ra, ra2 = ra * ra;
    const float m = dot(o, o), n = dot(o, d);
    const float br = Ra + ra;
    const float hb = n * n - m + br * br;
    if (hb < 0.0f) return -1.0f;
    const float k = (m - ra2 - Ra2) / 2.0f;
    const float k3 = n;
    const float k2 = n * n + Ra2 * d.y * d.y + k;
    const float k1 = k * n + Ra2 * o.y * d.y;
    const float k0 = k * k + Ra2 * o.y * o.y - Ra2 * ra2;
    float K3 = k3, K2 = k2, K1 = k1, K0 = k0;
    if (fabsf(k3 * (k3 * k3 - k2) + k1) < 0.01f) {
        po = -1.0f;
        const float ik0 = 1.0f / k0;
        K1 = k3 * ik0;
        K2 = k2 * ik0;
        K3 = k1 * ik0;
        K0 = ik0;
    }
    const float c2 = (2.0f * K2 - 3.0f * K3 * K3) / 3.0f;
    const float c1 = 2.0f * (K3 * (K3 * K3 - K2) + K1);
    const float c0 = (K3 * (K3 * (-3.0f * K3 * K3 + 4.0f * K2) - 8.0f * K1) + 4.0f * K0) / 3.0f;
    const float Q = c2 * c2 + c0;
    const float R = 3.0f * c0 * c2 - c2 * c2 * c2 - c1 * c1;
    const float h = R * R - Q * Q * Q;
    float z;
    if (h < 0.0f) {
        const float sQ = sqrtf(Q);
        z = 2.0f * sQ * cosf(acosf(clampf(R / (sQ * Q), -1.0f, 1.0f)) / 3.0f);
    } else {
        const float sQ = powf(sqrtf(h) + fabsf(R), 1.0f / 3.0f);
        z = sgn(R) * fabsf(sQ + Q / sQ);
    }
    z = c2 - z;
    float d1 = z - 3.0f * c2;
    float d2 = z * z - 3.0f * c0;
    if (fabsf(d1) < 1.0e-4f) {
        if (d2 < 0.0f) return -1.0f;
        d2 = sqrtf(d2);
    } else {
        if (d1 < 0.0f) return -1.0f;
        d1 = sqrtf(d1 / 2.0f);
        d2 = c1 / d1;
    }
    float result = 1e20f;
    float h1 = d1 * d1 - z + d2;
    if (h1 > 0.0f) {
        h1 = sqrtf(h1);
        float t1 = -d1 - h1 - K3; if (po < 0.0f) t1 = 2.0f / t1;
        float t2 = -d1 + h1 - K3; if (po < 0.0f) t2 = 2.0f / t2;
        if (t1 > 0.001f) result = t1;
        if (t2 > 0.001f) result = mn(result, t2);
    }
    float h2 = d1 * d1 - z - d2;
    if (h2 > 0.0f) {
        h2 = sqrtf(h2);
        float t3 = d1 - h2 - K3; if (po < 0.0f) t3 = 2.0f / t3;
        float t4 = d1 + h2 - K3; if (po < 0.0f) t4 = 2.0f / t4;
        if (t3 > 0.001f) result = mn(result, t3);
        if (t4 > 0.001f) result = mn(result, t4);
    }
    return result > 1e10f ? -1.0f : result;
}

// ----------------------------------------------------- shadow analytics -----
// trace_sphere_shadow's interval test, shader.wgsl:874-902
__device__ __forceinline__ bool sphere_blocks(v3 o, v3 d, float target) {
    const float a = dot(d, d), hb = dot(o, d), c = dot(o, o) - 1.0f;
    const float disc = hb * hb - a * c;
    if (disc < 0.0f) return false;
    const float sq = sqrtf(disc);
    const float t1 = (-hb - sq) / a, t2 = (-hb + sq) / a;
    return mx(0.001f, t1) < mn(target, t2);
}
// trace_cube_shadow, shader.wgsl:924-953
__device__ __forceinline__ bool cube_blocks(v3 o, v3 d, float target) {
    const v3 inv = splat(1.0f) / d;
    const v3 tn = (splat(-1.0f) - o) * inv;
    const v3 tf = (splat(1.0f) - o) * inv;
    const float t1 = hmax(vmin(tn, tf)), t2 = hmin(vmax(tn, tf));
    return mx(0.001f, t1) < mn(target, t2);
}
// trace_cylinder_shadow, shader.wgsl:961-1016
__device__ __forceinline__ bool cylinder_blocks(v3 o, v3 d, float top, float target) {
    const float k = (top - 1.0f) / 1.0f;
    const float oer = 1.0f + k * o.y;
    const float A = d.x * d.x + d.z * d.z - k * k * d.y * d.y;
    const float B = 2.0f * (o.x * d.x + o.z * d.z - oer * k * d.y);
    const float C = o.x * o.x + o.z * o.z - oer * oer;
    float h1 = 1e10f, h2 = -1e10f;
    uint32_t hits = 0;
    const float disc = B * B - 4.0f * A * C;
    if (disc >= 0.0f) {
        const float sq = sqrtf(disc);
        const float ts0 = (-B - sq) / (2.0f * A), ts1 = (-B + sq) / (2.0f * A);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const float t = i == 0 ? ts0 : ts1;
            const float y = o.y + t * d.y;
            if (y >= 0.0f && y <= 1.0f) { h1 = mn(h1, t); h2 = mx(h2, t); ++hits; }
        }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const float cy = i == 0 ? 0.0f : 1.0f;
        const float rr = i == 0 ? 1.0f : top;
        const float t = (cy - o.y) / d.y;
        const v3 p = o + t * d;
        if (p.x * p.x + p.z * p.z <= rr * rr) { h1 = mn(h1, t); h2 = mx(h2, t); ++hits; }
    }
    return hits >= 2u && mx(0.001f, h1) < mn(target, h2);
}
// ------------------------------------------- translucent shadow surfaces --
// evaluate_shadow_attenuation (shader.wgsl:777-811): a surface whose material
// has no albedo texture stops the light; one with a texture passes (1 - alpha)
// of it (alpha of the texel at uv * uv_scale) and stops it above alpha 0.9999.
// Returns true when the light is stopped; tr accumulates the product.
__device__ __forceinline__ bool shadow_surface(const DevScene& S, int mat, v2 uv, float& tr) {
    const pt_material& m = S.mats[mat];
    if (m.albedo_idx < 0) return true;
    const float a = sample_texture(S, m.albedo_idx, v2{uv.x * m.uv_scale[0], uv.y * m.uv_scale[1]}).w;
    tr *= 1.0f - a;
    return a > 0.9999f;
}
// the texture coordinates trace_*_shadow evaluates at (shader.wgsl:893-901,
// 904-922 get_cube_uv, 958-962 get_frustum_uv, 1090-1094 get_torus_uv)
__device__ __forceinline__ v2 sphere_uv(v3 p) {
    return v2{0.5f + atan2f(p.z, p.x) / kTWO_PI, 0.5f + asinf(clampf(p.y, -1.0f, 1.0f)) / kPI};
}
__device__ __forceinline__ v2 cube_uv(v3 p) {
    const v3 a = mk(fabsf(p.x), fabsf(p.y), fabsf(p.z));
    const float m = hmax(a);
    if (m == a.x) {
        const float s = sgn(p.x);
        return v2{(-s * p.z) * 0.5f + 0.5f, p.y * 0.5f + 0.5f};
    }
    if (m == a.y) {
        const float s = sgn(p.y);
        return v2{p.x * 0.5f + 0.5f, (s * p.z) * 0.5f + 0.5f};
    }
    const float s = sgn(p.z);
    return v2{(s * p.x) * 0.5f + 0.5f, p.y * 0.5f + 0.5f};
}
__device__ __forceinline__ v2 frustum_uv(v3 p, float r0, float r1) {
    if (p.y > 0.999f) return v2{p.x / (r1 * 2.0f) + 0.5f, p.z / (r1 * 2.0f) + 0.5f};
    if (p.y < 0.001f) return v2{p.x / (r0 * 2.0f) + 0.5f, p.z / (r0 * 2.0f) + 0.5f};
    return v2{atan2f(p.z, p.x) / kTWO_PI + 0.5f, p.y};
}
__device__ __forceinline__ v2 torus_uv(v3 p) {
    return v2{(atan2f(p.z, p.x) / kTWO_PI) + 0.5f, (atan2f(p.y, sqrtf(p.x * p.x + p.z * p.z) - 1.0f) / kTWO_PI) + 0.5f};
}
// entry and exit of the clipped interval [max(0.001, t1), min(target, t2)]
__device__ __forceinline__ bool interval_shadow(const DevScene& S, int mat, v3 o, v3 d, float h1, float h2, int kind,
                                                float top, float& tr) {
    if (!(h1 < h2)) return false;
    const v3 p1 = o + d * h1, p2 = o + d * h2;
    const v2 u1 = kind == 0 ? sphere_uv(p1) : kind == 1 ? cube_uv(p1) : frustum_uv(p1, 1.0f, top);
    if (shadow_surface(S, mat, u1, tr)) return true;
    const v2 u2 = kind == 0 ? sphere_uv(p2) : kind == 1 ? cube_uv(p2) : frustum_uv(p2, 1.0f, top);
    return shadow_surface(S, mat, u2, tr);
}
// trace_sphere_shadow with attenuation (shader.wgsl:874-902)
__device__ __forceinline__ bool sphere_shadow(const DevScene& S, int mat, v3 o, v3 d, float target, float& tr) {
    const float a = dot(d, d), hb = dot(o, d), c = dot(o, o) - 1.0f;
    const float disc = hb * hb - a * c;
    if (disc < 0.0f) return false;
    const float sq = sqrtf(disc);
    const float t1 = (-hb - sq) / a, t2 = (-hb + sq) / a;
    return interval_shadow(S, mat, o, d, mx(0.001f, t1), mn(target, t2), 0, 0.0f, tr);
}
// trace_cube_shadow with attenuation (shader.wgsl:926-954)
__device__ __forceinline__ bool cube_shadow(const DevScene& S, int mat, v3 o, v3 d, float target, float& tr) {
    const v3 inv = splat(1.0f) / d;
    const v3 tn = (splat(-1.0f) - o) * inv;
    const v3 tf = (splat(1.0f) - o) * inv;
    const float t1 = hmax(vmin(tn, tf)), t2 = hmin(vmax(tn, tf));
    return interval_shadow(S, mat, o, d, mx(0.001f, t1), mn(target, t2), 1, 0.0f, tr);
}
// trace_cylinder_shadow with attenuation (shader.wgsl:964-1016)
__device__ __forceinline__ bool cylinder_shadow(const DevScene& S, int mat, v3 o, v3 d, float top, float target,
                                                float& tr) {
    const float k = (top - 1.0f) / 1.0f;
    const float oer = 1.0f + k * o.y;
    const float A = d.x * d.x + d.z * d.z - k * k * d.y * d.y;
    const float B = 2.0f * (o.x * d.x + o.z * d.z - oer * k * d.y);
    const float C = o.x * o.x + o.z * o.z - oer * oer;
    float h1 = 1e10f, h2 = -1e10f;
    uint32_t hits = 0;
    const float disc = B * B - 4.0f * A * C;
    if (disc >= 0.0f) {
        const float sq = sqrtf(disc);
        const float ts0 = (-B - sq) / (2.0f * A), ts1 = (-B + sq) / (2.0f * A);
        for (int i = 0; i < 2; ++i) {
            const float t = i == 0 ? ts0 : ts1;
            const float y = o.y + t * d.y;
            if (y >= 0.0f && y <= 1.0f) { h1 = mn(h1, t); h2 = mx(h2, t); ++hits; }
        }
    }
    for (int i = 0; i < 2; ++i) {
        const float cy = i == 0 ? 0.0f : 1.0f;
        const float rr = i == 0 ? 1.0f : top;
        const float t = (cy - o.y) / d.y;
        const v3 p = o + t * d;
        if (p.x * p.x + p.z * p.z <= rr * rr) { h1 = mn(h1, t); h2 = mx(h2, t); ++hits; }
    }
    if (hits < 2u) return false;
    return interval_shadow(S, mat, o, d, mx(0.001f, h1), mn(target, h2), 2, top, tr);
}
// a triangle of trace_mesh_shadow (shader.wgsl:840-852) at barycentrics (u, v):
// its interpolated texture coordinates
__device__ __forceinline__ bool tri_shadow(const DevScene& S, int mat, uint32_t prim, float u, float v, float& tr) {
    const pt_triangle& t = S.tris[prim];
    const float b0 = 1.0f - u - v;
    return shadow_surface(S, mat, v2{t.uv0[0] * b0 + t.uv1[0] * u + t.uv2[0] * v, t.uv0[1] * b0 + t.uv1[1] * u + t.uv2[1] * v},
                          tr);
}

// solve_cubic_real / solve_quartic_real, shader.wgsl:1019-1087
__device__ __forceinline__ float cubic_root(float a, float b, float c) {
    const float q = (a * a - 3.0f * b) / 9.0f;
    const float r = (2.0f * a * a * a - 9.0f * a * b + 27.0f * c) / 54.0f;
    const float r2 = r * r, q3 = q * q * q;
    if (r2 < q3) {
        const float th = acosf(clampf(r / sqrtf(q3), -1.0f, 1.0f));
        return -2.0f * sqrtf(q) * cosf(th / 3.0f) - a / 3.0f;
    }
    const float aa = -sgn(r) * powf(fabsf(r) + sqrtf(r2 - q3), 1.0f / 3.0f);
    float bb = 0.0f;
    if (aa != 0.0f) bb = q / aa;
    return (aa + bb) - a / 3.0f;
}
// trace_torus_shadow (shader.wgsl:1096-1132); with ATT each interval's entry
// and exit attenuate as evaluate_shadow_attenuation says (S, mat, tr used
// only then).
template <bool ATT>
static __device__ __noinline__ bool torus_shadow(v3 wo, v3 wd, const float* inv, float rminor, float target,
                                                 const DevScene* S, int mat, float* tr) {
    v3 o = xform_point(inv, wo), d = xform_dir(inv, wd);
    const float rs = len(d);
    d = d / rs;
    const float R = 1.0f, r = rminor;
    const float K = dot(o, o) + R * R - r * r;
    const float G = dot(o, d);
    const float a = 4.0f * G;
    const float b = 2.0f * K + 4.0f * G * G - 4.0f * R * R * (1.0f - d.y * d.y);
    const float c = 4.0f * K * G - 8.0f * R * R * o.y * d.y;
    const float e = K * K - 4.0f * R * R * (o.x * o.x + o.z * o.z);
    float roots[4] = {-1.0f, -1.0f, -1.0f, -1.0f};
    uint32_t count = 0;
    const float a2 = a * a;
    const float y = cubic_root(-b, a * c - 4.0f * e, 4.0f * b * e - a2 * e - c * c);
    const float R2 = 0.25f * a2 - b + y;
    if (R2 < 0.0f) return false;
    const float Rr = sqrtf(R2);
    float D2, E2;
    if (Rr < 1e-6f) {
        D2 = 0.75f * a2 - 2.0f * b + 2.0f * sqrtf(y * y - 4.0f * e);
        E2 = 0.75f * a2 - 2.0f * b - 2.0f * sqrtf(y * y - 4.0f * e);
    } else {
        D2 = 0.75f * a2 - R2 - 2.0f * b + (4.0f * a * b - 8.0f * c - a2 * a) / (4.0f * Rr);
        E2 = 0.75f * a2 - R2 - 2.0f * b - (4.0f * a * b - 8.0f * c - a2 * a) / (4.0f * Rr);
    }
    if (D2 >= 0.0f) {
        const float D = sqrtf(D2);
        roots[0] = -0.25f * a + 0.5f * Rr + 0.5f * D;
        roots[1] = -0.25f * a + 0.5f * Rr - 0.5f * D;
        count += 2u;
    }
    if (E2 >= 0.0f) {
        const float E = sqrtf(E2);
        const float x0 = -0.25f * a - 0.5f * Rr + 0.5f * E;
        const float x1 = -0.25f * a - 0.5f * Rr - 0.5f * E;
        if (count == 0) { roots[0] = x0; roots[1] = x1; } else { roots[2] = x0; roots[3] = x1; }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i + 1; j < 4; ++j)
            if (roots[i] > roots[j]) { const float t = roots[i]; roots[i] = roots[j]; roots[j] = t; }
#pragma unroll
    for (int i = 0; i < 4; i += 2) {
        const float tn = roots[i], tf = roots[i + 1];
        if (tn < 0.0f || tn > target * rs) continue;
        const float h1 = mx(0.001f, tn) / rs;
        const float h2 = mn(target * rs, tf) / rs;
        if (h1 < h2) {
            if (!ATT) return true;
            if (shadow_surface(*S, mat, torus_uv(o + d * (h1 * rs)), *tr)) return true;
            if (shadow_surface(*S, mat, torus_uv(o + d * (h2 * rs)), *tr)) return true;
        }
    }
    return false;
}
__device__ __forceinline__ bool torus_blocks(v3 wo, v3 wd, const float* inv, float rminor, float target) {
    return torus_shadow<false>(wo, wd, inv, rminor, target, nullptr, 0, nullptr);
}

}  // namespace pt
