// traverse.h — closest-hit and any-hit traversal of planes, the TLAS, every
// BLAS and the list meshes.
//
// The reference nests a TLAS loop and a per-mesh BLAS loop, each with its own
// 64-entry private stack (shader.wgsl:473-542, 694-752, 813-872, 1134-1184).
// Here both levels run in one flattened loop over one LDS stack: each
// iteration does one unit of work (pop a TLAS node, intersect one TLAS-leaf
// object, or pop a BLAS node), so lanes of a wave stay in the same loop body
// whichever level they are at.
//
// Two traversals share the loop (template parameter W):
//   * reference order (W = false, child-pair nodes): the per-lane visit
//     order, the culling against the running closest t / target distance,
//     the strict `<` ties and the per-pop counter are exactly the
//     reference's, so hits match it bit for bit;
//   * wide (W = true, 4-wide nodes): children visited nearest first.  Every
//     box and triangle test is the reference's arithmetic, so the closest
//     hit is the same except where two candidates tie to within the
//     rounding of the box and triangle tests, where the visit order picks
//     the winner.  Shadow rays give the same answer either way (any
//     occluder ends the query).
#pragma once
#include "geometry.h"

namespace pt {

enum HitKind : int { kMiss = 0, kPlaneHit = 1, kObjectHit = 2, kListMeshHit = 3 };

// Identity of the closest hit; attributes are computed once for the winner
// (resolve_hit) instead of for every candidate, which gives the same values.
struct HitRec {
    float t;
    uint32_t count;
    int kind;
    int idx;    // plane / object / list-mesh index
    int prim;   // absolute triangle index, -1 otherwise
    float bu, bv;
};

__device__ __forceinline__ uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }

// Inner-node visit.  The reference pushes the accepted children far-first
// and pops the near one at once (shader.wgsl:533-539, 743-749); here the child
// it would pop next is returned in `next` (kept in a register) and only the
// other accepted child goes to the stack.  Same visit order, half the stack
// traffic.  Returns false when neither child is accepted.
template <bool CN>
__device__ __forceinline__ bool visit_inner(const CNode* nodes, uint32_t ref, v3 o, v3 inv, float limit, LdsStack& st,
                                            Counters& c, uint32_t& next) {
    if (CN) c.inner++;
    const CNode n = load_cnode(nodes + ref);
    v3 lmin, lmax, rmin, rmax;
    cnode_boxes(n, lmin, lmax, rmin, rmax);
    const float lt = slab(o, inv, lmin, lmax);
    const float rt = slab(o, inv, rmin, rmax);
    const bool lok = lt >= 0.0f && lt < limit;
    const bool rok = rt >= 0.0f && rt < limit;
    const bool left_top = lt < rt;  // ties: the right child ends on top
    const uint32_t top = left_top ? n.d.x : n.d.y, bot = left_top ? n.d.y : n.d.x;
    const bool top_ok = left_top ? lok : rok, bot_ok = left_top ? rok : lok;
    if (top_ok) {
        if (bot_ok) st.push(bot);
        next = top;
        return true;
    }
    next = bot;
    return bot_ok;
}

__device__ __forceinline__ void order2(float& ka, uint32_t& ra, float& kb, uint32_t& rb) {
    const bool sw = kb < ka;
    const float k = sw ? kb : ka;
    const uint32_t r = sw ? rb : ra;
    kb = sw ? ka : kb;
    rb = sw ? ra : rb;
    ka = k;
    ra = r;
}

// Wide-node visit: all children's slabs from one 128 B fetch, accepted
// children sorted by entry distance (a 4-input sorting network; rejected ones
// sort last as +inf); the nearest is returned in `next`, the others are
// pushed farthest first.  Shadow queries keep the order too: it finds
// occluders near the origin first.  Returns false when no child is accepted.
template <bool CN>
__device__ __forceinline__ bool visit_wide(const WNode* nodes, uint32_t ref, v3 o, v3 inv, float limit, LdsStack& st,
                                           Counters& c, uint32_t& next) {
    if (CN) c.inner++;
    const WNode* p = nodes + ref;
    const float4 lx = p->lox, ly = p->loy, lz = p->loz, hx = p->hix, hy = p->hiy, hz = p->hiz;
    const uint4 r = p->ref;
    const uint32_t cnt = p->meta.x;
    const float inf = __builtin_inff();
    // The reference's slab test on each child box, all four computed
    // unconditionally (a branch on the child count costs more than the two
    // extra tests) and then masked.
    auto key = [&](float ax, float ay, float az, float bx, float by, float bz) {
        const float t = slab(o, inv, mk(ax, ay, az), mk(bx, by, bz));
        return (t >= 0.0f && t < limit) ? t : inf;
    };
    float k0 = key(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x);
    float k1 = key(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y);
    float k2 = key(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z);
    float k3 = key(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w);
    k2 = cnt > 2u ? k2 : inf;
    k3 = cnt > 3u ? k3 : inf;
    uint32_t r0 = r.x, r1 = r.y, r2 = r.z, r3 = r.w;
    order2(k0, r0, k1, r1);
    order2(k2, r2, k3, r3);
    order2(k0, r0, k2, r2);
    order2(k1, r1, k3, r3);
    order2(k1, r1, k2, r2);
    if (k3 != inf) st.push(r3);
    if (k2 != inf) st.push(r2);
    if (k1 != inf) st.push(r1);
    next = r0;
    return k0 != inf;
}

// Wide-node visit on the quantised nodes: the four child boxes decoded as
// org + q * 2^e (the builder's arithmetic), then the reference's slab test,
// sort and pushes as visit_wide.  A flagged node is read from its WNode.
__device__ __forceinline__ float qdec(float org, uint32_t word, int i, float s) {
    return qnode_bound(org, (word >> (8 * i)) & 255u, s);
}

template <bool CN>
__device__ __forceinline__ bool visit_qwide(const QNode* qn, const WNode* wn, uint32_t ref, v3 o, v3 inv, float limit,
                                            LdsStack& st, Counters& c, uint32_t& next) {
    const QNode* p = qn + ref;
    const float4 g = p->org;
    const uint32_t meta = f2u(g.w);
    if (meta & kQFullBit) return visit_wide<CN>(wn, ref, o, inv, limit, st, c, next);
    if (CN) c.inner++;
    const uint4 q0 = p->q0, q1 = p->q1, r = p->ref;
    const float inf = __builtin_inff();
    const float sx = ldexpf(1.0f, int(meta & 255u) - 128);
    const float sy = ldexpf(1.0f, int((meta >> 8) & 255u) - 128);
    const float sz = ldexpf(1.0f, int((meta >> 16) & 255u) - 128);
    const uint32_t cnt = (meta >> 24) & 7u;
    // one child at a time: its six bounds decoded, then the slab test
    auto key = [&](int i) {
        const v3 lo = mk(qdec(g.x, q0.x, i, sx), qdec(g.y, q0.y, i, sy), qdec(g.z, q0.z, i, sz));
        const v3 hi = mk(qdec(g.x, q0.w, i, sx), qdec(g.y, q1.x, i, sy), qdec(g.z, q1.y, i, sz));
        const float t = slab(o, inv, lo, hi);
        return (t >= 0.0f && t < limit) ? t : inf;
    };
    float k0 = key(0), k1 = key(1), k2 = key(2), k3 = key(3);
    k2 = cnt > 2u ? k2 : inf;
    k3 = cnt > 3u ? k3 : inf;
    uint32_t r0 = r.x, r1 = r.y, r2 = r.z, r3 = r.w;
    order2(k0, r0, k1, r1);
    order2(k2, r2, k3, r3);
    order2(k0, r0, k2, r2);
    order2(k1, r1, k3, r3);
    order2(k1, r1, k2, r2);
    if (k3 != inf) st.push(r3);
    if (k2 != inf) st.push(r2);
    if (k1 != inf) st.push(r1);
    next = r0;
    return k0 != inf;
}

// Node kinds of the traversal template parameter W.
enum : int { kTravPairs = 0, kTravWide = 1, kTravQuant = 2 };

// One inner visit of either tree: W selects the child-pair, wide or
// quantised wide nodes.
template <int W, bool CN>
__device__ __forceinline__ bool visit_node(const CNode* cn, const WNode* wn, const QNode* qn, uint32_t ref, v3 o,
                                           v3 inv, float limit, LdsStack& st, Counters& c, uint32_t& next) {
    if (W == kTravQuant) return visit_qwide<CN>(qn, wn, ref, o, inv, limit, st, c, next);
    if (W == kTravWide) return visit_wide<CN>(wn, ref, o, inv, limit, st, c, next);
    return visit_inner<CN>(cn, ref, o, inv, limit, st, c, next);
}
template <int W>
__device__ __forceinline__ uint32_t root_ref(const Root& r) { return W ? r.wref : r.ref; }

// BLAS-only closest-hit traversal (list meshes, shader.wgsl:767-771).
template <bool CN, int W>
__device__ void blas_closest(const DevScene& S, uint32_t root_slot, const float* inv_m, uint32_t tri_off, v3 o, v3 d,
                             int idx, int kind, LdsStack& st, HitRec& h, Counters& c) {
    const v3 lo = xform_point(inv_m, o), ld = xform_dir(inv_m, d);
    const v3 linv = splat(1.0f) / ld;
    const Root& R = S.roots[root_slot];
    const float bt = slab(lo, linv, ld3(R.mn), ld3(R.mx));
    if (bt < 0.0f || bt >= h.t) return;
    st.sp = 0;
    uint32_t ref = root_ref<W>(R);
    for (;;) {
        h.count++;
        if (CN) c.pops++;
        bool have = false;
        if (ref & kLeafBit) {
            uint32_t s0, n0;
            leaf_range(ref, S.bvh_raw, s0, n0);
            for (uint32_t k = tri_off + s0; k < tri_off + s0 + n0; ++k) {
                if (CN) c.tris++;
                float u, v;
                const float4* g = (W ? S.tri_geo_w : S.tri_geo) + 3u * k;
                const float4 a = g[0];
                const float t = tri_hit_v(lo, ld, a, g[1], g[2], u, v);
                if (t >= 0.0f && t < h.t) {
                    h.t = t; h.kind = kind; h.idx = idx; h.prim = int(f2u(a.w)); h.bu = u; h.bv = v;
                }
            }
        } else {
            have = visit_node<W, CN>(S.bnodes, S.wbnodes, S.qbnodes, ref, lo, linv, h.t, st, c, ref);
        }
        if (!have) {
            if (st.sp == 0) break;
            ref = st.pop();
        }
    }
}

template <bool CN, int W>
__device__ bool blas_any(const DevScene& S, uint32_t root_slot, const float* inv_m, uint32_t tri_off, v3 o, v3 d,
                         float tdist, LdsStack& st, Counters& c, float* tr = nullptr, int mat = -1) {
    const v3 lo = xform_point(inv_m, o), ld = xform_dir(inv_m, d);
    const v3 linv = splat(1.0f) / ld;
    const Root& R = S.roots[root_slot];
    const float bt = slab(lo, linv, ld3(R.mn), ld3(R.mx));
    if (bt < 0.0f || bt >= tdist) return false;
    st.sp = 0;
    uint32_t ref = root_ref<W>(R);
    for (;;) {
        if (CN) c.pops++;
        bool have = false;
        if (ref & kLeafBit) {
            uint32_t s0, n0;
            leaf_range(ref, S.bvh_raw, s0, n0);
            for (uint32_t k = tri_off + s0; k < tri_off + s0 + n0; ++k) {
                if (CN) c.tris++;
                float u, v;
                const float4* g = (W ? S.tri_geo_w : S.tri_geo) + 3u * k;
                const float t = tri_hit(lo, ld, g, u, v);
                if (t >= 0.0f && t < tdist && (!tr || tri_shadow(S, mat, f2u(g[0].w), u, v, *tr))) return true;
            }
        } else {
            have = visit_node<W, CN>(S.bnodes, S.wbnodes, S.qbnodes, ref, lo, linv, tdist, st, c, ref);
        }
        if (!have) {
            if (st.sp == 0) break;
            ref = st.pop();
        }
    }
    return false;
}

// trace_scene (shader.wgsl:754-774) and trace_scene_shadow (1206-1229) as
// one resumable traversal.  begin_closest / begin_any test the planes (and,
// for closest hits, the TLAS root box); each step() does one unit of
// TLAS/BLAS work (pop a node, visit it, or intersect one TLAS-leaf object)
// and returns false when the trees are done; finish() walks the list meshes.
// A kernel can run other lanes' work between steps; closest_hit runs them
// back to back.  `ref` is the node visited next, or have == false when the
// next work is a TLAS-leaf object or a stack pop.
//
// Any-hit queries (shadow rays) keep their target distance in h.t, so the
// culling tests are shared; an occluder that stops the light sets `blocked`
// and ends the query.  Without translucent albedo textures every occluder
// stops it (evaluate_shadow_attenuation, shader.wgsl:777-792), so the
// transmittance is 0 or 1; with them (kTranslucentBit) a textured surface
// multiplies `trans` by (1 - alpha) and the walk goes on.
template <uint32_t F, bool CN, int W>
struct Trav {
    v3 o, d, inv;     // world-space ray
    v3 lo, ld, linv;  // the ray in the current level's space
    uint32_t ref, cur, end, tri_off;
    int blas_base, obj, target;
    bool have, in_blas, any, blocked;
    float trans;

    // an any-hit triangle of object `obj`: true when it stops the light
    __device__ __forceinline__ bool tri_stops(const DevScene& S, Flags<F> fl, uint32_t prim, float u, float v) {
        return !fl.has(kTranslucentBit) || tri_shadow(S, S.objs[obj].material_idx, prim, u, v, trans);
    }

    __device__ __forceinline__ void start(const DevScene& S, LdsStack& st) {
        inv = splat(1.0f) / d;
        st.sp = 0;
        ref = root_ref<W>(S.roots[0]);
        have = true;
        in_blas = false;
        blas_base = 0;
        cur = end = tri_off = 0;
        obj = -1;
        lo = o;
        ld = d;
        linv = inv;
    }

    // Returns true when the trees must be traversed.
    __device__ __forceinline__ bool begin_closest(const DevScene& S, Flags<F> fl, v3 o_, v3 d_, LdsStack& st,
                                                  HitRec& h) {
        o = o_;
        d = d_;
        any = false;
        blocked = false;
        target = -1;
        h.t = 1e10f;
        h.count = 0;
        h.kind = kMiss;
        h.idx = -1;
        h.prim = -1;
        h.bu = h.bv = 0.0f;
        if (fl.has(PT_HAS_PLANES)) {
            for (uint32_t i = 0; i < S.n_planes; ++i) {
                const pt_plane& p = S.planes[i];
                const float t = plane_t(ld3(p.normal), p.d, o, d);
                if (t > 0.001f && t < h.t) { h.t = t; h.kind = kPlaneHit; h.idx = int(i); h.prim = -1; }
            }
        }
        if (!fl.has(kAnyObject)) return false;
        const Root& R = S.roots[0];
        const float bt = slab(o, splat(1.0f) / d, ld3(R.mn), ld3(R.mx));
        if (bt < 0.0f || bt >= h.t) return false;
        start(S, st);
        return true;
    }

    // Shadow query toward tdist, skipping object `target`.  Returns true
    // when the trees must be traversed (false: decided, see `blocked`).
    __device__ __forceinline__ bool begin_any(const DevScene& S, Flags<F> fl, v3 o_, v3 d_, int target_, float tdist,
                                              LdsStack& st, HitRec& h) {
        o = o_;
        d = d_;
        any = true;
        blocked = false;
        trans = 1.0f;
        target = target_;
        h.t = tdist;
        if (fl.has(PT_HAS_PLANES)) {
            for (uint32_t i = 0; i < S.n_planes; ++i) {
                const pt_plane& p = S.planes[i];
                const v3 n = ld3(p.normal);
                const float denom = dot(n, d);
                if (fabsf(denom) > 1e-6f) {
                    const float t = (p.d - dot(n, o)) / denom;
                    if (t > 0.001f && t < tdist) {
                        if (fl.has(kTranslucentBit)) {  // trace_plane_shadow's frame (shader.wgsl:1186-1203)
                            const v3 hp = o + d * t;
                            v3 a = mk(1.0f, 0.0f, 0.0f);
                            if (fabsf(n.x) > 0.9999f) a = mk(0.0f, 0.0f, 1.0f);
                            const v3 T = normalize(cross(n, a));
                            const v3 B = normalize(cross(n, T));
                            if (!shadow_surface(S, p.material_idx, v2{dot(hp, T), dot(hp, B)}, trans)) continue;
                        }
                        blocked = true;
                        return false;
                    }
                }
            }
        }
        if (!fl.has(kAnyObject)) return false;
        start(S, st);  // no root box test for shadow rays (shader.wgsl:1134-1144)
        return true;
    }

    __device__ __forceinline__ bool step(const DevScene& S, Flags<F> fl, LdsStack& st, HitRec& h, Counters& c) {
        if (!have) {
            if (in_blas) {
                if (st.sp == blas_base) {
                    in_blas = false;
                    lo = o;
                    linv = inv;
                    return true;
                }
                ref = st.pop();
            } else if (cur < end) {
                const uint32_t i = cur++;
                if (any && int(i) == target) return true;
                if (CN) c.objs++;
                const pt_object& ob = S.objs[i];
                const int type = ob.object_type;
                if (fl.has(PT_HAS_MESHES) && type == PT_OBJ_MESH) {
                    const v3 mo = xform_point(ob.inv_matrix, o), md = xform_dir(ob.inv_matrix, d);
                    const v3 minv = splat(1.0f) / md;
                    const Root& br = S.roots[S.obj_root[i]];
                    const float b2 = slab(mo, minv, ld3(br.mn), ld3(br.mx));
                    if (!(b2 < 0.0f || b2 >= h.t)) {
                        lo = mo;
                        ld = md;
                        linv = minv;
                        tri_off = f2u(ob.param1);
                        obj = int(i);
                        blas_base = st.sp;
                        in_blas = true;
                        ref = root_ref<W>(br);
                        have = true;
                    }
                } else if (any && fl.has(kTranslucentBit)) {
                    const int mt = ob.material_idx;
                    bool b = false;
                    if (fl.has(PT_HAS_SPHERES) && type == PT_OBJ_SPHERE)
                        b = sphere_shadow(S, mt, xform_point(ob.inv_matrix, o), xform_dir(ob.inv_matrix, d), h.t, trans);
                    else if (fl.has(PT_HAS_CUBES) && type == PT_OBJ_CUBE)
                        b = cube_shadow(S, mt, xform_point(ob.inv_matrix, o), xform_dir(ob.inv_matrix, d), h.t, trans);
                    else if (fl.has(PT_HAS_CYLINDERS) && type == PT_OBJ_CYLINDER)
                        b = cylinder_shadow(S, mt, xform_point(ob.inv_matrix, o), xform_dir(ob.inv_matrix, d),
                                            ob.param0, h.t, trans);
                    else if (fl.has(PT_HAS_TORI) && type == PT_OBJ_TORUS)
                        b = torus_shadow<true>(o, d, ob.inv_matrix, ob.param0, h.t, &S, mt, &trans);
                    if (b) {
                        blocked = true;
                        return false;
                    }
                } else if (any) {
                    bool b = false;
                    if (fl.has(PT_HAS_SPHERES) && type == PT_OBJ_SPHERE)
                        b = sphere_blocks(xform_point(ob.inv_matrix, o), xform_dir(ob.inv_matrix, d), h.t);
                    else if (fl.has(PT_HAS_CUBES) && type == PT_OBJ_CUBE)
                        b = cube_blocks(xform_point(ob.inv_matrix, o), xform_dir(ob.inv_matrix, d), h.t);
                    else if (fl.has(PT_HAS_CYLINDERS) && type == PT_OBJ_CYLINDER)
                        b = cylinder_blocks(xform_point(ob.inv_matrix, o), xform_dir(ob.inv_matrix, d), ob.param0, h.t);
                    else if (fl.has(PT_HAS_TORI) && type == PT_OBJ_TORUS)
                        b = torus_blocks(o, d, ob.inv_matrix, ob.param0, h.t);
                    if (b) {
                        blocked = true;
                        return false;
                    }
                } else if (fl.has(PT_HAS_SPHERES) && type == PT_OBJ_SPHERE) {
                    const float t = unit_sphere(xform_point(ob.inv_matrix, o), xform_dir(ob.inv_matrix, d));
                    if (t > 0.001f && t < h.t) { h.t = t; h.kind = kObjectHit; h.idx = int(i); h.prim = -1; }
                } else if (fl.has(PT_HAS_CUBES) && type == PT_OBJ_CUBE) {
                    const float t = unit_cube(xform_point(ob.inv_matrix, o), xform_dir(ob.inv_matrix, d));
                    if (t > 0.001f && t < h.t) { h.t = t; h.kind = kObjectHit; h.idx = int(i); h.prim = -1; }
                } else if (fl.has(PT_HAS_CYLINDERS) && type == PT_OBJ_CYLINDER) {
                    int which;
                    v3 ps;
                    const float t = cone_cylinder(xform_point(ob.inv_matrix, o), xform_dir(ob.inv_matrix, d), 1.0f,
                                                  ob.param0, which, ps);
                    if (which >= 0 && t < h.t) { h.t = t; h.kind = kObjectHit; h.idx = int(i); h.prim = -1; }
                } else if (fl.has(PT_HAS_TORI) && type == PT_OBJ_TORUS) {
                    v3 tl = xform_dir(ob.inv_matrix, d);
                    const float rs = len(tl);
                    tl = tl / rs;
                    const float t = torus_t(xform_point(ob.inv_matrix, o), tl, 1.0f, ob.param0 * 1.0f);
                    if (t > 0.0f) {
                        const float wt = t / rs;
                        if (wt < h.t) { h.t = wt; h.kind = kObjectHit; h.idx = int(i); h.prim = -1; }
                    }
                }
                return true;
            } else {
                if (st.sp == 0) return false;
                ref = st.pop();
            }
        }
        // visit `ref` (the reference's pop)
        if (!any) h.count++;
        if (CN) c.pops++;
        have = false;
        if (ref & kLeafBit) {
            uint32_t s0, n0;
            if (in_blas) {
                leaf_range(ref, S.bvh_raw, s0, n0);
                const uint32_t kb = tri_off + s0, ke = kb + n0;
                for (uint32_t k = kb; k < ke; k += 2u) {
                    // both triangles' vertices are requested before either test
                    const bool two = k + 1u < ke;
                    const float4* g0 = (W ? S.tri_geo_w : S.tri_geo) + 3u * k;
                    const float4* g1 = two ? g0 + 3 : g0;
                    const float4 a0 = g0[0], b0 = g0[1], c0 = g0[2];
                    const float4 a1 = g1[0], b1 = g1[1], c1 = g1[2];
                    if (CN) c.tris++;
                    float u, v;
                    float t = tri_hit_v(lo, ld, a0, b0, c0, u, v);
                    if (t >= 0.0f && t < h.t) {
                        if (any) {
                            if (tri_stops(S, fl, f2u(a0.w), u, v)) {
                                blocked = true;
                                return false;
                            }
                        } else {
                            h.t = t; h.kind = kObjectHit; h.idx = obj; h.prim = int(f2u(a0.w)); h.bu = u; h.bv = v;
                        }
                    }
                    if (two) {
                        if (CN) c.tris++;
                        t = tri_hit_v(lo, ld, a1, b1, c1, u, v);
                        if (t >= 0.0f && t < h.t) {
                            if (any) {
                                if (tri_stops(S, fl, f2u(a1.w), u, v)) {
                                    blocked = true;
                                    return false;
                                }
                            } else {
                                h.t = t; h.kind = kObjectHit; h.idx = obj; h.prim = int(f2u(a1.w)); h.bu = u;
                                h.bv = v;
                            }
                        }
                    }
                }
            } else {
                leaf_range(ref, S.tlas_raw, s0, n0);
                cur = s0;
                end = s0 + n0;
            }
        } else {
            // one call site for both levels: (lo, linv) is the ray in the current level's space
            have = visit_node<W, CN>(in_blas ? S.bnodes : S.tnodes, in_blas ? S.wbnodes : S.wtnodes,
                                     in_blas ? S.qbnodes : S.qtnodes, ref, lo, linv, h.t, st, c, ref);
        }
        return true;
    }

    __device__ __forceinline__ void finish(const DevScene& S, Flags<F> fl, LdsStack& st, HitRec& h, Counters& c) {
        if (fl.has(PT_HAS_LIST_MESHES)) {
            for (uint32_t m = 0; m < S.n_meshes; ++m) {
                const pt_mesh_instance& mi = S.meshes[m];
                if (any) {
                    if (blocked) break;
                    blocked = blas_any<CN, W>(S, S.mesh_root[m], mi.inv_matrix, mi.tri_offset, o, d, h.t, st, c,
                                              fl.has(kTranslucentBit) ? &trans : nullptr, mi.material_idx);
                } else {
                    blas_closest<CN, W>(S, S.mesh_root[m], mi.inv_matrix, mi.tri_offset, o, d, int(m), kListMeshHit,
                                        st, h, c);
                }
            }
        }
    }
};

template <uint32_t F, bool CN, int W>
__device__ void closest_hit(const DevScene& S, Flags<F> fl, v3 o, v3 d, LdsStack& st, HitRec& h, Counters& c) {
    Trav<F, CN, W> q;
    if (q.begin_closest(S, fl, o, d, st, h))
        while (q.step(S, fl, st, h, c)) {
        }
    q.finish(S, fl, st, h, c);
}

// trace_scene_shadow (shader.wgsl:1206-1229): the light that reaches tdist
// past the occluders (object `target` skipped): 0 when one stops it, else the
// product of the translucent surfaces' (1 - alpha), 1 without any.
template <uint32_t F, bool CN, int W>
__device__ float transmittance(const DevScene& S, Flags<F> fl, v3 o, v3 d, int target, float tdist, LdsStack& st,
                               Counters& c) {
    Trav<F, CN, W> q;
    HitRec h;
    if (q.begin_any(S, fl, o, d, target, tdist, st, h))
        while (q.step(S, fl, st, h, c)) {
        }
    q.finish(S, fl, st, h, c);
    return q.blocked ? 0.0f : q.trans;
}

// trace_scene_shadow (shader.wgsl:1206-1229): true when an occluder lies
// before tdist (object `target` skipped).
template <uint32_t F, bool CN, int W>
__device__ bool occluded(const DevScene& S, Flags<F> fl, v3 o, v3 d, int target, float tdist, LdsStack& st,
                         Counters& c) {
    Trav<F, CN, W> q;
    HitRec h;
    if (q.begin_any(S, fl, o, d, target, tdist, st, h))
        while (q.step(S, fl, st, h, c)) {
        }
    q.finish(S, fl, st, h, c);
    return q.blocked;
}

// Attributes of the winning hit (the values the reference leaves in its
// SurfaceHit after the last accepted candidate).
struct SurfHit {
    float t;
    int m_idx;
    int o_idx;
    v3 n;
    v2 uv;
    v3 tg, bt;  // tangent frame of sphere, cube and plane hits (with uv only)
};

template <bool WANT_UV>
__device__ SurfHit resolve_hit(const DevScene& S, v3 o, v3 d, const HitRec& h) {
    SurfHit r;
    r.t = h.t;
    r.m_idx = -1;
    r.o_idx = -1;
    r.n = splat(0.0f);
    r.uv = v2{0.0f, 0.0f};
    r.tg = splat(0.0f);
    r.bt = splat(0.0f);
    if (h.kind == kMiss) return r;
    if (h.kind == kPlaneHit) {  // trace_plane, shader.wgsl:622-636
        const pt_plane& p = S.planes[h.idx];
        r.m_idx = p.material_idx;
        r.n = ld3(p.normal);
        if (WANT_UV) {
            const v3 hp = o + d * h.t;
            v3 tg = mk(1.0f, 0.0f, 0.0f);
            if (fabsf(r.n.x) > 0.9f) tg = mk(0.0f, 0.0f, 1.0f);
            const v3 T = normalize(cross(r.n, tg));
            const v3 B = normalize(cross(r.n, T));
            r.uv = v2{dot(hp, T), dot(hp, B)};
            r.tg = T;
            r.bt = B;
        }
        return r;
    }
    const float* inv;
    if (h.kind == kListMeshHit) {
        inv = S.meshes[h.idx].inv_matrix;
        r.m_idx = S.meshes[h.idx].material_idx;
        r.o_idx = -1;
    } else {
        const pt_object& ob = S.objs[h.idx];
        inv = ob.inv_matrix;
        r.m_idx = ob.material_idx;
        r.o_idx = h.idx;
    }
    if (h.prim >= 0) {  // trace_mesh, shader.wgsl:509-523
        const pt_triangle& tr = S.tris[h.prim];
        const float u = h.bu, v = h.bv;
        const float b0 = 1.0f - u - v;
        const v3 ln = normalize(ld3(tr.n0) * b0 + ld3(tr.n1) * u + ld3(tr.n2) * v);
        r.n = normalize(xform_tdir(inv, ln));
        if (WANT_UV) r.uv = v2{tr.uv0[0] * b0 + tr.uv1[0] * u + tr.uv2[0] * v, tr.uv0[1] * b0 + tr.uv1[1] * u + tr.uv2[1] * v};
        return r;
    }
    const pt_object& ob = S.objs[h.idx];
    const v3 lo = xform_point(inv, o), ld = xform_dir(inv, d);
    const float t = h.t;
    switch (ob.object_type) {
        case PT_OBJ_SPHERE: {  // shader.wgsl:552-572
            const v3 ln = lo + ld * t;
            if (WANT_UV) {
                const float phi = atan2f(ln.z, ln.x);
                const float th = asinf(clampf(ln.y, -1.0f, 1.0f));
                r.uv = v2{0.5f + phi / kTWO_PI, 0.5f + th / kPI};
                v3 lt = mk(-ln.z, 0.0f, ln.x);
                if (fabsf(ln.y) > 0.9999f) lt = mk(1.0f, 0.0f, 0.0f);
                lt = normalize(lt);
                r.tg = normalize(xform_normal(inv, lt));
                r.bt = normalize(xform_normal(inv, cross(ln, lt)));
            }
            r.n = normalize(xform_normal(inv, ln));
            break;
        }
        case PT_OBJ_CUBE: {  // shader.wgsl:585-618
            const v3 lh = lo + ld * t;
            const v3 a = mk(fabsf(lh.x), fabsf(lh.y), fabsf(lh.z));
            const float m = hmax(a);
            v3 ln, lt, lb;
            if (m == a.x) {
                const float s = sgn(lh.x);
                ln = mk(s, 0.0f, 0.0f);
                lt = mk(0.0f, 0.0f, -s);
                lb = mk(0.0f, 1.0f, 0.0f);
                r.uv = v2{(-s * lh.z) * 0.5f + 0.5f, lh.y * 0.5f + 0.5f};
            } else if (m == a.y) {
                const float s = sgn(lh.y);
                ln = mk(0.0f, s, 0.0f);
                lt = mk(1.0f, 0.0f, 0.0f);
                lb = mk(0.0f, 0.0f, s);
                r.uv = v2{lh.x * 0.5f + 0.5f, (s * lh.z) * 0.5f + 0.5f};
            } else {
                const float s = sgn(lh.z);
                ln = mk(0.0f, 0.0f, s);
                lt = mk(s, 0.0f, 0.0f);
                lb = mk(0.0f, 1.0f, 0.0f);
                r.uv = v2{(s * lh.x) * 0.5f + 0.5f, lh.y * 0.5f + 0.5f};
            }
            r.n = normalize(xform_normal(inv, ln));
            if (WANT_UV) {
                r.tg = normalize(xform_normal(inv, lt));
                r.bt = normalize(xform_normal(inv, lb));
            }
            break;
        }
        case PT_OBJ_CYLINDER: {  // hit_cylinder normals/uvs, shader.wgsl:345-368
            int which;
            v3 ps;
            cone_cylinder(lo, ld, 1.0f, ob.param0, which, ps);
            const float r0 = 1.0f, r1 = ob.param0, k = (r1 - r0) / 1.0f;
            v3 ln;
            const v3 p = lo + t * ld;
            if (which == 0) {
                const float y = lo.y + t * ld.y;
                const float slant = -k * (r0 + k * y);
                ln = normalize(mk(p.x, slant, p.z));
                if (WANT_UV) r.uv = v2{atan2f(p.z, p.x) / kTWO_PI + 0.5f, y};
            } else {
                ln = mk(0.0f, float(which - 1) * 2.0f - 1.0f, 0.0f);
                const float m2 = mx(r0, r1) * 2.0f;
                r.uv = v2{p.x / m2 + 0.5f, p.z / m2 + 0.5f};
            }
            r.n = normalize(xform_normal(inv, ln));
            break;
        }
        case PT_OBJ_TORUS: {  // shader.wgsl:666-690 (normal uses mat3, not its transpose)
            const float rs = len(ld);
            const v3 dn = ld / rs;
            const float tl = torus_t(lo, dn, 1.0f, ob.param0 * 1.0f);
            const v3 p = lo + dn * tl;
            const v3 ring = normalize(mk(p.x, 0.0f, p.z));
            const v3 ln = normalize(p - ring);
            r.n = normalize(xform_mat3(inv, ln));
            if (WANT_UV) {
                r.uv = v2{(atan2f(p.z, p.x) / kTWO_PI) + 0.5f,
                          (atan2f(p.y, sqrtf(p.x * p.x + p.z * p.z) - 1.0f) / kTWO_PI) + 0.5f};
            }
            break;
        }
        default: break;
    }
    return r;
}

// The hit as shading needs it: texture coordinates only where a material
// may use a texture (never in the kernels specialised for the configuration
// flag sets).
template <uint32_t F>
__device__ __forceinline__ SurfHit resolve_for_shading(const DevScene& S, Flags<F> fl, v3 o, v3 d, const HitRec& h) {
    if constexpr (F != kDynamicFlags && !(F & kTexturesBit)) {
        return resolve_hit<false>(S, o, d, h);
    } else {
        return fl.has(kTexturesBit) ? resolve_hit<true>(S, o, d, h) : resolve_hit<false>(S, o, d, h);
    }
}

}  // namespace pt
