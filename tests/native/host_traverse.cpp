// host_traverse.cpp — the kernels' traversal code (dev/traverse.h) compiled
// for the CPU, over the scene layout pt_layout.cpp builds for the device.
// Test-only: lets the CPU test suite compare the child-pair, the wide and the
// quantised wide traversal with the oracle, and check the wide stack bound,
// without a GPU.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "pt_layout.h"
#include "dev/traverse.h"

using namespace pt;

namespace {

struct HostScene {
    HostLayout L;
    std::vector<float4> geo;
    std::vector<QNode> qtn, qbn;  // the wide nodes quantised as the device builder does
    DevScene S{};
};

bool make_scene(const pt_scene_desc* d, HostScene& h, std::string& err) {
    if (build_layout(d, h.L, err)) return false;
    build_tri_geo(d, h.geo);
    DevScene& S = h.S;
    S.mats = static_cast<const pt_material*>(d->materials.ptr);
    S.objs = static_cast<const pt_object*>(d->objects.ptr);
    S.obj_root = h.L.obj_root.data();
    S.tnodes = h.L.tn.data();
    S.bnodes = h.L.bn.data();
    S.wtnodes = h.L.wtn.data();
    S.wbnodes = h.L.wbn.data();
    for (const WNode& w : h.L.wtn) h.qtn.push_back(quantize_node(w));
    for (const WNode& w : h.L.wbn) h.qbn.push_back(quantize_node(w));
    S.qtnodes = h.qtn.data();
    S.qbnodes = h.qbn.data();
    S.roots = h.L.roots.data();
    S.tlas_raw = static_cast<const pt_bvh_node*>(d->tlas.ptr);
    S.bvh_raw = static_cast<const pt_bvh_node*>(d->bvh.ptr);
    S.tri_geo = h.geo.data();
    S.tri_geo_w = h.geo.data();
    S.tris = static_cast<const pt_triangle*>(d->triangles.ptr);
    S.planes = static_cast<const pt_plane*>(d->planes.ptr);
    S.meshes = static_cast<const pt_mesh_instance*>(d->meshes.ptr);
    S.mesh_root = h.L.mesh_root.data();
    S.lights = static_cast<const pt_light*>(d->lights.ptr);
    S.n_planes = uint32_t(d->planes.count);
    S.n_meshes = uint32_t(d->meshes.count);
    S.n_lights = uint32_t(d->lights.count);
    S.n_objs = uint32_t(d->objects.count);
    S.flags = d->flags;
    S.bounce_limit = d->bounce_limit;
    S.stack_cap = stack_need(h.L.tlas_depth, h.L.blas_depth);
    S.wstack_cap = stack_need(h.L.tlas_wneed, h.L.blas_wneed);
    return true;
}

}  // namespace

extern "C" {

// Layout summary: [n_cnodes_tlas, n_cnodes_blas, n_wnodes_tlas, n_wnodes_blas,
// stack_cap, wstack_cap, tlas_depth, blas_depth].
int ht_layout_info(const pt_scene_desc* d, uint32_t* info) {
    HostScene h;
    std::string err;
    if (!make_scene(d, h, err)) return -1;
    info[0] = uint32_t(h.L.tn.size());
    info[1] = uint32_t(h.L.bn.size());
    info[2] = uint32_t(h.L.wtn.size());
    info[3] = uint32_t(h.L.wbn.size());
    info[4] = h.S.stack_cap;
    info[5] = h.S.wstack_cap;
    info[6] = h.L.tlas_depth;
    info[7] = h.L.blas_depth;
    return 0;
}

// Copies the wide nodes (TLAS then BLAS) into out (128 B each) when out is
// non-null; returns their number.
long ht_wide_nodes(const pt_scene_desc* d, void* out_tlas, void* out_blas) {
    HostScene h;
    std::string err;
    if (!make_scene(d, h, err)) return -1;
    if (out_tlas) std::memcpy(out_tlas, h.L.wtn.data(), h.L.wtn.size() * sizeof(WNode));
    if (out_blas) std::memcpy(out_blas, h.L.wbn.data(), h.L.wbn.size() * sizeof(WNode));
    return long(h.L.wtn.size() + h.L.wbn.size());
}

// The quantised nodes (TLAS then BLAS, 64 B each), as ht_wide_nodes.
long ht_quant_nodes(const pt_scene_desc* d, void* out_tlas, void* out_blas) {
    HostScene h;
    std::string err;
    if (!make_scene(d, h, err)) return -1;
    if (out_tlas) std::memcpy(out_tlas, h.qtn.data(), h.qtn.size() * sizeof(QNode));
    if (out_blas) std::memcpy(out_blas, h.qbn.data(), h.qbn.size() * sizeof(QNode));
    return long(h.qtn.size() + h.qbn.size());
}

// One node quantised: a WNode (128 B) in, a QNode (64 B) out.
void ht_quantize_node(const void* wnode, void* qnode) {
    WNode w;
    std::memcpy(&w, wnode, sizeof(w));
    const QNode q = quantize_node(w);
    std::memcpy(qnode, &q, sizeof(q));
}

// Closest hits of n rays (6 floats each).  wide selects the traversal (0
// child pairs, 1 wide, 2 quantised wide); stack_cap 0 uses the layout's
// bound, otherwise the given capacity.
int ht_trace_rays(const pt_scene_desc* d, const float* rays, uint32_t n, int wide, uint32_t stack_cap, pt_hit* out) {
    HostScene h;
    std::string err;
    if (!make_scene(d, h, err)) return -1;
    const uint32_t cap = stack_cap ? stack_cap : (wide ? h.S.wstack_cap : h.S.stack_cap);
    std::vector<uint32_t> stack(cap + 1);
    const Flags<kDynamicFlags> fl{h.S.flags};
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = rays + size_t(i) * 6;
        const v3 o = mk(r[0], r[1], r[2]), dd = mk(r[3], r[4], r[5]);
        LdsStack st = lds_stack_of(stack.data(), 1, cap);
        Counters c;
        c.clear();
        HitRec hr;
        if (wide == kTravQuant)
            closest_hit<kDynamicFlags, false, kTravQuant>(h.S, fl, o, dd, st, hr, c);
        else if (wide)
            closest_hit<kDynamicFlags, false, kTravWide>(h.S, fl, o, dd, st, hr, c);
        else
            closest_hit<kDynamicFlags, false, kTravPairs>(h.S, fl, o, dd, st, hr, c);
        const SurfHit s = resolve_hit<false>(h.S, o, dd, hr);
        pt_hit& x = out[i];
        std::memset(&x, 0, sizeof(x));
        x.t = hr.t;
        x.o_idx = s.o_idx;
        x.m_idx = s.m_idx;
        x.prim_id = hr.prim;
        x.plane_idx = hr.kind == kPlaneHit ? hr.idx : -1;
        x.count = hr.count;
        x.n[0] = s.n.x; x.n[1] = s.n.y; x.n[2] = s.n.z;
    }
    return 0;
}

// Shadow queries: occ[i] = 1 when the ray is blocked before target_dist.
int ht_trace_shadow_rays(const pt_scene_desc* d, const float* rays, const int32_t* target, const float* dist,
                         uint32_t n, int wide, uint8_t* occ) {
    HostScene h;
    std::string err;
    if (!make_scene(d, h, err)) return -1;
    const uint32_t cap = wide ? h.S.wstack_cap : h.S.stack_cap;
    std::vector<uint32_t> stack(cap + 1);
    const Flags<kDynamicFlags> fl{h.S.flags};
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = rays + size_t(i) * 6;
        const v3 o = mk(r[0], r[1], r[2]), dd = mk(r[3], r[4], r[5]);
        LdsStack st = lds_stack_of(stack.data(), 1, cap);
        Counters c;
        c.clear();
        const bool b = wide == kTravQuant
                           ? occluded<kDynamicFlags, false, kTravQuant>(h.S, fl, o, dd, target[i], dist[i], st, c)
                       : wide ? occluded<kDynamicFlags, false, kTravWide>(h.S, fl, o, dd, target[i], dist[i], st, c)
                              : occluded<kDynamicFlags, false, kTravPairs>(h.S, fl, o, dd, target[i], dist[i], st, c);
        occ[i] = b ? 1 : 0;
    }
    return 0;
}

// Transmittance of shadow queries (0 when blocked), with the scene's
// textures: albedo alphas below 1 make the queries attenuate.
int ht_shadow_transmittance(const pt_scene_desc* d, const float* rays, const int32_t* target, const float* dist,
                            uint32_t n, int wide, float* out) {
    HostScene h;
    std::string err;
    if (!make_scene(d, h, err)) return -1;
    h.S.tex = reinterpret_cast<const uint32_t*>(d->textures);
    h.S.tex_layers = d->texture_layers;
    h.S.flags |= kTexturesBit | kTranslucentBit;
    const uint32_t cap = wide ? h.S.wstack_cap : h.S.stack_cap;
    std::vector<uint32_t> stack(cap + 1);
    const Flags<kDynamicFlags> fl{h.S.flags};
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = rays + size_t(i) * 6;
        const v3 o = mk(r[0], r[1], r[2]), dd = mk(r[3], r[4], r[5]);
        LdsStack st = lds_stack_of(stack.data(), 1, cap);
        Counters c;
        c.clear();
        out[i] = wide == kTravQuant
                     ? transmittance<kDynamicFlags, false, kTravQuant>(h.S, fl, o, dd, target[i], dist[i], st, c)
                 : wide ? transmittance<kDynamicFlags, false, kTravWide>(h.S, fl, o, dd, target[i], dist[i], st, c)
                        : transmittance<kDynamicFlags, false, kTravPairs>(h.S, fl, o, dd, target[i], dist[i], st, c);
    }
    return 0;
}

}  // extern "C"
