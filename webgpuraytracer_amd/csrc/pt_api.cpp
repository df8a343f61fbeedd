// pt_api.cpp — the C ABI of libpt (declared in include/pt_abi.h).
//
// Host-side responsibilities: validate the tracer.js arrays, build the device
// layout (child-pair BVH nodes, root table, 48 B triangle records), own every
// device buffer, launch the kernels on the context's stream and report errors
// as status codes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "pt_abi.h"
#include "pt_internal.h"
#include "pt_ctx.h"
#include "pt_layout.h"

using namespace pt;

namespace {

constexpr uint32_t kMaxStack = 128;  // combined TLAS+BLAS entries per lane

}  // namespace

namespace {

bool check_array(const pt_array& a, uint32_t stride, bool allow_empty) {
    if (a.stride != stride) return false;
    if (a.count == 0) return allow_empty;
    return a.ptr != nullptr;
}

int validate_desc(pt_ctx* c, const pt_scene_desc* d) {
    if (!d) return c->fail(PT_ERR_ARG, "scene descriptor is NULL");
    struct { const pt_array* a; uint32_t stride; const char* name; } arrs[] = {
        {&d->materials, sizeof(pt_material), "materials"}, {&d->objects, sizeof(pt_object), "objects"},
        {&d->tlas, sizeof(pt_bvh_node), "tlas"},          {&d->meshes, sizeof(pt_mesh_instance), "meshes"},
        {&d->bvh, sizeof(pt_bvh_node), "bvh"},            {&d->triangles, sizeof(pt_triangle), "triangles"},
        {&d->planes, sizeof(pt_plane), "planes"},         {&d->lights, sizeof(pt_light), "lights"}};
    for (auto& x : arrs)  // every array holds at least one record, as WebGPU buffers do; with the
                          // device-built TLAS the descriptor's TLAS is not read and may be empty
        if (!check_array(*x.a, x.stride, x.a == &d->tlas && c->tlas_build == PT_TLAS_BUILD_DEVICE))
            return c->fail(PT_ERR_LAYOUT, "%s: expected a non-empty array with stride %u (got count %llu stride %u)",
                           x.name, x.stride, (unsigned long long)x.a->count, x.a->stride);
    if (d->texture_layers > 0 && !d->textures) return c->fail(PT_ERR_ARG, "texture_layers > 0 without texels");
    if (d->texture_layers > 4096) return c->fail(PT_ERR_ARG, "at most 4096 texture layers");
    const pt_material* m = static_cast<const pt_material*>(d->materials.ptr);
    const pt_object* objs = static_cast<const pt_object*>(d->objects.ptr);
    const pt_mesh_instance* mi = static_cast<const pt_mesh_instance*>(d->meshes.ptr);
    for (uint64_t i = 0; i < d->materials.count; ++i) {
        // normal and height maps (and parallax mapping) need the hit's tangent frame, which the reference
        // sets only on sphere, cube and plane hits (shader.wgsl:560-572, 598-617, 630-634): other hits
        // keep the frame of an earlier candidate, so materials with these maps may be used by spheres,
        // cubes and planes only
        if (m[i].normal_idx >= 0 || m[i].height_idx >= 0) {
            for (uint64_t k = 0; k < d->objects.count; ++k)
                if (objs[k].material_idx == int32_t(i) && objs[k].object_type != PT_OBJ_SPHERE &&
                    objs[k].object_type != PT_OBJ_CUBE)
                    return c->fail(PT_ERR_UNSUPPORTED,
                                   "material %llu: normal and height maps on object %llu, which is not a sphere, "
                                   "cube or plane", (unsigned long long)i, (unsigned long long)k);
            if (d->flags & PT_HAS_LIST_MESHES)
                for (uint64_t k = 0; k < d->meshes.count; ++k)
                    if (mi[k].material_idx == int32_t(i))
                        return c->fail(PT_ERR_UNSUPPORTED, "material %llu: normal and height maps on mesh %llu",
                                       (unsigned long long)i, (unsigned long long)k);
            if (m[i].height_idx >= 0 && !(m[i].height_params[2] > 0.0f && std::isfinite(m[i].height_params[2])))
                return c->fail(PT_ERR_ARG, "material %llu: height map with %g layers", (unsigned long long)i,
                               double(m[i].height_params[2]));
        }
        const int32_t idx[6] = {m[i].albedo_idx,   m[i].roughness_idx, m[i].metallic_idx,
                                m[i].emissive_idx, m[i].normal_idx,    m[i].height_idx};
        for (int32_t t : idx)
            if (t >= 0 && uint32_t(t) >= d->texture_layers)
                return c->fail(PT_ERR_LAYOUT, "material %llu: texture %d outside the %u layers",
                               (unsigned long long)i, t, d->texture_layers);
    }
    if (d->bounce_limit < 0) return c->fail(PT_ERR_ARG, "bounce_limit < 0");
    if (!d->sky_rgba16f || d->sky_width == 0 || d->sky_height == 0) return c->fail(PT_ERR_ARG, "sky texels missing");
    if (d->sky_cdf) {
        const uint64_t need = uint64_t(d->sky_height) + 1 + uint64_t(d->sky_height) * (d->sky_width + 1);
        if (d->sky_cdf_count != need)
            return c->fail(PT_ERR_LAYOUT, "sky_cdf has %llu floats, expected %llu", (unsigned long long)d->sky_cdf_count,
                           (unsigned long long)need);
    }
    // index sanity for everything the kernels dereference
    const uint64_t nm = d->materials.count, no = d->objects.count;
    const pt_object* ob = static_cast<const pt_object*>(d->objects.ptr);
    const uint32_t any_obj = PT_HAS_SPHERES | PT_HAS_CUBES | PT_HAS_CYLINDERS | PT_HAS_TORI | PT_HAS_MESHES;
    if (d->flags & any_obj) {
        for (uint64_t i = 0; i < no; ++i) {
            if (ob[i].material_idx < 0 || uint64_t(ob[i].material_idx) >= nm)
                return c->fail(PT_ERR_LAYOUT, "object %llu: material index out of range", (unsigned long long)i);
            if (ob[i].light_idx >= 0 && uint64_t(ob[i].light_idx) >= d->lights.count)
                return c->fail(PT_ERR_LAYOUT, "object %llu: light index out of range", (unsigned long long)i);
        }
    }
    if (d->flags & PT_HAS_PLANES) {
        const pt_plane* p = static_cast<const pt_plane*>(d->planes.ptr);
        for (uint64_t i = 0; i < d->planes.count; ++i)
            if (p[i].material_idx < 0 || uint64_t(p[i].material_idx) >= nm)
                return c->fail(PT_ERR_LAYOUT, "plane %llu: material index out of range", (unsigned long long)i);
    }
    if (d->flags & PT_HAS_LIST_MESHES) {
        const pt_mesh_instance* mi = static_cast<const pt_mesh_instance*>(d->meshes.ptr);
        for (uint64_t i = 0; i < d->meshes.count; ++i)
            if (mi[i].material_idx < 0 || uint64_t(mi[i].material_idx) >= nm)
                return c->fail(PT_ERR_LAYOUT, "mesh %llu: material index out of range", (unsigned long long)i);
    }
    if (d->flags & PT_HAS_LIGHTS) {
        const pt_light* L = static_cast<const pt_light*>(d->lights.ptr);
        for (uint64_t i = 0; i < d->lights.count; ++i)
            if (L[i].obj_idx >= 0 && uint64_t(L[i].obj_idx) >= no)
                return c->fail(PT_ERR_LAYOUT, "light %llu: object index out of range", (unsigned long long)i);
    }
    return PT_OK;
}

// The quantised copies of both wide trees (PT_TRAVERSAL_WIDE_QUANTIZED), made
// on the device after every upload, build and update.
int quantize_trees(pt_ctx* c) {
    int rc;
    if ((rc = c->alloc(c->qtnodes, std::max<uint64_t>(c->n_wtn, 1) * sizeof(QNode), "quantised tlas nodes")) ||
        (rc = c->alloc(c->qbnodes, std::max<uint64_t>(c->n_wbn, 1) * sizeof(QNode), "quantised bvh nodes")))
        return rc;
    hipError_t e = quantize_wide_nodes(static_cast<const WNode*>(c->wtnodes.p), static_cast<QNode*>(c->qtnodes.p),
                                       uint32_t(c->n_wtn), c->stream);
    if (e == hipSuccess)
        e = quantize_wide_nodes(static_cast<const WNode*>(c->wbnodes.p), static_cast<QNode*>(c->qbnodes.p),
                                uint32_t(c->n_wbn), c->stream);
    return e == hipSuccess ? PT_OK : c->hip_fail(e, "node quantisation");
}

void refresh_dev(pt_ctx* c) {
    DevScene& S = c->dev;
    S.mats = static_cast<const pt_material*>(c->mats.p);
    S.objs = static_cast<const pt_object*>(c->objs.p);
    S.obj_root = static_cast<const uint32_t*>(c->obj_root.p);
    S.tnodes = static_cast<const CNode*>(c->tnodes.p);
    S.bnodes = static_cast<const CNode*>(c->bnodes.p);
    S.wtnodes = static_cast<const WNode*>(c->wtnodes.p);
    S.wbnodes = static_cast<const WNode*>(c->wbnodes.p);
    S.qtnodes = static_cast<const QNode*>(c->qtnodes.p);
    S.qbnodes = static_cast<const QNode*>(c->qbnodes.p);
    S.roots = static_cast<const Root*>(c->roots.p);
    S.tlas_raw = static_cast<const pt_bvh_node*>(c->tlas_raw.p);
    S.bvh_raw = static_cast<const pt_bvh_node*>(c->bvh_raw.p);
    S.tri_geo = static_cast<const float4*>(c->tri_geo.p);
    S.tri_geo_w = static_cast<const float4*>(c->tri_geo_w.p ? c->tri_geo_w.p : c->tri_geo.p);
    S.tris = static_cast<const pt_triangle*>(c->tris.p);
    S.planes = static_cast<const pt_plane*>(c->planes.p);
    S.meshes = static_cast<const pt_mesh_instance*>(c->meshes.p);
    S.mesh_root = static_cast<const uint32_t*>(c->mesh_root.p);
    S.lights = static_cast<const pt_light*>(c->lights.p);
    S.sky = static_cast<const uint16_t*>(c->sky.p);
    S.cdf = static_cast<const float*>(c->cdf.p);
    S.n_planes = uint32_t(c->n_planes);
    S.n_meshes = uint32_t(c->n_meshes);
    S.n_lights = uint32_t(c->n_lights);
    S.n_objs = uint32_t(c->n_objs);
    S.sky_w = c->sky_w;
    S.sky_h = c->sky_h;
    S.flags = c->flags | (c->has_textures ? kTexturesBit : 0u) | (c->translucent ? kTranslucentBit : 0u);
    S.tex = static_cast<const uint32_t*>(c->tex.p);
    S.tex_layers = c->tex_layers;
    S.bounce_limit = c->bounce_limit;
    S.stack_cap = c->stack_cap;
    S.wstack_cap = c->wstack_cap;
}

// Per texture layer: does any texel have an alpha below 255?  Scanned once at
// upload; pt_update_objects keeps the uploaded texels and this table.
void scan_layer_alpha(pt_ctx* c, const pt_scene_desc* d) {
    c->layer_alpha.assign(d->texture_layers, 0);
    for (uint32_t l = 0; l < d->texture_layers; ++l) {
        const uint8_t* L = d->textures + size_t(l) * 512 * 512 * 4;
        for (size_t k = 3; k < size_t(512) * 512 * 4; k += 4)
            if (L[k] != 255) { c->layer_alpha[l] = 1; break; }
    }
}

// Some material's albedo texture has an alpha below 1: shadow rays then
// attenuate through such surfaces (evaluate_shadow_attenuation,
// shader.wgsl:777-811) instead of stopping at the first occluder.
bool translucent_albedo(const pt_ctx* c, const pt_scene_desc* d) {
    const pt_material* m = static_cast<const pt_material*>(d->materials.ptr);
    for (uint64_t i = 0; i < d->materials.count; ++i)
        if (m[i].albedo_idx >= 0 && size_t(m[i].albedo_idx) < c->layer_alpha.size() && c->layer_alpha[m[i].albedo_idx])
            return true;
    return false;
}

bool uses_textures(const pt_scene_desc* d) {
    const pt_material* m = static_cast<const pt_material*>(d->materials.ptr);
    for (uint64_t i = 0; i < d->materials.count; ++i)
        if (m[i].albedo_idx >= 0 || m[i].roughness_idx >= 0 || m[i].metallic_idx >= 0 || m[i].emissive_idx >= 0 ||
            m[i].normal_idx >= 0 || m[i].height_idx >= 0)
            return true;
    return false;
}

int upload_dynamic(pt_ctx* c, const pt_scene_desc* d, bool first) {
    HostLayout L;
    std::string err;
    int rc = build_layout(d, L, err, c->tlas_build == PT_TLAS_BUILD_DEVICE);
    if (rc) return c->fail(rc, "%s", err.c_str());
    std::vector<Root>& roots = L.roots;
    const uint32_t td = L.tlas_depth;
    if (!first) {
        // the fast path may only reuse BLAS layouts built at upload
        for (auto& kv : L.slots)
            if (c->blas_slot.find(kv.first) == c->blas_slot.end())
                return c->fail(PT_ERR_CAPACITY, "objects reference a BLAS root (%u) absent at upload", kv.first);
        // remap slots to the uploaded numbering, keep the uploaded BLAS nodes
        std::vector<uint32_t> remap(roots.size(), 0);
        for (auto& kv : L.slots) remap[kv.second] = c->blas_slot[kv.first];
        for (auto& s : L.obj_root) s = remap[s];
        for (auto& s : L.mesh_root) s = remap[s];
        std::vector<Root> merged = c->host_roots;
        merged[0] = roots[0];
        roots = merged;
        const uint32_t need = stack_need(td, c->blas_depth), wneed = stack_need(L.tlas_wneed, c->blas_wneed);
        if (need > kMaxStack || wneed > kMaxStack)
            return c->fail(PT_ERR_UNSUPPORTED, "TLAS depth %u exceeds the %u-entry stack", td, kMaxStack);
        c->stack_cap = need;
        c->wstack_cap = wneed;
    } else {
        const uint32_t need = stack_need(td, L.blas_depth), wneed = stack_need(L.tlas_wneed, L.blas_wneed);
        if (need > kMaxStack || wneed > kMaxStack)
            return c->fail(PT_ERR_UNSUPPORTED, "BVH depth %u+%u exceeds the %u-entry stack", td, L.blas_depth,
                           kMaxStack);
        c->stack_cap = need;
        c->wstack_cap = wneed;
        c->blas_slot = L.slots;
        c->blas_depth = L.blas_depth;
        c->blas_wneed = L.blas_wneed;
        if ((rc = c->upload(c->bnodes, L.bn.data(), L.bn.size() * sizeof(CNode), "bvh nodes"))) return rc;
        if ((rc = c->upload(c->wbnodes, L.wbn.data(), L.wbn.size() * sizeof(WNode), "wide bvh nodes"))) return rc;
        c->n_wbn = L.wbn.size();
    }
    c->tlas_depth = td;
    c->tlas_wneed = L.tlas_wneed;
    c->host_roots = roots;
    if ((rc = c->upload(c->tnodes, L.tn.data(), L.tn.size() * sizeof(CNode), "tlas nodes"))) return rc;
    if ((rc = c->upload(c->wtnodes, L.wtn.data(), L.wtn.size() * sizeof(WNode), "wide tlas nodes"))) return rc;
    c->n_wtn = L.wtn.size();
    if ((rc = c->upload(c->roots, roots.data(), roots.size() * sizeof(Root), "roots"))) return rc;
    if ((rc = c->upload(c->obj_root, L.obj_root.data(), L.obj_root.size() * 4, "object roots"))) return rc;
    if ((rc = c->upload(c->mesh_root, L.mesh_root.data(), L.mesh_root.size() * 4, "mesh roots"))) return rc;
    if ((rc = c->upload(c->mats, d->materials.ptr, d->materials.count * sizeof(pt_material), "materials"))) return rc;
    if ((rc = c->upload(c->objs, d->objects.ptr, d->objects.count * sizeof(pt_object), "objects"))) return rc;
    if ((rc = c->upload(c->tlas_raw, d->tlas.ptr, d->tlas.count * sizeof(pt_bvh_node), "tlas"))) return rc;
    if ((rc = c->upload(c->planes, d->planes.ptr, d->planes.count * sizeof(pt_plane), "planes"))) return rc;
    if ((rc = c->upload(c->meshes, d->meshes.ptr, d->meshes.count * sizeof(pt_mesh_instance), "meshes"))) return rc;
    c->n_mats = d->materials.count;
    c->n_objs = d->objects.count;
    c->n_tlas = d->tlas.count;
    c->n_planes = d->planes.count;
    c->n_meshes = d->meshes.count;
    return c->sync("upload");
}

// The wide traversal's bottom-level trees rebuilt on the device from the
// uploaded triangles (PT_BLAS_BUILD_DEVICE); the child-pair trees and the
// uploaded triangle order stay for the reference-order traversal.  Each BLAS
// must cover one contiguous triangle range, reached through one triangle
// offset.
uint32_t float_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

int device_blas(pt_ctx* c, const pt_scene_desc* d) {
    if (c->blas_slot.empty()) return PT_OK;
    const pt_bvh_node* bl = static_cast<const pt_bvh_node*>(d->bvh.ptr);
    const pt_object* ob = static_cast<const pt_object*>(d->objects.ptr);
    const pt_mesh_instance* mi = static_cast<const pt_mesh_instance*>(d->meshes.ptr);
    std::map<uint32_t, uint32_t> tri_off;  // slot -> triangle offset
    auto note = [&](uint32_t node_offset, uint32_t off) -> int {
        const uint32_t slot = c->blas_slot.at(node_offset);
        auto it = tri_off.find(slot);
        if (it != tri_off.end() && it->second != off)
            return c->fail(PT_ERR_UNSUPPORTED, "BLAS at node %u is used with two triangle offsets", node_offset);
        tri_off[slot] = off;
        return PT_OK;
    };
    int rc;
    if (d->flags & PT_HAS_MESHES)
        for (uint64_t i = 0; i < d->objects.count; ++i)
            if (ob[i].object_type == PT_OBJ_MESH && (rc = note(float_bits(ob[i].param0), float_bits(ob[i].param1))))
                return rc;
    if (d->flags & PT_HAS_LIST_MESHES)
        for (uint64_t i = 0; i < d->meshes.count; ++i)
            if ((rc = note(mi[i].node_offset, mi[i].tri_offset))) return rc;
    std::vector<BlasBuildJob> jobs;
    std::vector<uint32_t> job_slot;
    uint64_t total = 0;
    for (const auto& kv : c->blas_slot) {
        const uint32_t slot = kv.second;
        if (!tri_off.count(slot)) continue;
        // the leaves' triangle range (starts relative to the offset)
        uint64_t lo = UINT64_MAX, hi = 0, count = 0;
        std::vector<uint32_t> st{kv.first};
        while (!st.empty()) {
            const uint32_t i = st.back();
            st.pop_back();
            const pt_bvh_node& n = bl[i];
            if (n.count > 0) {
                lo = std::min<uint64_t>(lo, n.next);
                hi = std::max<uint64_t>(hi, uint64_t(n.next) + n.count);
                count += n.count;
            } else {
                st.push_back(i + 1);
                st.push_back(kv.first + n.next);
            }
        }
        if (count == 0 || count != hi - lo || hi - lo >= (1u << 25))
            return c->fail(PT_ERR_UNSUPPORTED, "BLAS at node %u does not cover one contiguous triangle range", kv.first);
        jobs.push_back(BlasBuildJob{uint32_t(tri_off[slot] + lo), uint32_t(hi - lo), uint32_t(lo)});
        job_slot.push_back(slot);
        total += hi - lo;
    }
    if (jobs.empty()) return PT_OK;
    uint32_t max_n = 0;
    for (const auto& j : jobs) max_n = std::max(max_n, j.n);
    DevBuf scratch, ctr, wn;
    const size_t sb = blas_build_scratch_bytes(max_n);
    if ((rc = c->alloc(scratch, sb, "bvh build scratch")) || (rc = c->alloc(ctr, 64, "bvh build counters")) ||
        (rc = c->alloc(wn, size_t(total) * sizeof(WNode), "device bvh nodes")) ||
        (rc = c->alloc(c->tri_geo_w, size_t(d->triangles.count) * 3 * sizeof(float4), "triangle geometry (wide)"))) {
        scratch.release();
        ctr.release();
        wn.release();
        return rc;
    }
    std::vector<BlasBuildResult> res(jobs.size());
    hipError_t e = hipMemcpyAsync(c->tri_geo_w.p, c->tri_geo.p, size_t(d->triangles.count) * 3 * sizeof(float4),
                                  hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess)
        e = build_blas_device(static_cast<const float4*>(c->tri_geo.p), static_cast<const pt_triangle*>(c->tris.p),
                              jobs.data(), uint32_t(jobs.size()), scratch.p, sb, static_cast<WNode*>(wn.p),
                              static_cast<float4*>(c->tri_geo_w.p), static_cast<uint32_t*>(ctr.p), res.data(),
                              c->stream);
    scratch.release();
    ctr.release();
    if (e != hipSuccess) {
        wn.release();
        return c->hip_fail(e, "device bvh build");
    }
    uint32_t need = 0;
    for (size_t j = 0; j < jobs.size(); ++j) {
        c->host_roots[job_slot[j]].wref = res[j].root_ref;
        need = std::max(need, res[j].need);
    }
    const uint32_t wneed = stack_need(c->tlas_wneed, need);
    if (wneed > kMaxStack) {
        wn.release();
        return c->fail(PT_ERR_UNSUPPORTED, "device-built BVH needs %u stack entries (at most %u)", wneed, kMaxStack);
    }
    c->wbnodes.release();
    c->wbnodes = wn;
    c->n_wbn = total;
    c->blas_wneed = need;
    c->wstack_cap = wneed;
    return c->upload(c->roots, c->host_roots.data(), c->host_roots.size() * sizeof(Root), "roots");
}

// Both forms of the top-level tree built on the device from the uploaded
// objects (PT_TLAS_BUILD_DEVICE), after upload_dynamic has uploaded the
// objects, their BLAS root slots and the BLAS roots.
int device_tlas(pt_ctx* c, const pt_scene_desc* d) {
    const uint32_t any_obj = PT_HAS_SPHERES | PT_HAS_CUBES | PT_HAS_CYLINDERS | PT_HAS_TORI | PT_HAS_MESHES;
    if (!(d->flags & any_obj)) return PT_OK;
    const uint64_t n = d->objects.count;
    if (n == 0) return c->fail(PT_ERR_LAYOUT, "the scene flags name objects but there are none");
    if (n >= (1u << 25)) return c->fail(PT_ERR_UNSUPPORTED, "the device-built TLAS takes fewer than 2^25 objects");
    const size_t nodes = std::max<uint64_t>(n - 1, 1);
    const size_t sb = tlas_build_scratch_bytes(uint32_t(n));
    DevBuf scratch, ctr;
    int rc;
    if ((rc = c->alloc(scratch, sb, "tlas build scratch")) || (rc = c->alloc(ctr, 64, "tlas build counters")) ||
        (rc = c->alloc(c->tnodes, nodes * sizeof(CNode), "tlas nodes")) ||
        (rc = c->alloc(c->wtnodes, nodes * sizeof(WNode), "wide tlas nodes"))) {
        scratch.release();
        ctr.release();
        return rc;
    }
    TlasBuildResult res{};
    const hipError_t e = build_tlas_device(static_cast<const pt_object*>(c->objs.p), uint32_t(n),
                                           static_cast<const uint32_t*>(c->obj_root.p),
                                           static_cast<const Root*>(c->roots.p), scratch.p, sb,
                                           static_cast<CNode*>(c->tnodes.p), static_cast<WNode*>(c->wtnodes.p),
                                           static_cast<uint32_t*>(ctr.p), &res, c->stream);
    scratch.release();
    ctr.release();
    if (e != hipSuccess) return c->hip_fail(e, "device tlas build");
    const uint32_t need = stack_need(res.depth, c->blas_depth), wneed = stack_need(res.need, c->blas_wneed);
    if (need > kMaxStack || wneed > kMaxStack)
        return c->fail(PT_ERR_UNSUPPORTED, "device-built TLAS depth %u needs more than %u stack entries", res.depth,
                       kMaxStack);
    c->n_wtn = nodes;
    c->host_roots[0] = res.root;
    c->tlas_depth = res.depth;
    c->tlas_wneed = res.need;
    c->stack_cap = need;
    c->wstack_cap = wneed;
    return c->upload(c->roots, c->host_roots.data(), c->host_roots.size() * sizeof(Root), "roots");
}

}  // namespace

extern "C" {

int pt_create(pt_ctx** out, int device) {
    if (!out) return PT_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return PT_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return PT_ERR_ARG;
    auto* c = new (std::nothrow) pt_ctx();
    if (!c) return PT_ERR_OOM;
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess || c->n_cu <= 0) { pt_destroy(c); return PT_ERR_HIP; }
    if (pt::noalloc_setup(c) != PT_OK) { pt_destroy(c); return PT_ERR_HIP; }
    *out = c;
    return PT_OK;
}

void pt_destroy(pt_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    pt::unet_free(c->unet);
    c->unet = nullptr;
    c->release_all();
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (hipEvent_t ev : c->bounce_ev) (void)hipEventDestroy(ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* pt_last_error(const pt_ctx* c) { return c ? c->err.c_str() : "null context"; }

int pt_set_option(pt_ctx* c, int key, int64_t value) {
    if (!c) return PT_ERR_ARG;
    switch (key) {
        case PT_OPT_BACKEND:
            if (value != PT_BACKEND_MEGAKERNEL && value != PT_BACKEND_WAVEFRONT)
                return c->fail(PT_ERR_ARG, "unknown backend %lld", (long long)value);
            c->backend = int(value);
            return PT_OK;
        case PT_OPT_COUNT_STATS: c->count_stats = value != 0; return PT_OK;
        case PT_OPT_WAVEFRONT_POOL:
            if (value < 1024 || value > (1ll << 28)) return c->fail(PT_ERR_ARG, "pool size must be in [1024, 2^28]");
            c->wf_pool_opt = uint64_t(value);
            return PT_OK;
        case PT_OPT_WAVEFRONT_REFILL:
            if (value < 1 || value > 64) return c->fail(PT_ERR_ARG, "refill threshold must be in [1, 64]");
            c->wf_refill = uint32_t(value);
            return PT_OK;
        case PT_OPT_LDS_STACK:
            if (value < 0 || value > kMaxStack)
                return c->fail(PT_ERR_ARG, "LDS stack entries must be in [0, %u]", kMaxStack);
            c->lds_stack_opt = uint32_t(value);
            return PT_OK;
        case PT_OPT_TRAVERSAL:
            if (value != PT_TRAVERSAL_REFERENCE && value != PT_TRAVERSAL_WIDE && value != PT_TRAVERSAL_WIDE_QUANTIZED)
                return c->fail(PT_ERR_ARG, "unknown traversal %lld", (long long)value);
            c->traversal = int(value);
            return PT_OK;
        case PT_OPT_BLAS_BUILD:
            if (value != PT_BLAS_BUILD_REFERENCE && value != PT_BLAS_BUILD_DEVICE)
                return c->fail(PT_ERR_ARG, "unknown BLAS build %lld", (long long)value);
            c->blas_build = int(value);
            return PT_OK;
        case PT_OPT_TLAS_BUILD:
            if (value != PT_TLAS_BUILD_REFERENCE && value != PT_TLAS_BUILD_DEVICE)
                return c->fail(PT_ERR_ARG, "unknown TLAS build %lld", (long long)value);
            c->tlas_build = int(value);
            return PT_OK;
        case PT_OPT_SCHEDULE:
            if (value != PT_SCHEDULE_SYNC && value != PT_SCHEDULE_REGEN)
                return c->fail(PT_ERR_ARG, "unknown schedule %lld", (long long)value);
            c->schedule = int(value);
            return PT_OK;
        case PT_OPT_DENOISE_MATH:
            if (value != PT_DENOISE_MATH_FUSED && value != PT_DENOISE_MATH_EXACT)
                return c->fail(PT_ERR_ARG, "unknown denoise math %lld", (long long)value);
            c->denoise_math = int(value);
            return PT_OK;
        case PT_OPT_BOUNCE_TIMING: c->bounce_timing = value != 0; return PT_OK;
        case PT_OPT_KERNEL_FLAGS:
        case PT_OPT_SPECIALISED: return c->fail(PT_ERR_ARG, "option %d is read-only", key);
        default: return c->fail(PT_ERR_ARG, "unknown option %d", key);
    }
}

int pt_get_option(pt_ctx* c, int key, int64_t* value) {
    if (!c) return PT_ERR_ARG;
    if (!value) return c->fail(PT_ERR_ARG, "value is NULL");
    switch (key) {
        case PT_OPT_BACKEND: *value = c->backend; return PT_OK;
        case PT_OPT_COUNT_STATS: *value = c->count_stats ? 1 : 0; return PT_OK;
        case PT_OPT_WAVEFRONT_POOL: *value = int64_t(c->wf_pool_opt); return PT_OK;
        case PT_OPT_SCHEDULE: *value = c->schedule; return PT_OK;
        case PT_OPT_TRAVERSAL: *value = c->traversal; return PT_OK;
        case PT_OPT_LDS_STACK: *value = c->lds_stack_opt; return PT_OK;
        case PT_OPT_WAVEFRONT_REFILL: *value = c->wf_refill; return PT_OK;
        case PT_OPT_BLAS_BUILD: *value = c->blas_build; return PT_OK;
        case PT_OPT_TLAS_BUILD: *value = c->tlas_build; return PT_OK;
        case PT_OPT_DENOISE_MATH: *value = c->denoise_math; return PT_OK;
        case PT_OPT_BOUNCE_TIMING: *value = c->bounce_timing ? 1 : 0; return PT_OK;
        case PT_OPT_KERNEL_FLAGS:
        case PT_OPT_SPECIALISED:
            if (!c->has_scene) return c->fail(PT_ERR_STATE, "no scene uploaded");
            *value = key == PT_OPT_KERNEL_FLAGS ? int64_t(c->dev.flags) : (pt::specialised_flag_set(c->dev.flags) ? 1 : 0);
            return PT_OK;
        default: return c->fail(PT_ERR_ARG, "unknown option %d", key);
    }
}

int pt_upload_scene(pt_ctx* c, const pt_scene_desc* d) {
    if (!c) return PT_ERR_ARG;
    (void)hipSetDevice(c->device);
    int rc = validate_desc(c, d);
    if (rc) return rc;
    c->has_scene = false;
    c->flags = d->flags;
    c->bounce_limit = d->bounce_limit;
    if ((rc = upload_dynamic(c, d, true))) return rc;
    // what pt_update_objects may rewrite in place, however its counts shrink and regrow meanwhile
    c->cap_mats = d->materials.count;
    c->cap_objs = d->objects.count;
    c->cap_tlas = d->tlas.count;
    c->cap_planes = d->planes.count;
    c->cap_meshes = d->meshes.count;

    // triangles: 48 B geometry records + the original 128 B records
    const pt_triangle* tr = static_cast<const pt_triangle*>(d->triangles.ptr);
    std::vector<float4> geo;
    build_tri_geo(d, geo);
    if ((rc = c->upload(c->tri_geo, geo.data(), geo.size() * sizeof(float4), "triangle geometry"))) return rc;
    if ((rc = c->upload(c->tris, tr, d->triangles.count * sizeof(pt_triangle), "triangles"))) return rc;
    if ((rc = c->upload(c->bvh_raw, d->bvh.ptr, d->bvh.count * sizeof(pt_bvh_node), "bvh"))) return rc;
    c->tri_geo_w.release();
    c->tex_layers = d->texture_layers;
    c->has_textures = uses_textures(d);
    scan_layer_alpha(c, d);
    c->translucent = translucent_albedo(c, d);
    if (d->texture_layers > 0) {
        if ((rc = c->upload(c->tex, d->textures, size_t(d->texture_layers) * 512 * 512 * 4, "textures"))) return rc;
    } else {
        c->tex.release();
    }
    if (c->tlas_build == PT_TLAS_BUILD_DEVICE && (rc = device_tlas(c, d))) return rc;
    if (c->blas_build == PT_BLAS_BUILD_DEVICE && (rc = device_blas(c, d))) return rc;
    if ((rc = c->upload(c->lights, d->lights.ptr, d->lights.count * sizeof(pt_light), "lights"))) return rc;
    c->n_bvh = d->bvh.count;
    c->n_tris = d->triangles.count;
    c->n_lights = d->lights.count;
    // sky + CDF
    c->sky_w = d->sky_width;
    c->sky_h = d->sky_height;
    const size_t texels = size_t(d->sky_width) * d->sky_height;
    if ((rc = c->upload(c->sky, d->sky_rgba16f, texels * 8, "sky"))) return rc;
    const size_t cdf_n = size_t(d->sky_height) + 1 + size_t(d->sky_height) * (d->sky_width + 1);
    if (d->sky_cdf) {
        if ((rc = c->upload(c->cdf, d->sky_cdf, cdf_n * 4, "sky cdf"))) return rc;
        c->sky_total_lum = d->sky_total_lum;
        c->sky_total_weight = d->sky_total_lum;
    } else {
        if ((rc = c->alloc(c->cdf, cdf_n * 4, "sky cdf"))) return rc;
        if ((rc = c->alloc(c->cdf_rows, (size_t(d->sky_height) + 1) * sizeof(double), "sky cdf rows"))) return rc;
        double* rows = static_cast<double*>(c->cdf_rows.p);
        hipError_t e = launch_sky_cdf(static_cast<const uint16_t*>(c->sky.p), d->sky_width, d->sky_height,
                                      d->sky_exposure, static_cast<float*>(c->cdf.p), rows, rows + d->sky_height,
                                      c->stream);
        if (e != hipSuccess) return c->hip_fail(e, "sky cdf kernels");
        double total = 0;
        e = hipMemcpyAsync(&total, rows + d->sky_height, sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e != hipSuccess) return c->hip_fail(e, "sky cdf total");
        if ((rc = c->sync("sky cdf"))) return rc;
        c->sky_total_weight = total;
        c->sky_total_lum = float(total);
    }
    if ((rc = quantize_trees(c)) || (rc = c->sync("upload"))) return rc;
    refresh_dev(c);
    c->has_scene = true;
    return PT_OK;
}

int pt_update_objects(pt_ctx* c, const pt_scene_desc* d) {
    if (!c) return PT_ERR_ARG;
    if (!c->has_scene) return c->fail(PT_ERR_STATE, "no scene uploaded");
    (void)hipSetDevice(c->device);
    int rc = validate_desc(c, d);
    if (rc) return rc;
    if (d->flags != c->flags) return c->fail(PT_ERR_FLAGS, "specialisation flags changed");
    if (d->objects.count > c->cap_objs || d->tlas.count > c->cap_tlas || d->planes.count > c->cap_planes ||
        d->meshes.count > c->cap_meshes || d->materials.count > c->cap_mats || d->bvh.count != c->n_bvh ||
        d->triangles.count != c->n_tris || d->texture_layers != c->tex_layers)
        return c->fail(PT_ERR_CAPACITY, "scene arrays outgrew the uploaded buffers");
    c->bounce_limit = d->bounce_limit;
    if ((rc = upload_dynamic(c, d, false))) return rc;
    c->has_textures = uses_textures(d);  // the texels stay those of the upload
    c->translucent = translucent_albedo(c, d);
    if (c->tlas_build == PT_TLAS_BUILD_DEVICE && (rc = device_tlas(c, d))) return rc;
    if ((rc = quantize_trees(c))) return rc;
    refresh_dev(c);
    return c->sync("update");
}

int pt_resize(pt_ctx* c, uint32_t w, uint32_t h, int with_gbuffer) {
    if (!c) return PT_ERR_ARG;
    if (w == 0 || h == 0 || uint64_t(w) * h > (1ull << 31)) return c->fail(PT_ERR_ARG, "bad frame size %ux%u", w, h);
    (void)hipSetDevice(c->device);
    const size_t px = size_t(w) * h;
    int rc;
    if ((rc = c->alloc(c->accum, px * 16, "accum"))) return rc;
    if ((rc = c->alloc(c->out, px * 4, "output"))) return rc;
    hipError_t e = hipMemsetAsync(c->accum.p, 0, px * 16, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->out.p, 0, px * 4, c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "clear frame");
    if (with_gbuffer) {
        if ((rc = c->alloc(c->gbuf, px * 48, "gbuffer"))) return rc;
        e = hipMemsetAsync(c->gbuf.p, 0, px * 48, c->stream);
        if (e != hipSuccess) return c->hip_fail(e, "clear gbuffer");
    }
    c->width = w;
    c->height = h;
    c->with_gbuffer = with_gbuffer != 0 || c->with_gbuffer;
    return c->sync("resize");
}

static int check_params(pt_ctx* c, const pt_scene_params* p) {
    if (!p) return c->fail(PT_ERR_ARG, "params is NULL");
    if (!c->has_scene) return c->fail(PT_ERR_STATE, "no scene uploaded");
    if (c->width == 0) return c->fail(PT_ERR_STATE, "no frame: call pt_resize first");
    if (p->width != c->width || p->height != c->height)
        return c->fail(PT_ERR_ARG, "params frame %ux%u does not match the resized frame %ux%u", p->width, p->height,
                       c->width, c->height);
    if ((c->flags & PT_MIS_SKYBOX) && (p->sky_width != c->sky_w || p->sky_height != c->sky_h ||
                                       p->sky_cdf_index != c->sky_h + 1))
        return c->fail(PT_ERR_ARG, "params sky fields do not match the uploaded sky (%ux%u, cdf index %u)", c->sky_w,
                       c->sky_h, c->sky_h + 1);
    return PT_OK;
}

// Wavefront backend: the call's passes in batches of whole passes, each
// batch bounce_limit x (extend, shade, shadow) + accumulate.
// Allocates the path pool and fills plan.w; launches nothing.
static int prepare_wavefront(pt_ctx* c, pt::TracePlan& plan) {
    const TraceArgs& a = plan.a;
    const uint64_t npix = uint64_t(a.sec_w) * a.sec_h;
    uint64_t per = std::max<uint64_t>(1, c->wf_pool_opt / std::max<uint64_t>(npix, 1));
    per = std::min<uint64_t>(per, a.n_passes);
    const uint64_t pool = per * npix;
    // shadow-queue entries keep the path index below bit 29
    if (pool >= (1ull << 29))
        return c->fail(PT_ERR_ARG, "section of %llu pixels is too large for the wavefront backend (2^29 paths)",
                       (unsigned long long)npix);
    const uint32_t P = uint32_t(pool);
    int rc;
    if (c->wf_pool < P) {
        struct { DevBuf* b; size_t each; const char* what; } bufs[] = {
            {&c->wf_o, 16, "wf origin"}, {&c->wf_d, 16, "wf direction"}, {&c->wf_t, 16, "wf throughput"},
            {&c->wf_r, 16, "wf radiance"}, {&c->wf_meta, 4, "wf meta"}, {&c->wf_ha, 16, "wf hit"},
            {&c->wf_hb, 8, "wf bary"}, {&c->wf_media, 28 * 4, "wf media"}, {&c->wf_q0, 4, "wf queue"},
            {&c->wf_q1, 4, "wf queue"}, {&c->wf_sq, 4, "wf shadow queue"}, {&c->wf_so, 32, "wf shadow o"},
            {&c->wf_sd, 32, "wf shadow d"}, {&c->wf_sc, 32, "wf shadow L"}};
        for (auto& b : bufs) {
            b.b->release();
            if ((rc = c->alloc(*b.b, size_t(P) * b.each, b.what))) { c->wf_pool = 0; return rc; }
        }
        if ((rc = c->alloc(c->wf_cnt, 64, "wf counters"))) return rc;
        c->wf_em.release();
        c->wf_stp.release();
        c->wf_pool = P;
    }
    // the shadow slots' throughput (translucent occluders) and the deferred
    // emission of height-field-blocked surfaces, for the scenes that have them
    if ((c->dev.flags & pt::kTranslucentBit) && c->wf_stp.bytes < size_t(c->wf_pool) * 32u &&
        (rc = c->alloc(c->wf_stp, size_t(c->wf_pool) * 32u, "wf shadow tp")))
        return rc;
    if ((c->dev.flags & PT_HAS_HEIGHTMAPS) && c->wf_em.bytes < size_t(c->wf_pool) * 16u &&
        (rc = c->alloc(c->wf_em, size_t(c->wf_pool) * 16u, "wf deferred emission")))
        return rc;
    WfArgs& w = plan.w;
    w = WfArgs{};
    w.P = a.P;
    w.sec_w = a.sec_w;
    w.sec_h = a.sec_h;
    w.pool = c->wf_pool;
    w.seeds = a.seeds;
    w.accum = a.accum;
    w.out = a.out;
    w.stats = a.stats;
    w.wide = a.wide;
    w.o_pdf = static_cast<float4*>(c->wf_o.p);
    w.d_wsum = static_cast<float4*>(c->wf_d.p);
    w.tp_rng = static_cast<float4*>(c->wf_t.p);
    w.rad = static_cast<float4*>(c->wf_r.p);
    w.meta = static_cast<uint32_t*>(c->wf_meta.p);
    w.hit_a = static_cast<int4*>(c->wf_ha.p);
    w.hit_b = static_cast<float2*>(c->wf_hb.p);
    w.media = static_cast<float*>(c->wf_media.p);
    w.queue[0] = static_cast<uint32_t*>(c->wf_q0.p);
    w.queue[1] = static_cast<uint32_t*>(c->wf_q1.p);
    w.shadow_queue = static_cast<uint32_t*>(c->wf_sq.p);
    w.sh_o = static_cast<float4*>(c->wf_so.p);
    w.sh_d = static_cast<float4*>(c->wf_sd.p);
    w.sh_c = static_cast<float4*>(c->wf_sc.p);
    w.sh_tp = (c->dev.flags & pt::kTranslucentBit) ? static_cast<float4*>(c->wf_stp.p) : nullptr;
    w.sh_em = (c->dev.flags & PT_HAS_HEIGHTMAPS) ? static_cast<float4*>(c->wf_em.p) : nullptr;
    w.counters = static_cast<uint32_t*>(c->wf_cnt.p);
    // persistent traversal launches: grid, stack split, refill threshold
    const bool wide = a.wide && !c->count_stats;
    const uint32_t total = wide ? c->wstack_cap : c->stack_cap;
    w.stack_lds = c->lds_stack_opt ? std::min<uint32_t>(total, c->lds_stack_opt) : total;
    int bpc = 0;
    hipError_t e = wf_trace_blocks_per_cu(c->dev, c->count_stats, a.wide, w.stack_lds, &bpc);
    if (e != hipSuccess) return c->hip_fail(e, "wavefront occupancy");
    if (bpc <= 0) return c->fail(PT_ERR_CAPACITY, "wavefront traversal kernel cannot be resident");
    // the resident grid, or fewer workgroups when a batch has fewer paths than lanes
    const uint64_t resident = uint64_t(bpc) * uint64_t(c->n_cu);
    w.trace_grid = uint32_t(std::min<uint64_t>(resident, (pool + 255u) / 256u));
    w.spill_stride = w.trace_grid * 256u;
    if ((rc = c->alloc(c->wf_spill, size_t(w.spill_stride) * (total - w.stack_lds) * 4u, "wf stack spill"))) return rc;
    w.stack_spill = static_cast<uint32_t*>(c->wf_spill.p);
    w.refill = c->wf_refill;
    plan.per = per;
    return PT_OK;
}

}  // extern "C"

namespace pt {

int check_trace_params(pt_ctx* c, const pt_scene_params* p) { return check_params(c, p); }

int prepare_trace(pt_ctx* c, const pt_scene_params* p, uint32_t sw, uint32_t sh, uint32_t n_passes,
                  const uint32_t* seeds, pt_hit* hits_dev, uint32_t n_rec, TracePlan& plan) {
    int rc;
    if ((rc = c->alloc(c->stats, kStCount * 8, "stats"))) return rc;
    plan = TracePlan{};
    TraceArgs& a = plan.a;
    a.P = *p;
    const uint32_t sx = p->section[0], sy = p->section[1];
    if (sx >= p->width || sy >= p->height) return c->fail(PT_ERR_ARG, "section origin outside the frame");
    a.sec_w = sw ? std::min(sw, p->width - sx) : p->width - sx;
    a.sec_h = sh ? std::min(sh, p->height - sy) : p->height - sy;
    a.n_passes = n_passes;
    a.seeds = nullptr;
    if (seeds && n_passes) {
        if ((rc = c->upload(c->seeds, seeds, size_t(n_passes) * 4, "seeds"))) return rc;
        a.seeds = static_cast<const uint32_t*>(c->seeds.p);
    }
    a.accum = static_cast<float4*>(c->accum.p);
    a.out = static_cast<uchar4*>(c->out.p);
    a.stats = static_cast<unsigned long long*>(c->stats.p);
    a.hits = hits_dev;
    a.n_rec = n_rec;
    a.schedule = c->schedule;
    a.wide = c->traversal;
    plan.wavefront = c->backend == PT_BACKEND_WAVEFRONT && !hits_dev;
    if (plan.wavefront) return prepare_wavefront(c, plan);
    plan.cn = c->count_stats && !hits_dev;
    a.n_tiles = ((a.sec_w + 7u) / 8u) * ((a.sec_h + 7u) / 8u);
    // the kernel variant: wide unless counting or recording (both use the reference order)
    const uint32_t total = (a.wide && !plan.cn && !hits_dev) ? c->wstack_cap : c->stack_cap;
    a.stack_lds = c->lds_stack_opt ? std::min<uint32_t>(total, c->lds_stack_opt) : total;
    int bpc = 0;
    hipError_t e = trace_blocks_per_cu(c->dev, a, plan.cn, &bpc);
    if (e != hipSuccess) return c->hip_fail(e, "trace occupancy");
    if (bpc <= 0)
        return c->fail(PT_ERR_CAPACITY, "trace kernel cannot be resident (LDS stack of %u entries)", c->stack_cap);
    plan.grid = uint32_t(std::min<uint64_t>(a.n_tiles, uint64_t(bpc) * uint64_t(c->n_cu)));
    a.media_stride = plan.grid * 64u;
    if ((rc = c->alloc(c->mk_media, size_t(a.media_stride) * kLaneMediaFloats * 4u, "medium stacks"))) return rc;
    if ((rc = c->alloc(c->mk_tiles, 8 * 4, "tile counters"))) return rc;
    if ((rc = c->alloc(c->mk_spill, size_t(a.media_stride) * (total - a.stack_lds) * 4u, "stack spill"))) return rc;
    a.stack_spill = static_cast<uint32_t*>(c->mk_spill.p);
    a.media = static_cast<float*>(c->mk_media.p);
    a.tile_ctr = static_cast<uint32_t*>(c->mk_tiles.p);
    return PT_OK;
}

hipError_t enqueue_trace(pt_ctx* c, TracePlan& plan, hipStream_t s, bool timed) {
    hipError_t e = launch_zero_u64(plan.a.stats, kStCount, s);
    if (plan.wavefront) {
        WfArgs w = plan.w;
        const TraceArgs& a = plan.a;
        const uint64_t npix = uint64_t(a.sec_w) * a.sec_h;
        // per-bounce timing (measurement only): events around the bounces, a wait per batch
        const size_t nb = size_t(std::max(c->dev.bounce_limit, 0));
        const bool per_bounce = timed && c->bounce_timing;
        c->bounce_ms.assign(nb, 0.0);
        while (per_bounce && e == hipSuccess && c->bounce_ev.size() < nb + 1) {
            hipEvent_t ev = nullptr;
            e = hipEventCreate(&ev);
            if (e == hipSuccess) c->bounce_ev.push_back(ev);
        }
        if (e == hipSuccess && timed) e = hipEventRecord(c->ev0, s);
        for (uint32_t first = 0; e == hipSuccess && first < a.n_passes; first += uint32_t(plan.per)) {
            w.first_pass = first;
            w.batch_passes = uint32_t(std::min<uint64_t>(plan.per, a.n_passes - first));
            w.n_items = uint32_t(w.batch_passes * npix);
            w.write_output = first + w.batch_passes >= a.n_passes ? 1 : 0;
            e = launch_wavefront_batch(c->dev, w, c->count_stats, s, per_bounce ? c->bounce_ev.data() : nullptr);
            if (e == hipSuccess && per_bounce) e = hipEventSynchronize(c->bounce_ev[nb]);
            for (size_t b = 0; e == hipSuccess && per_bounce && b < nb; ++b) {
                float ms = 0.0f;
                e = hipEventElapsedTime(&ms, c->bounce_ev[b], c->bounce_ev[b + 1]);
                c->bounce_ms[b] += ms;
            }
        }
    } else {
        if (e == hipSuccess) e = launch_zero_u64(reinterpret_cast<unsigned long long*>(plan.a.tile_ctr), 4, s);
        if (e == hipSuccess && timed) e = hipEventRecord(c->ev0, s);
        if (e == hipSuccess) e = launch_trace(c->dev, plan.a, plan.grid, plan.cn, s);
    }
    if (e == hipSuccess && timed) e = hipEventRecord(c->ev1, s);
    return e;
}

int read_stats(pt_ctx* c, pt_stats* stats, float ms, uint32_t launches) {
    unsigned long long v[kStCount];
    const hipError_t e = hipMemcpy(v, c->stats.p, sizeof(v), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return c->hip_fail(e, "stats readback");
    std::memset(stats, 0, sizeof(*stats));
    stats->extension_rays = v[kStExt];
    stats->shadow_rays = v[kStShadow];
    stats->paths = v[kStPaths];
    stats->node_pops = v[kStPops];
    stats->inner_pops = v[kStInner];
    stats->object_visits = v[kStObjs];
    stats->tri_tests = v[kStTris];
    stats->shade_events = v[kStShade];
    stats->kernel_ms = ms;
    stats->launches = launches;
    return PT_OK;
}

}  // namespace pt

extern "C" {

static int run_trace(pt_ctx* c, const pt_scene_params* p, uint32_t sw, uint32_t sh, uint32_t n_passes,
                     const uint32_t* seeds, pt_stats* stats, pt_hit* hits_dev, uint32_t n_rec,
                     float* rays_dev = nullptr) {
    pt::TracePlan plan;
    int rc = pt::prepare_trace(c, p, sw, sh, n_passes, seeds, hits_dev, n_rec, plan);
    if (rc) return rc;
    plan.a.rec_rays = rays_dev;
    const hipError_t e = pt::enqueue_trace(c, plan, c->stream, true);
    if (e != hipSuccess) return c->hip_fail(e, "trace launch");
    if ((rc = c->sync("trace"))) return rc;
    if (stats) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
        return pt::read_stats(c, stats, ms, 1);
    }
    return PT_OK;
}

int pt_get_bounce_stats(pt_ctx* c, uint64_t* ext_rays, uint64_t* shadow_rays, double* ms, uint32_t n) {
    if (!c) return PT_ERR_ARG;
    if (n > uint32_t(pt::kStatBounces)) return c->fail(PT_ERR_ARG, "at most %d bounces are counted", pt::kStatBounces);
    if (!c->stats.p) return c->fail(PT_ERR_STATE, "no trace call yet");
    unsigned long long v[pt::kStCount];
    int rc = c->sync("bounce stats");
    if (rc) return rc;
    const hipError_t e = hipMemcpy(v, c->stats.p, sizeof(v), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return c->hip_fail(e, "bounce stats readback");
    for (uint32_t b = 0; b < n; ++b) {
        if (ext_rays) ext_rays[b] = v[pt::kStExtB0 + b];
        if (shadow_rays) shadow_rays[b] = v[pt::kStShadowB0 + b];
        if (ms) ms[b] = b < c->bounce_ms.size() ? c->bounce_ms[b] : 0.0;
    }
    return PT_OK;
}

int pt_trace(pt_ctx* c, const pt_scene_params* p, uint32_t sw, uint32_t sh, uint32_t n_passes, const uint32_t* seeds,
             pt_stats* stats) {
    if (!c) return PT_ERR_ARG;
    (void)hipSetDevice(c->device);
    int rc = check_params(c, p);
    if (rc) return rc;
    if (n_passes == 0) { if (stats) std::memset(stats, 0, sizeof(*stats)); return PT_OK; }
    return run_trace(c, p, sw, sh, n_passes, seeds, stats, nullptr, 0);
}

int pt_trace_hits(pt_ctx* c, const pt_scene_params* p, uint32_t n_bounces, pt_hit* hits) {
    return pt_trace_hits_rays(c, p, n_bounces, hits, nullptr);
}

int pt_trace_hits_rays(pt_ctx* c, const pt_scene_params* p, uint32_t n_bounces, pt_hit* hits, float* rays) {
    if (!c) return PT_ERR_ARG;
    (void)hipSetDevice(c->device);
    int rc = check_params(c, p);
    if (rc) return rc;
    if (!hits || n_bounces == 0 || n_bounces > 64) return c->fail(PT_ERR_ARG, "bad hit buffer / bounce count");
    const size_t n = size_t(p->width) * p->height * n_bounces;
    std::vector<pt_hit> init(n);
    for (auto& h : init) {
        std::memset(&h, 0, sizeof(h));
        h.t = -1.0f;
        h.o_idx = h.m_idx = h.prim_id = h.plane_idx = -1;
    }
    if ((rc = c->upload(c->rays_out, init.data(), n * sizeof(pt_hit), "hit records"))) return rc;
    if (rays) {  // bounces a path never reached keep zero rays
        std::vector<float> zero(n * 6, 0.0f);
        if ((rc = c->upload(c->rays_aux2, zero.data(), n * 6 * sizeof(float), "ray records"))) return rc;
    }
    pt_scene_params q = *p;
    q.section[0] = q.section[1] = 0;
    if ((rc = run_trace(c, &q, 0, 0, 1, nullptr, nullptr, static_cast<pt_hit*>(c->rays_out.p), n_bounces,
                        rays ? static_cast<float*>(c->rays_aux2.p) : nullptr)))
        return rc;
    hipError_t e = hipMemcpy(hits, c->rays_out.p, n * sizeof(pt_hit), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return c->hip_fail(e, "hit readback");
    if (rays && (e = hipMemcpy(rays, c->rays_aux2.p, n * 6 * sizeof(float), hipMemcpyDeviceToHost)) != hipSuccess)
        return c->hip_fail(e, "ray readback");
    return PT_OK;
}

int pt_gbuffer(pt_ctx* c, const pt_scene_params* p) {
    if (!c) return PT_ERR_ARG;
    (void)hipSetDevice(c->device);
    int rc = check_params(c, p);
    if (rc) return rc;
    if (!c->gbuf.p) return c->fail(PT_ERR_STATE, "frame has no G-buffer: pt_resize(..., with_gbuffer=1)");
    hipError_t e = launch_gbuffer(c->dev, *p, static_cast<pt_gbuffer_texel*>(c->gbuf.p), static_cast<uchar4*>(c->out.p),
                                  c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "gbuffer launch");
    return c->sync("gbuffer");
}

int pt_denoise(pt_ctx* c, const pt_denoise_params* u, int stride_mode) {
    if (!c) return PT_ERR_ARG;
    if (!u) return c->fail(PT_ERR_ARG, "params is NULL");
    if (stride_mode != 0 && stride_mode != 1) return c->fail(PT_ERR_ARG, "stride_mode must be 0 or 1");
    if (c->width == 0 || !c->gbuf.p) return c->fail(PT_ERR_STATE, "denoise needs a frame with a G-buffer");
    if (u->width != c->width || u->height != c->height) return c->fail(PT_ERR_ARG, "denoise size does not match the frame");
    if (u->passes > 16) return c->fail(PT_ERR_ARG, "at most 16 a-trous passes");
    (void)hipSetDevice(c->device);
    const size_t px = size_t(c->width) * c->height;
    int rc;
    if ((rc = c->alloc(c->dn_a, px * 16, "denoise A")) || (rc = c->alloc(c->dn_b, px * 16, "denoise B")) ||
        (rc = c->alloc(c->dn_fin, px * 16, "denoise out")) || (rc = c->alloc(c->dn_guide, px * 48, "denoise guide")))
        return rc;
    DenoiseArgs a{};
    a.u = *u;
    a.accum = static_cast<const float4*>(c->accum.p);
    a.gbuf = static_cast<const pt_gbuffer_texel*>(c->gbuf.p);
    a.gn = static_cast<float4*>(c->dn_guide.p);
    a.ga = a.gn + px;
    a.gm = a.ga + px;
    a.exact = c->denoise_math == PT_DENOISE_MATH_EXACT ? 1 : 0;
    a.a = static_cast<float4*>(c->dn_a.p);
    a.b = static_cast<float4*>(c->dn_b.p);
    a.fin = static_cast<float4*>(c->dn_fin.p);
    a.out = static_cast<uchar4*>(c->out.p);
    hipError_t e = launch_denoise(a, stride_mode, c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "denoise launch");
    return c->sync("denoise");
}

static int copy_out(pt_ctx* c, void* dst, const DevBuf& b, uint64_t bytes, uint64_t need, const char* what) {
    if (!c) return PT_ERR_ARG;
    if (!dst || bytes < need) return c->fail(PT_ERR_ARG, "%s: destination smaller than %llu bytes", what, (unsigned long long)need);
    if (!b.p) return c->fail(PT_ERR_STATE, "%s: buffer not allocated", what);
    (void)hipSetDevice(c->device);
    int rc = c->sync(what);
    if (rc) return rc;
    hipError_t e = hipMemcpy(dst, b.p, need, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return c->hip_fail(e, what);
    return PT_OK;
}
static int copy_in(pt_ctx* c, DevBuf& b, const void* src, uint64_t bytes, uint64_t need, const char* what) {
    if (!c) return PT_ERR_ARG;
    if (!src || bytes != need) return c->fail(PT_ERR_ARG, "%s: expected exactly %llu bytes", what, (unsigned long long)need);
    if (!b.p) return c->fail(PT_ERR_STATE, "%s: buffer not allocated", what);
    (void)hipSetDevice(c->device);
    hipError_t e = hipMemcpy(b.p, src, need, hipMemcpyHostToDevice);
    if (e != hipSuccess) return c->hip_fail(e, what);
    return PT_OK;
}

int pt_read_accum(pt_ctx* c, float* dst, uint64_t bytes) {
    return copy_out(c, dst, c ? c->accum : DevBuf{}, bytes, c ? uint64_t(c->width) * c->height * 16 : 0, "read accum");
}
int pt_write_accum(pt_ctx* c, const float* src, uint64_t bytes) {
    if (!c) return PT_ERR_ARG;
    return copy_in(c, c->accum, src, bytes, uint64_t(c->width) * c->height * 16, "write accum");
}
int pt_read_output(pt_ctx* c, uint8_t* dst, uint64_t bytes) {
    return copy_out(c, dst, c ? c->out : DevBuf{}, bytes, c ? uint64_t(c->width) * c->height * 4 : 0, "read output");
}
int pt_read_gbuffer(pt_ctx* c, pt_gbuffer_texel* dst, uint64_t bytes) {
    return copy_out(c, dst, c ? c->gbuf : DevBuf{}, bytes, c ? uint64_t(c->width) * c->height * 48 : 0, "read gbuffer");
}
int pt_write_gbuffer(pt_ctx* c, const pt_gbuffer_texel* src, uint64_t bytes) {
    if (!c) return PT_ERR_ARG;
    return copy_in(c, c->gbuf, src, bytes, uint64_t(c->width) * c->height * 48, "write gbuffer");
}
int pt_read_denoised(pt_ctx* c, float* dst, uint64_t bytes) {
    return copy_out(c, dst, c ? c->dn_fin : DevBuf{}, bytes, c ? uint64_t(c->width) * c->height * 16 : 0, "read denoised");
}
int pt_read_sky_cdf(pt_ctx* c, float* dst, uint64_t count, float* total) {
    if (!c) return PT_ERR_ARG;
    const uint64_t need = uint64_t(c->sky_h) + 1 + uint64_t(c->sky_h) * (c->sky_w + 1);
    if (!c->has_scene) return c->fail(PT_ERR_STATE, "no scene uploaded");
    int rc = copy_out(c, dst, c->cdf, count * 4, need * 4, "read sky cdf");
    if (rc) return rc;
    if (total) *total = c->sky_total_lum;
    return PT_OK;
}

int pt_trace_rays(pt_ctx* c, const float* rays, uint32_t n, pt_hit* hits) {
    if (!c) return PT_ERR_ARG;
    if (!c->has_scene) return c->fail(PT_ERR_STATE, "no scene uploaded");
    if (n == 0) return PT_OK;
    if (!rays || !hits) return c->fail(PT_ERR_ARG, "rays/hits is NULL");
    (void)hipSetDevice(c->device);
    int rc;
    if ((rc = c->upload(c->rays_in, rays, size_t(n) * 24, "rays"))) return rc;
    if ((rc = c->alloc(c->rays_out, size_t(n) * sizeof(pt_hit), "hits"))) return rc;
    RayArgs a{};
    a.rays = static_cast<const float*>(c->rays_in.p);
    a.n = n;
    a.hits = static_cast<pt_hit*>(c->rays_out.p);
    a.wide = c->traversal;
    hipError_t e = launch_rays(c->dev, a, false, c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "rays launch");
    if ((rc = c->sync("rays"))) return rc;
    e = hipMemcpy(hits, c->rays_out.p, size_t(n) * sizeof(pt_hit), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return c->hip_fail(e, "hits readback");
    return PT_OK;
}

int pt_trace_shadow_rays(pt_ctx* c, const float* rays, const int32_t* target_idx, const float* target_dist, uint32_t n,
                         float* trans) {
    if (!c) return PT_ERR_ARG;
    if (!c->has_scene) return c->fail(PT_ERR_STATE, "no scene uploaded");
    if (n == 0) return PT_OK;
    if (!rays || !target_idx || !target_dist || !trans) return c->fail(PT_ERR_ARG, "NULL argument");
    (void)hipSetDevice(c->device);
    int rc;
    if ((rc = c->upload(c->rays_in, rays, size_t(n) * 24, "rays")) ||
        (rc = c->upload(c->rays_aux, target_idx, size_t(n) * 4, "targets")) ||
        (rc = c->upload(c->rays_aux2, target_dist, size_t(n) * 4, "distances")) ||
        (rc = c->alloc(c->rays_out, size_t(n) * 12, "transmittance")))
        return rc;
    RayArgs a{};
    a.rays = static_cast<const float*>(c->rays_in.p);
    a.target = static_cast<const int32_t*>(c->rays_aux.p);
    a.tdist = static_cast<const float*>(c->rays_aux2.p);
    a.n = n;
    a.trans = static_cast<float*>(c->rays_out.p);
    a.wide = c->traversal;
    hipError_t e = launch_rays(c->dev, a, true, c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "shadow rays launch");
    if ((rc = c->sync("shadow rays"))) return rc;
    e = hipMemcpy(trans, c->rays_out.p, size_t(n) * 12, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return c->hip_fail(e, "transmittance readback");
    return PT_OK;
}

int pt_synchronize(pt_ctx* c) {
    if (!c) return PT_ERR_ARG;
    return c->sync("synchronize");
}

}  // extern "C"
