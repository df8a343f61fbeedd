// pt_ctx.h — the context behind the C ABI's opaque pt_ctx handle: the
// device, stream and every device buffer of a scene, a frame, the wavefront
// pool and the learned denoiser.  Private to libpt.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "pt_abi.h"
#include "pt_internal.h"

namespace pt {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

struct UNet;               // pt_unet.cpp
void unet_free(UNet* u);

}  // namespace pt

struct pt_ctx {
    using DevBuf = pt::DevBuf;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // PT_OPT_BOUNCE_TIMING: events around the wavefront bounces, and the last
    // call's device time per bounce
    bool bounce_timing = false;
    std::vector<hipEvent_t> bounce_ev;
    std::vector<double> bounce_ms;
    std::string err;
    // options
    int backend = PT_BACKEND_WAVEFRONT;
    int schedule = PT_SCHEDULE_REGEN;
    int traversal = PT_TRAVERSAL_WIDE;
    bool count_stats = true;
    int denoise_math = PT_DENOISE_MATH_FUSED;
    // scene
    bool has_scene = false;
    uint32_t flags = 0;
    int bounce_limit = 8;
    DevBuf mats, objs, obj_root, tnodes, bnodes, wtnodes, wbnodes, roots, tlas_raw, bvh_raw, tri_geo, tris, planes, meshes, mesh_root,
        lights, sky, cdf, cdf_rows;
    uint64_t n_mats = 0, n_objs = 0, n_tlas = 0, n_planes = 0, n_meshes = 0, n_lights = 0, n_bvh = 0, n_tris = 0;
    // counts at the last pt_upload_scene: the most pt_update_objects accepts
    uint64_t cap_mats = 0, cap_objs = 0, cap_tlas = 0, cap_planes = 0, cap_meshes = 0;
    uint32_t sky_w = 1, sky_h = 1;
    float sky_total_lum = 0.f;
    double sky_total_weight = 0.0;
    std::map<uint32_t, uint32_t> blas_slot;   // node_offset -> root slot
    std::vector<pt::Root> host_roots;
    uint32_t tlas_depth = 0, blas_depth = 0, stack_cap = 2;
    uint32_t blas_wneed = 0, wstack_cap = 2, tlas_wneed = 0;
    int blas_build = PT_BLAS_BUILD_REFERENCE;  // where the wide BLAS trees come from
    int tlas_build = PT_TLAS_BUILD_REFERENCE;  // where both TLAS forms come from
    DevBuf tex;                      // material texture array (512 x 512 rgba8 layers)
    uint32_t tex_layers = 0;
    bool has_textures = false;       // some material samples a texture
    bool translucent = false;        // some albedo texture has an alpha below 1
    std::vector<uint8_t> layer_alpha;  // per uploaded texture layer: some alpha below 255
    DevBuf tri_geo_w;                          // device-built trees: triangle records in their leaf order
    DevBuf qtnodes, qbnodes;                   // the wide trees with quantised boxes
    uint64_t n_wtn = 0, n_wbn = 0;             // wide nodes in wtnodes / wbnodes
    pt::DevScene dev{};
    // frame
    uint32_t width = 0, height = 0;
    bool with_gbuffer = false;
    DevBuf accum, out, gbuf, dn_a, dn_b, dn_fin, dn_guide, stats, seeds, rays_in, rays_aux, rays_aux2, rays_out;
    // wavefront path pool
    uint64_t wf_pool_opt = 1ull << 27;  // paths per wavefront batch (~312 B each)
    uint32_t wf_pool = 0;
    DevBuf wf_o, wf_d, wf_t, wf_r, wf_meta, wf_ha, wf_hb, wf_media, wf_q0, wf_q1, wf_sq, wf_so, wf_sd, wf_stp, wf_sc,
        wf_cnt, wf_em;
    // persistent megakernel: per-lane medium stacks, per-XCD tile counters
    DevBuf mk_media, mk_tiles, mk_spill, wf_spill;
    uint32_t wf_refill = 56;      // wavefront traversal: idle lanes of a wave that trigger a refill
    uint32_t lds_stack_opt = 16;  // traversal stack entries per lane in LDS (0: all)
    int n_cu = 0;

    int fail(int code, const char* fmt, ...) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        err = buf;
        return code;
    }
    int hip_fail(hipError_t e, const char* what) {
        return fail(e == hipErrorOutOfMemory ? PT_ERR_OOM : PT_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    }
    int alloc(DevBuf& b, size_t bytes, const char* what) {
        if (b.p && b.bytes >= bytes && bytes > 0) return PT_OK;
        b.release();
        const size_t n = std::max<size_t>(bytes, 16);
        hipError_t e = hipMalloc(&b.p, n);
        if (e != hipSuccess) { b.p = nullptr; return hip_fail(e, what); }
        b.bytes = n;
        return PT_OK;
    }
    int upload(DevBuf& b, const void* src, size_t bytes, const char* what) {
        int rc = alloc(b, bytes, what);
        if (rc) return rc;
        if (bytes == 0) return PT_OK;
        hipError_t e = hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, stream);
        if (e != hipSuccess) return hip_fail(e, what);
        return PT_OK;
    }
    int sync(const char* what) {
        hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return hip_fail(e, what);
        return PT_OK;
    }
    pt::UNet* unet = nullptr;  // the learned denoiser, created on first use
    void release_all() {
        DevBuf* all[] = {&mats, &objs, &obj_root, &tnodes, &bnodes, &wtnodes, &wbnodes, &roots, &tlas_raw, &bvh_raw, &tri_geo, &tris,
                         &planes, &meshes, &mesh_root, &lights, &sky, &cdf, &cdf_rows, &tri_geo_w, &accum, &out, &gbuf, &dn_a,
                         &dn_b, &dn_fin, &dn_guide, &stats, &seeds, &rays_in, &rays_aux, &rays_aux2, &rays_out, &qtnodes, &qbnodes,
                         &wf_o, &wf_d, &wf_t, &wf_r, &wf_meta, &wf_ha, &wf_hb, &wf_media, &wf_q0, &wf_q1, &wf_sq,
                         &wf_so, &wf_sd, &wf_stp, &wf_sc, &wf_cnt, &wf_em, &mk_media, &mk_tiles, &mk_spill, &wf_spill,
                         &tex};
        for (DevBuf* b : all) b->release();
    }
};

namespace pt {

// A trace call split in two: prepare_trace allocates, uploads the seeds and
// fills the kernel arguments (outside any stream capture); enqueue_trace only
// launches kernels (zeroing the stats with a kernel), so a batch can record it
// into a graph.
struct TracePlan {
    TraceArgs a{};
    WfArgs w{};
    uint64_t per = 0;      // wavefront: passes per batch
    bool wavefront = false;
    bool cn = false;       // megakernel: counting variant
    uint32_t grid = 0;     // megakernel: persistent grid
};
int check_trace_params(pt_ctx* c, const pt_scene_params* p);
int prepare_trace(pt_ctx* c, const pt_scene_params* p, uint32_t sw, uint32_t sh, uint32_t n_passes,
                  const uint32_t* seeds, pt_hit* hits_dev, uint32_t n_rec, TracePlan& plan);
hipError_t enqueue_trace(pt_ctx* c, TracePlan& plan, hipStream_t s, bool timed);
int read_stats(pt_ctx* c, pt_stats* stats, float ms, uint32_t launches);
int noalloc_setup(pt_ctx* c);  // pt_noalloc.cpp: the device address of this device's scratch array

}  // namespace pt

