// pt_internal.h — interface between the C ABI (pt_api.cpp) and the kernels
// (pt_kernels.hip).  Not installed; not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include "pt_abi.h"
#include "dev/geometry.h"

namespace pt {

// Floats of one lane's medium stack (4 media x ior, sigma rgb, emission rgb).
constexpr uint32_t kLaneMediaFloats = 28;

// Counter slots in the device stats buffer (uint64 each).
// per-bounce ray counts (wavefront backend): kStatBounces slots each
constexpr int kStatBounces = 16;
enum {
    kStExt = 0, kStShadow, kStPaths, kStPops, kStInner, kStObjs, kStTris, kStShade,
    kStExtB0, kStShadowB0 = kStExtB0 + kStatBounces, kStCount = kStShadowB0 + kStatBounces
};

struct TraceArgs {
    pt_scene_params P;
    uint32_t sec_w, sec_h;
    uint32_t n_passes;
    const uint32_t* seeds;       // device, n_passes entries (NULL: P.seed)
    float4* accum;
    uchar4* out;
    unsigned long long* stats;   // kStCount counters
    pt_hit* hits;                // hit recording (NULL: off)
    float* rec_rays;             // with hits: each recorded bounce's ray, 6 floats (NULL: not kept)
    uint32_t n_rec;
    int schedule;                // PT_SCHEDULE_*
    int wide;                    // the traversal: PT_TRAVERSAL_* (0 child pairs, 1 wide, 2 wide quantised)
    uint32_t n_tiles;            // 8x8 tiles of the section
    uint32_t* tile_ctr;          // 8 per-XCD band counters, zero at launch
    float* media;                // kMediaFloats per lane of the grid, structure of arrays
    uint32_t media_stride;       // grid workgroups x 64
    uint32_t stack_lds;          // traversal stack entries per lane kept in LDS
    uint32_t* stack_spill;       // the rest: (stack capacity - stack_lds) per lane, same layout as media
};

struct RayArgs {
    const float* rays;           // 6 floats per ray
    const int32_t* target;       // shadow: per-ray ignored object
    const float* tdist;          // shadow: per-ray distance
    uint32_t n;
    pt_hit* hits;
    float* trans;
    int wide;                    // the traversal: PT_TRAVERSAL_* (0 child pairs, 1 wide, 2 wide quantised)
};

struct DenoiseArgs {
    pt_denoise_params u;
    const float4* accum;
    const pt_gbuffer_texel* gbuf;
    float4* gn;                  // guide planes, one float4 per pixel each (see dn_guide):
    float4* ga;                  //   (normal, |normal| > 0.1), (tonemap(albedo), dist/(dist+3)),
    float4* gm;                  //   (transmission, roughness, metallic, 0)
    float4* a;
    float4* b;
    float4* fin;                 // final fp32 result
    uchar4* out;
    int exact;                   // 1: the a-trous weights with the library exp/pow, one per factor
};

// Wavefront batch (pt_wavefront.hip): device state of `pool` path slots;
// slot i of a batch is pass first_pass + i / npix of section pixel i % npix.
struct WfArgs {
    pt_scene_params P;
    uint32_t sec_w, sec_h;
    uint32_t n_items;
    uint32_t pool;
    uint32_t first_pass, batch_passes;
    int write_output;
    const uint32_t* seeds;
    float4* accum;
    uchar4* out;
    unsigned long long* stats;
    float4* o_pdf;      // ray origin, last BSDF pdf
    float4* d_wsum;     // ray direction, last light weight sum
    float4* tp_rng;     // throughput, PCG state bits
    float4* rad;        // radiance
    uint32_t* meta;     // medium stack depth
    int4* hit_a;        // t bits, kind<<30 | index, triangle, node pops
    float2* hit_b;      // barycentrics
    float* media;       // 28 floats per slot, structure of arrays with stride pool
    uint32_t* queue[2];
    uint32_t* shadow_queue;
    float4* sh_o;       // per slot x2: origin, distance
    float4* sh_d;       //              direction, ignored object bits
    float4* sh_c;       //              throughput x clamped unoccluded radiance, w: throughput x 0 bits
                        //              (translucent occluders: unclamped radiance, bounce in w)
    float4* sh_tp;      //              translucent occluders only (else NULL): throughput
    float4* sh_em;      // per path, height-mapped scenes only (else NULL): emission of a surface the height
                        // field blocked, to be added after the shadow slots; w = 1 while pending
    uint32_t* counters; // per extension queue its length and its shadow queue's, two fetch cursors
    int wide;           // the traversal: PT_TRAVERSAL_* (0 child pairs, 1 wide, 2 wide quantised)
    uint32_t trace_grid;    // persistent traversal launches: workgroups (all resident)
    uint32_t refill;        // dynamic fetch: idle lanes of a wave that trigger a refill
    uint32_t stack_lds;     // traversal stack entries per lane in LDS
    uint32_t* stack_spill;  // the rest, one column per lane of trace_grid x 256
    uint32_t spill_stride;
};

// Launchers (pt_kernels.hip).  All asynchronous on `stream`.
// The trace kernel is persistent: launch_trace takes the grid size, normally
// trace_blocks_per_cu x CUs (capped by the tile count); a.media must hold
// kMediaFloats x grid x 64 floats.
hipError_t trace_blocks_per_cu(const DevScene& S, const TraceArgs& a, bool count_nodes, int* blocks_per_cu);
hipError_t launch_trace(const DevScene& S, const TraceArgs& a, uint32_t grid, bool count_nodes, hipStream_t stream);
hipError_t launch_gbuffer(const DevScene& S, const pt_scene_params& P, pt_gbuffer_texel* gbuf, uchar4* out,
                          hipStream_t stream);
hipError_t launch_rays(const DevScene& S, const RayArgs& a, bool shadow, hipStream_t stream);
hipError_t launch_denoise(const DenoiseArgs& a, int stride_mode, hipStream_t stream);
hipError_t launch_sky_cdf(const uint16_t* sky, uint32_t w, uint32_t h, float exposure, float* cdf, double* row_tmp,
                          double* total, hipStream_t stream);
hipError_t launch_zero_u64(unsigned long long* p, uint32_t n, hipStream_t stream);

// Bottom-level trees built on the device (pt_build.hip).  Job j covers the
// triangles [t0, t0 + n) of the scene; its leaves store starts relative to
// the triangle offset its objects use (leaf_base = t0 - that offset).
struct BlasBuildJob {
    uint32_t t0, n, leaf_base;
};
struct BlasBuildResult {
    uint32_t root_ref;  // wide node index, or a leaf reference
    uint32_t need;      // largest sum of (children - 1) along a path
    bool leaf_root;
};
size_t blas_build_scratch_bytes(uint32_t max_n);
// Writes the wide nodes of every job into wn (at most the jobs' total
// triangle count of them, allocated from index 0) and the triangle records
// of every job's range, in leaf order, into geo_out; ctr: 4 words of device
// scratch.  Synchronous.
hipError_t build_blas_device(const float4* geo, const pt_triangle* tris, const BlasBuildJob* jobs, uint32_t n_jobs,
                             void* scratch, size_t scratch_bytes, WNode* wn, float4* geo_out, uint32_t* ctr,
                             BlasBuildResult* res, hipStream_t stream);
// The top-level tree built on the device (pt_build.hip) over the objects'
// world boxes (from their inverse matrices and, for meshes, the BLAS root
// boxes in roots[obj_root[i]]), one object per leaf referenced by its own
// index: child-pair nodes into cn (n - 1 of them) and 4-wide nodes into wn
// (at most n - 1); ctr: 4 words of device scratch.  n <= 2^25.  Synchronous.
struct TlasBuildResult {
    Root root;       // box, child-pair ref and wide ref of the root
    uint32_t depth;  // child-pair tree depth
    uint32_t need;   // largest sum of (children - 1) along a wide path
};
size_t tlas_build_scratch_bytes(uint32_t n);
hipError_t build_tlas_device(const pt_object* objs, uint32_t n, const uint32_t* obj_root, const Root* roots,
                             void* scratch, size_t scratch_bytes, CNode* cn, WNode* wn, uint32_t* ctr,
                             TlasBuildResult* res, hipStream_t stream);
// bounce_ev: NULL, or bounce_limit + 1 events recorded around the bounces.
hipError_t launch_wavefront_batch(const DevScene& S, const WfArgs& a, bool count_nodes, hipStream_t stream,
                                  hipEvent_t* bounce_ev);
// QNodes from WNodes, index for index (pt_build.hip).
hipError_t quantize_wide_nodes(const WNode* in, QNode* out, uint32_t n, hipStream_t stream);
// True when the trace kernels have a variant compiled for this kernel flag
// word (both backends dispatch on the same sets); otherwise the run-time-flags
// variant runs.
bool specialised_flag_set(uint32_t flags);
// Resident workgroups per CU of the wavefront traversal kernels (the larger of the two).
hipError_t wf_trace_blocks_per_cu(const DevScene& S, bool count_nodes, int wide, uint32_t stack_lds,
                                  int* blocks_per_cu);

}  // namespace pt
