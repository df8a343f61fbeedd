/*
 * pt_abi.h — C ABI of the MI355X path-tracer hot path (libpt.so).
 *
 * This is the drop-in boundary for the reference's WebGPU compute dispatch:
 * every entry point replaces one method of tracer.js's Renderer / Denoiser
 * (cited per function).  Buffers cross the boundary as plain pointers and
 * byte counts in the layouts of pt_layouts.h; the callee copies what it needs
 * before returning (the queue.writeBuffer copy semantics), keeps no caller
 * pointer, and owns all device memory.  No exception, torch type or HIP type
 * crosses the boundary: every function returns PT_OK (0) or a negative
 * PT_ERR_* code, with a message in pt_last_error().
 *
 * Threading: a context is confined to one caller thread and is not
 * re-entrant.  Each context owns one HIP stream; calls are synchronous with
 * respect to the caller (they return after the device work has finished),
 * which mirrors `await device.queue.onSubmittedWorkDone()`.
 */
#ifndef PT_ABI_H
#define PT_ABI_H

#include "pt_layouts.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PT_OK = 0,
    PT_ERR_ARG = -1,         /* null pointer, zero size, bad enum */
    PT_ERR_LAYOUT = -2,      /* stride / byte count does not match pt_layouts.h */
    PT_ERR_CAPACITY = -3,    /* pt_update_objects: arrays outgrew the upload */
    PT_ERR_FLAGS = -4,       /* pt_update_objects: specialisation flags changed */
    PT_ERR_HIP = -5,         /* a HIP runtime call failed */
    PT_ERR_OOM = -6,         /* device allocation failed */
    PT_ERR_STATE = -7,       /* no scene uploaded / no frame allocated */
    PT_ERR_UNSUPPORTED = -8, /* feature outside this build (normal/height maps off spheres, cubes, planes) */
    PT_ERR_NO_DEVICE = -9    /* no HIP device visible */
};

typedef struct pt_ctx pt_ctx;

/* One typed array: `count` records of `stride` bytes each. */
typedef struct pt_array {
    const void* ptr;
    uint64_t count;
    uint32_t stride;
    uint32_t _pad;
} pt_array;

/* Everything Renderer.setScene uploads (tracer.js:1955-2102). */
typedef struct pt_scene_desc {
    pt_array materials;  /* pt_material,      stride 144 (binding 3)  */
    pt_array objects;    /* pt_object,        stride 96  (binding 4)  */
    pt_array tlas;       /* pt_bvh_node,      stride 32  (binding 5)  */
    pt_array meshes;     /* pt_mesh_instance, stride 80  (binding 6)  */
    pt_array bvh;        /* pt_bvh_node,      stride 32  (binding 7)  */
    pt_array triangles;  /* pt_triangle,      stride 128 (binding 8)  */
    pt_array planes;     /* pt_plane,         stride 32  (binding 9)  */
    pt_array lights;     /* pt_light,         stride 80  (binding 10) */
    /* Sky: rgba16float texels, W*H*4 halfs (binding 13). */
    uint32_t sky_width;
    uint32_t sky_height;
    const uint16_t* sky_rgba16f;
    /* Sky CDF in the HDRTexture.buildHDRCDF layout (tracer.js:135-204):
     * [marginal: H+1][conditional: H*(W+1)] f32.  When sky_cdf is NULL the
     * CDF and its total weight are built on the device from the texels
     * (sky_exposure applied as buildHDRCDF applies it). */
    const float* sky_cdf;
    uint64_t sky_cdf_count;
    float sky_exposure;
    float sky_total_lum;    /* ignored when the CDF is built on the device */
    /* Texture array, 512x512 rgba8 layers, rows top first (binding 11),
     * sampled as tracer.js sets it up (linear, repeat, level 0).  Materials
     * may use albedo (an alpha below 1 makes shadow rays attenuate through
     * the surface), roughness, metallic, emissive, normal and height
     * (parallax occlusion) maps; a material with a
     * normal or height map used by a cylinder, torus or mesh is rejected with
     * PT_ERR_UNSUPPORTED (only sphere, cube and plane hits carry a tangent
     * frame), a height map with height_params[2] <= 0 layers with
     * PT_ERR_ARG.  pt_update_objects keeps the uploaded texels. */
    const uint8_t* textures;
    uint32_t texture_layers;
    uint32_t flags;         /* PT_HAS_* / PT_MIS_SKYBOX */
    int32_t bounce_limit;   /* override constant BOUNCE_LIMIT */
    uint32_t _pad;
} pt_scene_desc;

/* Per-call counters.  Rays are counted the way the reference's loops issue
 * them: an extension ray is one trace_scene call, a shadow ray is one
 * trace_scene_shadow call. */
typedef struct pt_stats {
    uint64_t extension_rays;
    uint64_t shadow_rays;
    uint64_t paths;
    uint64_t node_pops;     /* BVH nodes popped, TLAS + BLAS (SurfaceHit.count) */
    uint64_t inner_pops;    /* popped nodes that were inner nodes */
    uint64_t object_visits; /* TLAS-leaf objects intersected */
    uint64_t tri_tests;     /* ray-triangle tests */
    uint64_t shade_events;  /* surface hits that were shaded */
    double kernel_ms;       /* summed device time of the trace kernels (HIP events) */
    uint32_t launches;
    uint32_t _pad;
} pt_stats;

/* Closest-hit record for pt_trace_rays (SurfaceHit, shader.wgsl:117-123,
 * plus the identities the shader does not keep). */
typedef struct pt_hit {
    float t;            /* 1e10 on a miss */
    int32_t o_idx;      /* object index, -1 for planes / list meshes / miss */
    int32_t m_idx;      /* material index, -1 on a miss */
    int32_t prim_id;    /* absolute triangle index for meshes, else -1 */
    int32_t plane_idx;  /* plane index when the closest hit is a plane, else -1 */
    uint32_t count;     /* BVH nodes popped */
    float n[3];
    float uv[2];
    float _pad[3];
} pt_hit;
PT_LAYOUT_ASSERT(sizeof(pt_hit) == 56, "pt_hit is 56 B");

/* Options (pt_set_option). */
enum {
    PT_OPT_BACKEND = 1,     /* PT_BACKEND_* (default WAVEFRONT) */
    PT_OPT_COUNT_STATS = 2, /* 1: count node/triangle work (default 1) */
    PT_OPT_WAVEFRONT_POOL = 3, /* paths per wavefront batch, whole passes (default 1<<27, ~312 B each) */
    PT_OPT_SCHEDULE = 4,       /* megakernel pass schedule, PT_SCHEDULE_* (default REGEN) */
    PT_OPT_TRAVERSAL = 5,      /* BVH traversal, PT_TRAVERSAL_* (default WIDE) */
    PT_OPT_LDS_STACK = 6,      /* traversal-stack entries per lane kept in LDS, the rest in device
                                  memory; 0 = all in LDS (default 16) */
    PT_OPT_WAVEFRONT_REFILL = 7, /* wavefront traversal: idle lanes of a wave that trigger fetching
                                   new rays, 1..64 (default 56) */
    PT_OPT_BLAS_BUILD = 8,       /* PT_BLAS_BUILD_*: where the wide traversal's bottom-level trees
                                   come from; read by the next pt_upload_scene */
    PT_OPT_TLAS_BUILD = 9,       /* PT_TLAS_BUILD_*: where the top-level tree comes from; read by
                                   every pt_upload_scene and pt_update_objects */
    PT_OPT_DENOISE_MATH = 12,    /* PT_DENOISE_MATH_*: how pt_denoise evaluates the a-trous weights */
    PT_OPT_BOUNCE_TIMING = 13,   /* 1: the wavefront backend records an event before each bounce's
                                   launches and waits after each batch, so that pt_get_bounce_stats
                                   reports device time per bounce; for measurement (default 0) */
    /* read-only (pt_get_option), describing the uploaded scene: */
    PT_OPT_KERNEL_FLAGS = 10,    /* the flag word the trace kernels run with: desc->flags plus internal
                                   bits (bit 20 material textures, bit 21 albedo alpha below 1) */
    PT_OPT_SPECIALISED = 11      /* 1 when a trace-kernel variant is compiled for that flag word, 0 when
                                   the run-time-flags variant runs */
};
/* REFERENCE: the uploaded tracer.js BVH, widened on the host (default).
   DEVICE: rebuilt on the GPU from the triangles (linear BVH over Morton
   codes, 4-wide); the same closest hits, different node counts.  The
   child-pair (reference-order) traversal always uses the uploaded BVH. */
enum { PT_BLAS_BUILD_REFERENCE = 0, PT_BLAS_BUILD_DEVICE = 1 };
/* REFERENCE: the descriptor's TLAS (SceneBVHBuilder, tracer.js:1509-1619),
   converted on the host (default).  DEVICE: both forms of the top-level tree
   (child-pair and 4-wide) built on the GPU over the objects' world boxes,
   one object per leaf (a linear BVH over Morton codes), at every upload and
   update; the descriptor's `tlas` array is ignored and may be empty, so a
   caller that moves objects sends only the objects.  The same closest hits;
   node counts are of the device-built tree. */
enum { PT_TLAS_BUILD_REFERENCE = 0, PT_TLAS_BUILD_DEVICE = 1 };
enum { PT_BACKEND_MEGAKERNEL = 0, PT_BACKEND_WAVEFRONT = 1 };
/* FUSED (default): the four Gaussian weights of a tap summed into one base-2
   exponent with precomputed reciprocals, and pow(dot(n, n'), sigma_normal)
   joined to it as sigma_normal * log2: one hardware exp2 and log2 per tap,
   within 1e-5 relative of EXACT on the output.  EXACT: each weight with the
   library exp / pow and a division, as denoiser.wgsl:117-189 writes it. */
enum { PT_DENOISE_MATH_FUSED = 0, PT_DENOISE_MATH_EXACT = 1 };
/* SYNC: the lanes of a wave run each pass together.  REGEN: a lane whose path
   ends starts its pixel's next pass at once.  Both produce the same bits. */
enum { PT_SCHEDULE_SYNC = 0, PT_SCHEDULE_REGEN = 1 };
/* REFERENCE: the reference's child-pair tree in its visit order; hits and the
   per-hit node counts match it bit for bit.  WIDE: the same boxes regrouped
   into 4-wide nodes, nearest child first; the same closest hit except where
   two candidates tie within rounding (then the visit order picks one), and
   node counts are of wide nodes.  WIDE_QUANTIZED: the same 4-wide trees
   with each child box stored as 8-bit offsets from the node's origin in
   power-of-two steps, rounded outward (64 B nodes instead of 128 B); every
   decoded box contains the exact one, so the same closest hits as WIDE up to
   ties.  Work counters (PT_OPT_COUNT_STATS) and pt_trace_hits always use
   REFERENCE. */
enum { PT_TRAVERSAL_REFERENCE = 0, PT_TRAVERSAL_WIDE = 1, PT_TRAVERSAL_WIDE_QUANTIZED = 2 };

/* Lifetime: `new Renderer` + Renderer.init (tracer.js:1719-1770). */
int pt_create(pt_ctx** out, int device);
void pt_destroy(pt_ctx* ctx);
const char* pt_last_error(const pt_ctx* ctx);
int pt_set_option(pt_ctx* ctx, int key, int64_t value);
/* The current value of an option, or of a read-only PT_OPT_KERNEL_FLAGS /
 * PT_OPT_SPECIALISED query (PT_ERR_STATE before the first upload). */
int pt_get_option(pt_ctx* ctx, int key, int64_t* value);

/* Per-bounce work of the last pt_trace call on the wavefront backend, for
 * bounces 0 .. n-1 (n <= 16): extension rays traced at bounce b (the camera
 * rays at b = 0), shadow rays traced for the NEE samples of bounce b's
 * shading, and with PT_OPT_BOUNCE_TIMING the device time of bounce b's three
 * launches (extension, shading, shadow) in ms, summed over the call's
 * batches (0 without it).  The megakernel reports zeros.  No counterpart in
 * the reference: a measurement entry point. */
int pt_get_bounce_stats(pt_ctx* ctx, uint64_t* ext_rays, uint64_t* shadow_rays, double* ms, uint32_t n);

/* Renderer.setScene (tracer.js:1955-2102): copy all scene arrays, the sky and
 * its CDF to the device and specialise the trace kernels on desc->flags. */
int pt_upload_scene(pt_ctx* ctx, const pt_scene_desc* desc);

/* Renderer.updateObjects fast path (tracer.js:2132-2164): rewrite objects,
 * TLAS, planes, list meshes and materials in place.  Returns PT_ERR_CAPACITY
 * when an array outgrew the last pt_upload_scene's (earlier updates that
 * shrank it do not lower the capacity) and PT_ERR_FLAGS when the flags
 * changed; the caller then calls pt_upload_scene, as the reference falls back
 * to a full setScene.  As in the reference, lights, triangles, BVH nodes, the
 * sky and the texels stay those of the upload: a caller that moves or removes
 * an emitter re-uploads. */
int pt_update_objects(pt_ctx* ctx, const pt_scene_desc* desc);

/* Renderer.resize (tracer.js:2166-2205): (re)allocate the fp32 RGBA
 * accumulation buffer (16 B/px), the rgba8 output and, if asked, the
 * 48 B/px G-buffer; clears the accumulation buffer. */
int pt_resize(pt_ctx* ctx, uint32_t width, uint32_t height, int with_gbuffer);

/* Renderer.renderSection + updateUniforms (tracer.js:2207-2309), n_passes
 * times: pass k runs shader.wgsl `main` over params->section ..
 * section+(section_w, section_h) (0,0 = whole frame) with sample_number =
 * params->sample_number + k and seed = seeds[k] (params->seed when seeds is
 * NULL).  Each pass sees its own sample number and seed — the reference's
 * tiled loop instead reuses the last uniforms for every rep in a tile.
 * stats may be NULL. */
int pt_trace(pt_ctx* ctx, const pt_scene_params* params, uint32_t section_w,
             uint32_t section_h, uint32_t n_passes, const uint32_t* seeds,
             pt_stats* stats);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): pt_trace with the same arguments and results, calling no
 * hipMalloc, hipFree or allocating routine.  It always runs the persistent
 * megakernel (bit-identical to either backend) with its scratch — work
 * counters, tile counters, the seed list, per-lane medium stacks and
 * traversal-stack spill — in a static 64 MiB device array per GPU, held until
 * the process ends; the persistent grid shrinks to what the array holds.  The
 * scene and frame must already be uploaded and sized (pt_upload_scene,
 * pt_resize).  At most 65536 passes per call (PT_ERR_CAPACITY beyond).  Calls
 * on one GPU share the array: calls that overlap on one GPU race. */
int pt_trace_noalloc(pt_ctx* ctx, const pt_scene_params* params, uint32_t section_w, uint32_t section_h,
                     uint32_t n_passes, const uint32_t* seeds, pt_stats* stats);

/* gbuf_main (shader.wgsl:2510-2603) over the whole frame: fills the G-buffer
 * and writes the four debug strips to the rgba8 output. */
int pt_gbuffer(pt_ctx* ctx, const pt_scene_params* params);

/* Denoiser.execute (tracer.js:2583-2637): prepass, `passes` à-trous passes,
 * composite into the rgba8 output.  stride_mode 0: pass i uses step 1<<i
 * (the intended schedule); 1: every pass uses 1<<(passes-1), which is what
 * the reference's single-submit uniform writes produce.  The accumulation
 * buffer and G-buffer are read, not modified. */
int pt_denoise(pt_ctx* ctx, const pt_denoise_params* params, int stride_mode);

/* Readback / restore (gui.js:1627-1664 copies the texture out). */
int pt_read_accum(pt_ctx* ctx, float* dst, uint64_t bytes);
int pt_write_accum(pt_ctx* ctx, const float* src, uint64_t bytes);
int pt_read_output(pt_ctx* ctx, uint8_t* dst, uint64_t bytes);
int pt_read_gbuffer(pt_ctx* ctx, pt_gbuffer_texel* dst, uint64_t bytes);
int pt_write_gbuffer(pt_ctx* ctx, const pt_gbuffer_texel* src, uint64_t bytes);
/* Reads the denoiser's final fp32 RGBA buffer (before the rgba8 store). */
int pt_read_denoised(pt_ctx* ctx, float* dst, uint64_t bytes);
/* Reads the device copy of the sky CDF ((H+1)+H*(W+1) floats) and the total
 * weight it was normalised by. */
int pt_read_sky_cdf(pt_ctx* ctx, float* dst, uint64_t count, float* total_lum);

/* Deterministic hit pass: one sample of shader.wgsl `main` per pixel with
 * params' sample_number and seed, recording each of the first n_bounces
 * trace_scene hits of every path in hits[pixel * n_bounces + b] (t = -1 and
 * ids -1 after the path has ended).  Accumulation and output are untouched.
 * Primary (b = 0) and secondary (b = 1) hits are what the parity tests
 * compare. */
int pt_trace_hits(pt_ctx* ctx, const pt_scene_params* params, uint32_t n_bounces, pt_hit* hits);

/* pt_trace_hits that also returns the ray each recorded trace_scene call
 * traced: rays[(pixel * n_bounces + b) * 6 ..] = origin.xyz, direction.xyz
 * (zeros after the path has ended; rays may be NULL).  The parity tests trace
 * the device's own secondary rays on the oracle, so a hit comparison does not
 * depend on how the two sides round sin/cos/pow while sampling the BSDF. */
int pt_trace_hits_rays(pt_ctx* ctx, const pt_scene_params* params, uint32_t n_bounces, pt_hit* hits, float* rays);

/* Closest hit (trace_scene, shader.wgsl:754-774) for n rays given as
 * origin.xyz, dir.xyz pairs (6 floats per ray). */
int pt_trace_rays(pt_ctx* ctx, const float* rays, uint32_t n, pt_hit* hits);

/* Shadow transmittance (trace_scene_shadow, shader.wgsl:1206-1229) for n rays
 * (6 floats each) with per-ray target object and distance; writes 3 floats
 * per ray. */
int pt_trace_shadow_rays(pt_ctx* ctx, const float* rays, const int32_t* target_idx,
                         const float* target_dist, uint32_t n, float* transmittance);

/* Blocks until the context's stream is idle. */
int pt_synchronize(pt_ctx* ctx);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* PT_ABI_H */
