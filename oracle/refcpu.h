/*
 * refcpu.h — the oracle: a CPU restatement of the reference's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (libpt, libptscene, the
 * Python package, the Node addon) links, loads or calls this code; only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark do, and
 * only as the checker or as the measured CPU baseline.
 *
 * What it restates (file:line into the reference):
 *   ref_trace_rays        trace_scene            shader.wgsl:754-774 (+ 473-752)
 *   ref_trace_shadow_rays trace_scene_shadow     shader.wgsl:1206-1229 (+ 777-1184)
 *   ref_render            main                   shader.wgsl:2273-2492
 *   ref_trace_hits        main's first bounces, recording each trace_scene hit
 *   ref_gbuffer           gbuf_main              shader.wgsl:2510-2603
 *   ref_denoise           prepass/atrous/composite denoiser.wgsl:55-214, pass
 *                         schedule Denoiser.execute tracer.js:2583-2637
 *   ref_build_sky_cdf     HDRTexture.buildHDRCDF tracer.js:135-204
 *   ref64_closest         brute-force closest hit in double precision over every
 *                         plane, object and triangle (no BVH): the FP64 arbiter
 *
 * Arithmetic rules (the WGSL leaves these to the implementation; the oracle
 * fixes them and the HIP kernels follow the same rules):
 *   - f32 everywhere, no contraction (compiled with -ffp-contract=off), each
 *     WGSL expression evaluated left to right;
 *   - dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z; normalize(v) = v / length(v);
 *     mat*vec sums the column products in column order;
 *   - min/max return the non-NaN operand (IEEE minNum/maxNum);
 *   - pow, exp, log, sin, cos, atan2, asin, acos, tan are the C library's
 *     (pow of a negative base with an integral exponent is the signed power);
 *   - textureSampleLevel on the sky is bilinear with fp32 weights, texel
 *     centres at (i+0.5)/size, clamp-to-edge, f16 texels widened exactly;
 *   - rgba8unorm stores round(clamp(x,0,1)*255) to nearest, NaN -> 0.
 */
#ifndef PT_REFCPU_H
#define PT_REFCPU_H

#include "pt_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ref_scene ref_scene;

int ref_scene_create(ref_scene** out, const pt_scene_desc* desc);
void ref_scene_destroy(ref_scene* s);

int ref_trace_rays(const ref_scene* s, const float* rays, uint32_t n, pt_hit* hits, pt_stats* stats);
int ref_trace_shadow_rays(const ref_scene* s, const float* rays, const int32_t* target_idx,
                          const float* target_dist, uint32_t n, float* transmittance);

/* n_passes progressive samples into accum (width*height*4 floats, in/out);
 * out_rgba8 may be NULL.  rows/row_stride restrict the run to rows
 * row0, row0+row_stride, ... (row_stride 1 = all rows of the section).
 * threads <= 0 uses every hardware thread. */
int ref_render(const ref_scene* s, const pt_scene_params* params, uint32_t section_w,
               uint32_t section_h, uint32_t n_passes, const uint32_t* seeds, float* accum,
               uint8_t* out_rgba8, pt_stats* stats, int threads, uint32_t row_stride);

/* One sample per pixel (sample_number/seed from params); records the hit of
 * each of the first n_bounces trace_scene calls per pixel into
 * hits[pixel * n_bounces + b] (t = -1, ids -1 when the path ended earlier). */
int ref_trace_hits(const ref_scene* s, const pt_scene_params* params, uint32_t n_bounces,
                   pt_hit* hits, int threads);
/* ref_trace_hits plus each recorded call's ray (6 floats, zeros after the
 * path ended) in rays[(pixel * n_bounces + b) * 6 ..]; rays may be NULL. */
int ref_trace_hits_rays(const ref_scene* s, const pt_scene_params* params, uint32_t n_bounces, pt_hit* hits,
                        float* rays, int threads);

int ref_gbuffer(const ref_scene* s, const pt_scene_params* params, pt_gbuffer_texel* gbuf,
                uint8_t* out_rgba8, int threads);

int ref_denoise(const pt_denoise_params* params, int stride_mode, const float* accum,
                const pt_gbuffer_texel* gbuf, float* out_rgba, uint8_t* out_rgba8);

int ref_build_sky_cdf(const uint16_t* rgba16f, uint32_t width, uint32_t height, float exposure,
                      float* cdf, double* total_weight);

/* sample_surface (shader.wgsl:2052-2175) n times at a surface point at the
 * origin with shading normal n, viewer direction v (pointing away from the
 * surface), material m, outside any medium.  Sample i draws from PCG state
 * (seed + i*7919)|1.  out[8*i..]: L.xyz, weight.xyz (f/pdf), pdf, status
 * (1: continued along L, 0: the path ended). */
int ref_bsdf_sample(const pt_material* m, const float n[3], const float v[3], uint32_t n_samples, uint32_t seed,
                    float* out);
/* eval_surface / pdf_surface for k directions L (3 floats each). */
int ref_bsdf_eval(const pt_material* m, const float n[3], const float v[3], const float* L, uint32_t k, float* f,
                  float* pdf);

/* sample_env_cdf (shader.wgsl:231-254) n times with uniforms (u.x, u.y) from
 * PCG state (seed + i*7919)|1: out[7*i..] = dir.xyz, pdf, color.rgb.  And
 * get_sky_pdf (shader.wgsl:208-229) for k directions. */
int ref_env_sample(const ref_scene* s, const pt_scene_params* params, uint32_t n, uint32_t seed, float* out);
int ref_env_pdf(const ref_scene* s, const pt_scene_params* params, const float* dirs, uint32_t k, float* pdf);

/* rand_pcg (shader.wgsl:146-154) n times from `state`: the output words
 * (before the float conversion) and the f32 values. */
int ref_pcg(uint32_t state, uint32_t n, uint32_t* words, float* values);

/* The camera ray of shader.wgsl `main` (2280-2287: seeded stream, warm-up
 * draw, jitter) for n pixels (x, y pairs), or with centre != 0 the pixel-centre
 * ray of gbuf_main (2517-2522).  rays: 6 floats each (origin, direction);
 * jitter (may be NULL): the 2 jitter values. */
int ref_camera_rays(const pt_scene_params* params, uint32_t n, const uint32_t* pixels, int centre, float* rays,
                    float* jitter);

/* BSDF building blocks on n argument triples:
 *   0 D_GGX(NdotH, alpha2)                       shader.wgsl:1700-1704
 *   1 fresnel_dielectric(cos, etai, etat)        :1721-1733
 *   2 G_Smith_GGX_div_NV(NdotV, NdotL, alpha2)   :1787-1791
 *   3 G_Smith_GGX_Uncorrelated_div_NV(..)        :1794-1798
 *   4 G1_GGX_div_NV(NdotV, alpha2)               :1801-1803
 *   5 mis_weight(pdf_a, pdf_b)                   :1648-1659 */
int ref_bsdf_terms(int term, const float* args, uint32_t n, float* out);

/* sample_light (shader.wgsl:1636-1646, sphere :1501-1549, box :1575-1624)
 * for light `light` seen from pos, n times, sample i from PCG state
 * (seed + i*7919)|1: out[8*i..] = dir.xyz, dist, pdf, color.rgb.  And the
 * pdf the renderer gives a BSDF direction that reaches the light
 * (get_light_pdf :1626-1634, 0 when the ray's closest hit is not the light). */
int ref_light_sample(const ref_scene* s, uint32_t light, const float pos[3], uint32_t n, uint32_t seed, float* out);
int ref_light_pdf(const ref_scene* s, uint32_t light, const float pos[3], const float* dirs, uint32_t k, float* pdf);

/* rays: 6 doubles; fills *hit with the double-precision closest hit (t in
 * hit->t rounded to float, exact value in *t_exact). */
int ref64_closest(const ref_scene* s, const double* ray, pt_hit* hit, double* t_exact);

#ifdef __cplusplus
}
#endif
#endif
