// pt_kernels.hip — gfx950 kernels of the tracer hot path.
//
//   trace_kernel    shader.wgsl `main` (2273-2492): one lane per pixel, all of
//                   a call's progressive passes in one launch (the running
//                   mean stays in registers; accum is read and written once)
//   gbuffer_kernel  shader.wgsl `gbuf_main` (2510-2603)
//   rays_kernel     trace_scene / trace_scene_shadow on caller rays
//   dn_*            denoiser.wgsl prepass / atrous / composite (55-214)
//   cdf_*           HDRTexture.buildHDRCDF (tracer.js:135-204) as a parallel
//                   double-precision scan, one workgroup per row
//
// Work mapping: a 64-lane workgroup (one wave64) shades an 8x8 pixel tile.
// Each of the 8 XCDs works on a contiguous band of tiles (the dispatcher
// deals workgroups round-robin over XCDs), which keeps neighbouring tiles —
// and the BVH nodes they touch — in one XCD's L2: trace_kernel is persistent
// and pulls tiles from per-XCD band counters, the other kernels remap
// workgroup ids.
#include "pt_internal.h"
#include "dev/path.h"

namespace pt {

constexpr int kTile = 8;
constexpr int kWave = 64;


__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb) {
    const uint32_t q = nb / 8u, r = nb % 8u;
    const uint32_t xcd = b % 8u, slot = b / 8u;
    const uint32_t base = xcd < r ? xcd * (q + 1u) : r * (q + 1u) + (xcd - r) * q;
    return base + slot;
}

__device__ __forceinline__ unsigned long long wave_sum(uint32_t v) {
    unsigned long long s = v;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    return s;
}

__device__ void flush_counters(const Counters& c, unsigned long long* stats, bool nodes) {
    const uint32_t lane = threadIdx.x & 63u;
    const unsigned long long e = wave_sum(c.ext), s = wave_sum(c.shadow), p = wave_sum(c.paths);
    if (lane == 0) {
        atomicAdd(stats + kStExt, e);
        atomicAdd(stats + kStShadow, s);
        atomicAdd(stats + kStPaths, p);
    }
    if (nodes) {
        const unsigned long long a = wave_sum(c.pops), b = wave_sum(c.inner), o = wave_sum(c.objs),
                                 t = wave_sum(c.tris), h = wave_sum(c.shade);
        if (lane == 0) {
            atomicAdd(stats + kStPops, a);
            atomicAdd(stats + kStInner, b);
            atomicAdd(stats + kStObjs, o);
            atomicAdd(stats + kStTris, t);
            atomicAdd(stats + kStShade, h);
        }
    }
}

__device__ void store_hit(pt_hit* o, const HitRec& h, const SurfHit& s) {
    pt_hit r;
    r.t = h.t;
    r.o_idx = s.o_idx;
    r.m_idx = s.m_idx;
    r.prim_id = h.prim;
    r.plane_idx = h.kind == kPlaneHit ? h.idx : -1;
    r.count = h.count;
    r.n[0] = s.n.x; r.n[1] = s.n.y; r.n[2] = s.n.z;
    r.uv[0] = s.uv.x; r.uv[1] = s.uv.y;
    r._pad[0] = r._pad[1] = r._pad[2] = 0.0f;
    *o = r;
}

// Megakernel NEE sink: traces the shadow ray at once and adds the
// contribution, exactly where evaluate_nee's result is added
// (shader.wgsl:2440, 2452).
template <uint32_t F, bool CN, int W>
struct ImmediateNee {
    const DevScene& S;
    Flags<F> fl;
    LdsStack& st;
    Counters& c;
    __device__ void operator()(int /*k*/, bool need_ray, v3 o, v3 d, float dist, int ignore, v3 inc, v3 tp, int bounce,
                               v3& rad) {
        if (need_ray) {
            c.shadow++;
            if (fl.has(kTranslucentBit)) {  // sample_color *= shadow_hit (shader.wgsl:2262-2267)
                const float tr = transmittance<F, CN, W>(S, fl, o, d, ignore, dist, st, c);
                inc = tr < 0.001f ? splat(0.0f) : inc * tr;
            } else if (occluded<F, CN, W>(S, fl, o, d, ignore, dist, st, c)) {
                inc = splat(0.0f);
            }
        }
        rad = rad + tp * firefly(inc, bounce);
    }
    __device__ void emit(v3 add, v3& rad) { rad = rad + add; }
};

// One path of shader.wgsl `main` (2279-2480).  rec: when non-null, the first
// n_rec trace_scene hits of the path are written there, and when rec_rays is
// non-null the ray each of them traced (origin, direction).
template <uint32_t F, bool CN, bool REC, int W>
__device__ v3 trace_path(const DevScene& S, Flags<F> fl, const pt_scene_params& PP, const Params& P, uint32_t px,
                         uint32_t py, uint32_t sample_number, uint32_t seed, LdsStack& st, float* media,
                         uint32_t media_stride, Counters& c, pt_hit* rec, float* rec_rays, uint32_t n_rec) {
    PathState ps;
    path_begin(PP, px, py, sample_number, seed, ps);
    Media md{media, media_stride, 0};
    ImmediateNee<F, CN, W> sink{S, fl, st, c};
    c.paths++;
    for (int bounce = 0; bounce < S.bounce_limit; ++bounce) {
        HitRec h;
        c.ext++;
        if (REC && rec_rays && uint32_t(bounce) < n_rec) {
            float* r = rec_rays + size_t(bounce) * 6u;
            r[0] = ps.ro.x; r[1] = ps.ro.y; r[2] = ps.ro.z;
            r[3] = ps.rd.x; r[4] = ps.rd.y; r[5] = ps.rd.z;
        }
        closest_hit<F, CN, W>(S, fl, ps.ro, ps.rd, st, h, c);
        const SurfHit sh = resolve_hit<REC>(S, ps.ro, ps.rd, h);
        if (REC && uint32_t(bounce) < n_rec) store_hit(rec + bounce, h, sh);
        if (!shade_bounce<F, CN>(S, fl, P, bounce, h, sh, ps, md, c, sink)) break;
    }
    return ps.rad;
}

// Pass schedules of trace_kernel.
//   kSync   every lane of a wave runs pass p, then pass p+1: a wave waits
//           for its longest path in every pass.
//   kRegen  a lane whose path ends starts its pixel's next pass at once, so
//           a wave only waits once, for the lane with the longest total.
//   kRecord kSync plus per-bounce hit recording.
// Each pixel still runs its passes in order and folds them into the running
// mean in order, so every schedule writes the same bits.
enum { kSync = 0, kRegen = 1, kRecord = 2 };
static_assert(kMediaFloats == int(kLaneMediaFloats), "medium stack size");

// Next 8x8 tile for a persistent trace workgroup.  The section's tiles are
// split into 8 contiguous row-major bands, one per XCD (workgroups are dealt
// round-robin over the XCDs, so workgroup b runs on XCD b % 8); a workgroup
// drains its own band first, keeping neighbouring tiles and the BVH nodes
// they share in one L2, then helps the other bands.  Called by all lanes;
// lane 0 does the atomics.  Returns n_tiles when every band is drained.
__device__ uint32_t next_tile(uint32_t* ctr, uint32_t n_tiles) {
    uint32_t t = n_tiles;
    if (threadIdx.x == 0) {
        const uint32_t home = blockIdx.x & 7u;
        for (uint32_t k = 0; k < 8u; ++k) {
            const uint32_t b = (home + k) & 7u;
            const uint32_t lo = uint32_t(uint64_t(n_tiles) * b / 8u);
            const uint32_t hi = uint32_t(uint64_t(n_tiles) * (b + 1u) / 8u);
            if (lo == hi || __hip_atomic_load(ctr + b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= hi - lo)
                continue;
            const uint32_t i = atomicAdd(ctr + b, 1u);
            if (i < hi - lo) {
                t = lo + i;
                break;
            }
        }
    }
    return __shfl(t, 0, kWave);
}

// Persistent: the grid holds as many workgroups as are resident at once and
// each loops over tiles from next_tile.  Lanes keep their medium stacks in
// A.media (structure of arrays, one column per lane of the grid) and the
// first A.stack_lds entries of their traversal stacks in LDS, the rest in
// A.stack_spill (same column layout).
//
// Occupancy: the kernel needs ~150-165 VGPRs, which allows 3 waves per SIMD.
// Capping it at 96 VGPRs (5 waves per SIMD) spills some values of the
// shading code but hides more of the traversal's dependent-load latency:
// on the 808k-triangle scene 4 waves/SIMD run 4% faster than 3, and 5 another
// 5% faster than 4.
#ifndef PT_TRACE_WAVES_PER_EU
#define PT_TRACE_WAVES_PER_EU 5
#endif
template <uint32_t F, bool CN, int MODE, int W>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(PT_TRACE_WAVES_PER_EU)))
trace_kernel(DevScene S, TraceArgs A) {
    extern __shared__ uint32_t lds_stack[];
    const Flags<F> fl{S.flags};
    const uint32_t tiles_x = (A.sec_w + kTile - 1) / kTile;
    const uint32_t lane = threadIdx.x;
    const uint32_t gid = blockIdx.x * kWave + lane;
    LdsStack st{lds_column(lds_stack + lane), kWave, A.stack_lds, 0, A.stack_spill + gid, A.media_stride,
                W ? S.wstack_cap : S.stack_cap};
    float* media = A.media + gid;
    const Params P{A.P.sky_total_lum, A.P.sky_width, A.P.sky_height, A.P.sky_cdf_index};
    Counters c;
    c.clear();
    for (uint32_t tile = next_tile(A.tile_ctr, A.n_tiles); tile < A.n_tiles;
         tile = next_tile(A.tile_ctr, A.n_tiles)) {
        const uint32_t px = A.P.section[0] + (tile % tiles_x) * kTile + (lane & 7u);
        const uint32_t py = A.P.section[1] + (tile / tiles_x) * kTile + (lane >> 3);
        const bool active = px < A.P.section[0] + A.sec_w && py < A.P.section[1] + A.sec_h && px < A.P.width &&
                            py < A.P.height;
        if (active) {
            const uint32_t idx = py * A.P.width + px;
            const float4 old = A.accum[idx];
            v3 acc = mk(old.x, old.y, old.z);
            if (MODE == kRegen) {
                ImmediateNee<F, CN, W> sink{S, fl, st, c};
                Media md{media, A.media_stride, 0};
                PathState ps;
                uint32_t pass = 0;
                int bounce = 0;
                path_begin(A.P, px, py, A.P.sample_number, A.seeds ? A.seeds[0] : A.P.seed, ps);
                c.paths++;
                for (;;) {
                    if (bounce < S.bounce_limit) {
                        HitRec h;
                        c.ext++;
                        closest_hit<F, CN, W>(S, fl, ps.ro, ps.rd, st, h, c);
                        const SurfHit sh = resolve_for_shading<F>(S, fl, ps.ro, ps.rd, h);
                        const bool more = shade_bounce<F, CN>(S, fl, P, bounce, h, sh, ps, md, c, sink);
                        ++bounce;
                        if (more && bounce < S.bounce_limit) continue;
                    }
                    const uint32_t sn = A.P.sample_number + pass;
                    acc = lerp(acc, ps.rad, 1.0f / float(sn + 1u));
                    if (++pass == A.n_passes) break;
                    path_begin(A.P, px, py, sn + 1u, A.seeds ? A.seeds[pass] : A.P.seed, ps);
                    md.n = 0;
                    bounce = 0;
                    c.paths++;
                }
            } else {
                for (uint32_t pass = 0; pass < A.n_passes; ++pass) {
                    const uint32_t sn = A.P.sample_number + pass;
                    const uint32_t seed = A.seeds ? A.seeds[pass] : A.P.seed;
                    pt_hit* rec = MODE == kRecord ? A.hits + size_t(idx) * A.n_rec : nullptr;
                    float* rec_rays = MODE == kRecord && A.rec_rays ? A.rec_rays + size_t(idx) * A.n_rec * 6u : nullptr;
                    const v3 r = trace_path<F, CN, MODE == kRecord, W>(S, fl, A.P, P, px, py, sn, seed, st, media,
                                                                    A.media_stride, c, rec, rec_rays, A.n_rec);
                    acc = lerp(acc, r, 1.0f / float(sn + 1u));
                }
            }
            if (MODE != kRecord) {
                A.accum[idx] = make_float4(acc.x, acc.y, acc.z, 1.0f);
                v3 o = acc * splat(A.P.exposure);
                o = o / (o + splat(0.3f));
                A.out[idx] = make_uchar4(unorm8(powf(o.x, 0.4545f)), unorm8(powf(o.y, 0.4545f)),
                                         unorm8(powf(o.z, 0.4545f)), 255);
            }
        }
    }
    if (A.stats) flush_counters(c, A.stats, CN);
}

// gbuf_main, shader.wgsl:2510-2603
template <uint32_t F>
__global__ void __launch_bounds__(kWave) gbuffer_kernel(DevScene S, pt_scene_params P, pt_gbuffer_texel* gbuf,
                                                        uchar4* out) {
    extern __shared__ uint32_t lds_stack[];
    const Flags<F> fl{S.flags};
    const uint32_t tiles_x = (P.width + kTile - 1) / kTile;
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t lane = threadIdx.x;
    const uint32_t px = (tile % tiles_x) * kTile + (lane & 7u);
    const uint32_t py = (tile / tiles_x) * kTile + (lane >> 3);
    if (px >= P.width || py >= P.height) return;
    LdsStack st = lds_stack_of(lds_stack + lane, blockDim.x, S.stack_cap);
    Counters c;
    c.clear();
    const uint32_t idx = py * P.width + px;
    v3 rd = camera_dir(P, (float(px) + 0.5f) / float(P.width), (float(py) + 0.5f) / float(P.height));
    v3 ro = ld3(P.eye);
    v3 n = splat(0.0f), albedo = splat(0.0f), emit = splat(0.0f), tint = splat(1.0f);
    float rough = 0.0f, metal = 0.0f, trans = 0.0f, dist = 0.0f;
    for (int bounce = 0; bounce < S.bounce_limit; ++bounce) {
        HitRec h;
        closest_hit<F, false, kTravPairs>(S, fl, ro, rd, st, h, c);
        const SurfHit sh = resolve_for_shading<F>(S, fl, ro, rd, h);
        if (sh.m_idx == -1) {
            n = splat(0.0f);
            emit = sample_sky(S, rd);
            albedo = splat(0.0f);
            break;
        }
        const pt_material& m = S.mats[sh.m_idx];
        trans = m.transmission;
        albedo = ld3(m.color);
        emit = ld3(m.emittance);
        rough = m.roughness;
        metal = m.metallic;
        v3 sn = sh.n;
        float alpha = 1.0f;
        if (fl.has(kTexturesBit)) {  // get_surface_context's maps, no parallax
            Surface sc = base_surface(m, sh.n);
            apply_textures(S, m, v2{sh.uv.x * m.uv_scale[0], sh.uv.y * m.uv_scale[1]}, sh.tg, sh.bt, sc);
            albedo = sc.albedo;
            emit = sc.emittance;
            rough = sc.roughness;
            metal = sc.metallic;
            sn = sc.normal;
            alpha = sc.alpha;
        }
        n = dot(sn, -rd) > 0.0f ? sn : -sn;
        dist += h.t;
        if (alpha < 1.0f) {  // seen through (shader.wgsl:2553-2557)
            tint = tint * (splat(1.0f) * (1.0f - alpha) + albedo * alpha);
            ro = ro + rd * h.t - n * 0.001f;
            continue;
        }
        if (m.metallic >= 0.99f && m.roughness <= 0.01f) {
            tint = tint * albedo;
            ro = ro + rd * h.t + n * 0.001f;
            rd = reflect(rd, n);
            continue;
        }
        break;
    }
    const v3 color = (albedo + emit) * tint;
    pt_gbuffer_texel g;
    g.normal[0] = n.x; g.normal[1] = n.y; g.normal[2] = n.z;
    g.roughness = rough;
    g.albedo[0] = color.x; g.albedo[1] = color.y; g.albedo[2] = color.z;
    g.metallic = metal;
    g.motion[0] = g.motion[1] = 0.0f;
    g.transmission = trans;
    g.dist = dist;
    gbuf[idx] = g;
    v3 o;
    if (px < P.width * 1u / 4u) {
        o = splat(dist / (dist + 3.0f));
    } else if (px < P.width * 2u / 4u) {
        o = n * 0.5f + splat(0.5f);
    } else if (px < P.width * 3u / 4u) {
        v3 t = color * splat(P.exposure);
        t = t / (t + splat(0.3f));
        o = mk(powf(t.x, 0.4545f), powf(t.y, 0.4545f), powf(t.z, 0.4545f));
    } else {
        o = mk(trans, rough, metal);
    }
    out[idx] = make_uchar4(unorm8(o.x), unorm8(o.y), unorm8(o.z), 255);
}

template <uint32_t F, int W>
__global__ void __launch_bounds__(kWave) rays_kernel(DevScene S, RayArgs A, int shadow) {
    extern __shared__ uint32_t lds_stack[];
    const Flags<F> fl{S.flags};
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    LdsStack st = lds_stack_of(lds_stack + threadIdx.x, blockDim.x, W ? S.wstack_cap : S.stack_cap);
    Counters c;
    c.clear();
    const float* r = A.rays + size_t(i) * 6;
    const v3 o = mk(r[0], r[1], r[2]), d = mk(r[3], r[4], r[5]);
    if (shadow) {
        const float v = transmittance<F, false, W>(S, fl, o, d, A.target[i], A.tdist[i], st, c);
        A.trans[3 * i] = v; A.trans[3 * i + 1] = v; A.trans[3 * i + 2] = v;
    } else {
        HitRec h;
        closest_hit<F, false, W>(S, fl, o, d, st, h, c);
        store_hit(A.hits + i, h, resolve_hit<true>(S, o, d, h));
    }
}

// ---------------------------------------------------------------- denoiser --
__device__ __forceinline__ v3 dn_tonemap(v3 c, float e) {
    v3 t = c * e;
    t = t / (t + splat(0.3f));
    return mk(powf(t.x, 0.4545f), powf(t.y, 0.4545f), powf(t.z, 0.4545f));
}

// Per-pixel guide values the a-trous taps compare, computed once instead of
// once per tap, in three structure-of-arrays planes: gn = (normal,
// |normal| > 0.1), ga = (tonemap(albedo), dist/(dist+3)), gm = (transmission,
// roughness, metallic, 0).
__global__ void dn_guide(DenoiseArgs A) {
    const uint32_t x = blockIdx.x * 16 + threadIdx.x, y = blockIdx.y * 16 + threadIdx.y;
    if (x >= A.u.width || y >= A.u.height) return;
    const uint32_t i = y * A.u.width + x;
    const pt_gbuffer_texel g = A.gbuf[i];
    const v3 n = ld3(g.normal);
    const v3 ta = dn_tonemap(ld3(g.albedo), A.u.exposure);
    A.gn[i] = make_float4(n.x, n.y, n.z, len(n) > 0.1f ? 1.0f : 0.0f);
    A.ga[i] = make_float4(ta.x, ta.y, ta.z, g.dist / (g.dist + 3.0f));
    A.gm[i] = make_float4(g.transmission, g.roughness, g.metallic, 0.0f);
}

// prepass_main, denoiser.wgsl:55-112
__global__ void dn_prepass(DenoiseArgs A) {
    const int W = int(A.u.width), H = int(A.u.height);
    const int x = int(blockIdx.x * 16 + threadIdx.x), y = int(blockIdx.y * 16 + threadIdx.y);
    if (x >= W || y >= H) return;
    const uint32_t i = uint32_t(y * W + x);
    float4 cc;
    if (A.u.flags & 1u) {
        float4 col[9];
        float lum[9];
        int k = 0;
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                const int cx = min(max(x + dx, 0), W - 1), cy = min(max(y + dy, 0), H - 1);
                col[k] = A.accum[cy * W + cx];
                lum[k] = dot(mk(col[k].x, col[k].y, col[k].z), mk(0.299f, 0.587f, 0.114f));
                ++k;
            }
#pragma unroll
        for (int a = 0; a < 5; ++a)
#pragma unroll
            for (int b = a + 1; b < 9; ++b)
                if (lum[b] < lum[a]) {
                    const float tl = lum[a]; lum[a] = lum[b]; lum[b] = tl;
                    const float4 tc = col[a]; col[a] = col[b]; col[b] = tc;
                }
        cc = col[4];
    } else {
        cc = A.accum[i];
    }
    const bool valid = cc.x >= 0.0f && cc.y >= 0.0f && cc.z >= 0.0f && cc.x <= 10000.0f && cc.y <= 10000.0f &&
                       cc.z <= 10000.0f;
    if (!valid) { cc.x = 0.0f; cc.y = 0.0f; cc.z = 0.0f; }
    v3 mc = dn_tonemap(mk(cc.x, cc.y, cc.z), A.u.exposure);
    const float4 g0 = A.gn[i], g1 = A.ga[i];
    const bool sky = len(mk(g0.x, g0.y, g0.z)) < 0.1f;
    if ((A.u.flags & 2u) && !sky) {
        const v3 ma = vmax(mk(g1.x, g1.y, g1.z), splat(0.001f));
        mc = mc / ma;
        mc = mk(clampf(mc.x, 0.0f, 10.0f), clampf(mc.y, 0.0f, 10.0f), clampf(mc.z, 0.0f, 10.0f));
    }
    A.a[i] = make_float4(mc.x, mc.y, mc.z, cc.w);
}

// atrous_main, denoiser.wgsl:117-189.
//
// A pass with stride s visits the 5x5 taps (x + dx*s, y + dy*s): on the
// sub-lattice of pixels congruent to (x, y) modulo s they are a dense 5x5
// stencil.  So each workgroup takes a 32x8 tile of one such sub-lattice (one
// phase (x mod s, y mod s)), stages it with a 2-point halo (36x12 points,
// clamped to the image edge as the taps are) in LDS — colour and the three
// guide planes, 27.6 KB — and every tap reads LDS.  The phases of one image
// region are consecutive workgroups of one XCD, so the strided global loads
// of the tile fill share that XCD's L2.
//
// EXACT: the reference's weights factor by factor, with the library exp and
// pow (one correctly rounded division per Gaussian).  Otherwise the four
// Gaussians are summed into one base-2 exponent with precomputed reciprocals
// and the normal weight pow(d, sigma_n) joins it as sigma_n * log2(d), so a
// tap costs one hardware exp2 and one log2: within a few ulp of each factor,
// inside what WGSL allows its exp and pow.
constexpr int kDnTx = 32, kDnTy = 8, kDnHx = kDnTx + 4, kDnHy = kDnTy + 4, kDnPts = kDnHx * kDnHy;

template <bool EXACT>
__global__ void __launch_bounds__(256) dn_atrous(DenoiseArgs A, const float4* __restrict__ in,
                                                 float4* __restrict__ outb, int step, uint32_t tiles_x, uint32_t phx,
                                                 uint32_t nph) {
    __shared__ float4 s_col[kDnPts], s_n[kDnPts], s_a[kDnPts], s_m[kDnPts];
    const int W = int(A.u.width), H = int(A.u.height);
    const uint32_t L = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t phase = L % nph, tile = L / nph;
    const int ph_x = int(phase % phx), ph_y = int(phase / phx);
    const int lx0 = int(tile % tiles_x) * kDnTx - 2, ly0 = int(tile / tiles_x) * kDnTy - 2;
    for (int k = int(threadIdx.x); k < kDnPts; k += 256) {
        const int px = min(max(ph_x + step * (lx0 + k % kDnHx), 0), W - 1);
        const int py = min(max(ph_y + step * (ly0 + k / kDnHx), 0), H - 1);
        const uint32_t si = uint32_t(py * W + px);
        s_col[k] = in[si];
        s_n[k] = A.gn[si];
        s_a[k] = A.ga[si];
        s_m[k] = A.gm[si];
    }
    __syncthreads();
    const int li = int(threadIdx.x) % kDnTx, lj = int(threadIdx.x) / kDnTx;
    const int x = ph_x + step * (lx0 + 2 + li), y = ph_y + step * (ly0 + 2 + lj);
    if (x >= W || y >= H) return;
    const int c = (lj + 2) * kDnHx + li + 2;
    const float4 cc = s_col[c], cg0 = s_n[c], cg1 = s_a[c], cg2 = s_m[c];
    const v3 cn = mk(cg0.x, cg0.y, cg0.z);
    const float kern[5] = {0.0625f, 0.25f, 0.375f, 0.25f, 0.0625f};
    float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
    float sw = 0.0f;
    const float preserve = (1.0f - cg2.y) * mx(cg2.z, cg2.x);
    const float att = lerp(1.0f, 0.0f, preserve * A.u.glossy_protect);
    const float sc2 = A.u.sigma_color * A.u.sigma_color + 1e-5f;
    const float sd2 = A.u.sigma_depth * A.u.sigma_depth + 1e-5f;
    const float sa2 = A.u.sigma_albedo * A.u.sigma_albedo + 1e-5f;
    const float sm2 = A.u.sigma_material * A.u.sigma_material + 1e-5f;
    constexpr float kLog2e = 1.4426950408889634f;
    const float ic = kLog2e / sc2, id = kLog2e / sd2, ia = kLog2e / sa2, im = kLog2e / sm2;
    const float sn = A.u.sigma_normal;
    // EXACT: one row of taps per iteration (unrolled, the library exp/pow of
    // 25 taps do not fit the register file)
    constexpr int kRowUnroll = EXACT ? 1 : 5;
#pragma unroll kRowUnroll
    for (int dy = -2; dy <= 2; ++dy) {
#pragma unroll
        for (int dx = -2; dx <= 2; ++dx) {
            const int t = c + dy * kDnHx + dx;
            const float4 s = s_col[t], g0 = s_n[t], g1 = s_a[t], g2 = s_m[t];
            const float hk = kern[dx + 2] * kern[dy + 2];
            const v3 cd = mk(cc.x, cc.y, cc.z) - mk(s.x, s.y, s.z);
            const float dd = cg1.w - g1.w;
            const v3 ad = mk(cg1.x, cg1.y, cg1.z) - mk(g1.x, g1.y, g1.z);
            const v3 md_ = mk(cg2.x, cg2.y, cg2.z) - mk(g2.x, g2.y, g2.z);
            const bool normals = cg0.w != 0.0f || g0.w != 0.0f;
            float w;
            if (EXACT) {
                const float wc = expf(-dot(cd, cd) / sc2);
                float wn = 1.0f;
                if (normals) wn = powf(mx(0.0f, dot(cn, mk(g0.x, g0.y, g0.z))), sn);
                const float wd = expf(-(dd * dd) / sd2);
                const float wa = expf(-dot(ad, ad) / sa2);
                const float wm = expf(-dot(md_, md_) / sm2);
                w = hk * wc * wn * wd * wa * wm;
            } else {
                float e = fmaf(cd.x, cd.x, fmaf(cd.y, cd.y, cd.z * cd.z)) * ic;
                e = fmaf(dd * dd, id, e);
                e = fmaf(fmaf(ad.x, ad.x, fmaf(ad.y, ad.y, ad.z * ad.z)), ia, e);
                e = fmaf(fmaf(md_.x, md_.x, fmaf(md_.y, md_.y, md_.z * md_.z)), im, e);
                if (normals && sn != 0.0f) {
                    const float nd = mx(0.0f, fmaf(cn.x, g0.x, fmaf(cn.y, g0.y, cn.z * g0.z)));
                    e = fmaf(-sn, __builtin_amdgcn_logf(nd), e);  // pow(nd, sn) = 2^(sn log2 nd)
                }
                w = hk * __builtin_amdgcn_exp2f(-e);
            }
            if (dx != 0 || dy != 0) w *= att;
            if (EXACT) {
                sum.x = sum.x + s.x * w;
                sum.y = sum.y + s.y * w;
                sum.z = sum.z + s.z * w;
                sum.w = sum.w + s.w * w;
            } else {
                sum.x = fmaf(s.x, w, sum.x);
                sum.y = fmaf(s.y, w, sum.y);
                sum.z = fmaf(s.z, w, sum.z);
                sum.w = fmaf(s.w, w, sum.w);
            }
            sw += w;
        }
    }
    const float den = sw + 1e-5f;
    outb[uint32_t(y * W + x)] = make_float4(sum.x / den, sum.y / den, sum.z / den, sum.w / den);
}

// composite_main, denoiser.wgsl:194-214
__global__ void dn_composite(DenoiseArgs A, const float4* __restrict__ in) {
    const uint32_t x = blockIdx.x * 16 + threadIdx.x, y = blockIdx.y * 16 + threadIdx.y;
    if (x >= A.u.width || y >= A.u.height) return;
    const uint32_t i = y * A.u.width + x;
    float4 f = in[i];
    const float4 g0 = A.gn[i], g1 = A.ga[i];
    const bool sky = len(mk(g0.x, g0.y, g0.z)) < 0.1f;
    if ((A.u.flags & 2u) && !sky) {
        const v3 ma = vmax(mk(g1.x, g1.y, g1.z), splat(0.001f));
        f.x = f.x * ma.x;
        f.y = f.y * ma.y;
        f.z = f.z * ma.z;
    }
    A.fin[i] = f;
    A.out[i] = make_uchar4(unorm8(f.x), unorm8(f.y), unorm8(f.z), unorm8(f.w));
}

// --------------------------------------------------------------- sky CDF ----
constexpr int kCdfThreads = 256;

__device__ __forceinline__ double half_bits_to_double(uint16_t b) {
    return double(float(__builtin_bit_cast(_Float16, b)));
}

__device__ double block_exclusive_scan(double v, double* sh, double* total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < kCdfThreads; off <<= 1) {
        const double add = t >= off ? sh[t - off] : 0.0;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    const double incl = sh[t];
    *total = sh[kCdfThreads - 1];
    __syncthreads();
    return incl - v;
}

// One workgroup per row: each lane sums a contiguous run of texels in double,
// the runs are combined by a block scan, then every lane rewrites its run as
// running sums, stored as f32 and divided (as doubles) by the row total.
__global__ void __launch_bounds__(kCdfThreads) cdf_rows(const uint16_t* sky, uint32_t W, uint32_t H, float exposure,
                                                        float* cdf, double* row_tot) {
    __shared__ double sh[kCdfThreads];
    const uint32_t y = blockIdx.x;
    const double thr = 65519.0;
    const double sin_t = sin(3.141592653589793 * (double(y) + 0.5) / double(H));
    const uint32_t per = (W + kCdfThreads - 1) / kCdfThreads;
    const uint32_t x0 = threadIdx.x * per, x1 = min(W, x0 + per);
    auto lum_at = [&](uint32_t x) {
        const uint16_t* t = sky + (size_t(y) * W + x) * 4;
        const double r = fmin(half_bits_to_double(t[0]) * double(exposure), thr);
        const double g = fmin(half_bits_to_double(t[1]) * double(exposure), thr);
        const double b = fmin(half_bits_to_double(t[2]) * double(exposure), thr);
        return (0.2126 * r + 0.7152 * g + 0.0722 * b) * sin_t;
    };
    double local = 0.0;
    for (uint32_t x = x0; x < x1; ++x) local += lum_at(x);
    double total;
    double run = block_exclusive_scan(local, sh, &total);
    float* row = cdf + (H + 1) + size_t(y) * (W + 1);
    if (threadIdx.x == 0) row[0] = 0.0f;
    for (uint32_t x = x0; x < x1; ++x) {
        run += lum_at(x);
        row[x + 1] = total > 0.0 ? float(double(float(run)) / total) : float(double(x + 1) / double(W));
    }
    if (threadIdx.x == 0) row_tot[y] = total;
}

__global__ void __launch_bounds__(kCdfThreads) cdf_marginal(uint32_t H, float* cdf, const double* row_tot,
                                                            double* total_out) {
    __shared__ double sh[kCdfThreads];
    const uint32_t per = (H + kCdfThreads - 1) / kCdfThreads;
    const uint32_t y0 = threadIdx.x * per, y1 = min(H, y0 + per);
    double local = 0.0;
    for (uint32_t y = y0; y < y1; ++y) local += row_tot[y];
    double total;
    double run = block_exclusive_scan(local, sh, &total);
    if (threadIdx.x == 0) { cdf[0] = 0.0f; *total_out = total; }
    for (uint32_t y = y0; y < y1; ++y) {
        run += row_tot[y];
        cdf[y + 1] = total > 0.0 ? float(double(float(run)) / total) : float(double(y + 1) / double(H));
    }
}

__global__ void zero_u64(unsigned long long* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// ---------------------------------------------------------------- launchers --
static size_t stack_bytes(const DevScene& S, uint32_t threads, bool wide) {
    return size_t(wide ? S.wstack_cap : S.stack_cap) * threads * 4u;
}

template <uint32_t F, bool CN, int MODE, int W>
static hipError_t trace_variant(const DevScene& S, const TraceArgs& a, uint32_t grid, int* blocks_per_cu,
                                hipStream_t s) {
    const size_t lds = size_t(a.stack_lds) * kWave * 4u;
    if (blocks_per_cu)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, trace_kernel<F, CN, MODE, W>, kWave, lds);
    hipLaunchKernelGGL((trace_kernel<F, CN, MODE, W>), dim3(grid), dim3(kWave), lds, s, S, a);
    return hipGetLastError();
}

// Variants: hit recording and work counting use the reference-order
// traversal (they report the reference's per-pop counts); rendering uses the
// traversal the caller chose.
template <uint32_t F, bool CN, int W>
static hipError_t trace_sched(const DevScene& S, const TraceArgs& a, uint32_t grid, int* bpc, hipStream_t s) {
    if (a.schedule == PT_SCHEDULE_REGEN) return trace_variant<F, CN, kRegen, W>(S, a, grid, bpc, s);
    return trace_variant<F, CN, kSync, W>(S, a, grid, bpc, s);
}

template <uint32_t F>
static hipError_t trace_f(const DevScene& S, const TraceArgs& a, bool cn, uint32_t grid, int* bpc, hipStream_t s) {
    if (a.hits) return trace_variant<F, false, kRecord, kTravPairs>(S, a, grid, bpc, s);
    if (cn) return trace_sched<F, true, kTravPairs>(S, a, grid, bpc, s);
    if (a.wide == kTravQuant) return trace_sched<F, false, kTravQuant>(S, a, grid, bpc, s);
    if (a.wide == kTravWide) return trace_sched<F, false, kTravWide>(S, a, grid, bpc, s);
    return trace_sched<F, false, kTravPairs>(S, a, grid, bpc, s);
}

static hipError_t trace_dispatch(const DevScene& S, const TraceArgs& a, bool cn, uint32_t grid, int* bpc,
                                 hipStream_t s) {
    if (!a.hits && S.flags == kFlagsC1) return trace_f<kFlagsC1>(S, a, cn, grid, bpc, s);
    if (!a.hits && S.flags == kFlagsC2) return trace_f<kFlagsC2>(S, a, cn, grid, bpc, s);
    if (!a.hits && S.flags == kFlagsC3) return trace_f<kFlagsC3>(S, a, cn, grid, bpc, s);
    if (!a.hits && S.flags == kFlagsC1T) return trace_f<kFlagsC1T>(S, a, cn, grid, bpc, s);
    if (!a.hits && S.flags == kFlagsC1H) return trace_f<kFlagsC1H>(S, a, cn, grid, bpc, s);
    if (!a.hits && S.flags == kFlagsC1TA) return trace_f<kFlagsC1TA>(S, a, cn, grid, bpc, s);
    if (!a.hits && S.flags == kFlagsC1HA) return trace_f<kFlagsC1HA>(S, a, cn, grid, bpc, s);
    return trace_f<kDynamicFlags>(S, a, cn, grid, bpc, s);
}

hipError_t trace_blocks_per_cu(const DevScene& S, const TraceArgs& a, bool count_nodes, int* blocks_per_cu) {
    *blocks_per_cu = 0;
    return trace_dispatch(S, a, count_nodes, 0, blocks_per_cu, nullptr);
}

hipError_t launch_trace(const DevScene& S, const TraceArgs& a, uint32_t grid, bool count_nodes, hipStream_t s) {
    if (a.n_tiles == 0 || grid == 0 || a.n_passes == 0) return hipSuccess;
    return trace_dispatch(S, a, count_nodes, grid, nullptr, s);
}

hipError_t launch_gbuffer(const DevScene& S, const pt_scene_params& P, pt_gbuffer_texel* gbuf, uchar4* out,
                          hipStream_t s) {
    const uint32_t tiles = ((P.width + kTile - 1) / kTile) * ((P.height + kTile - 1) / kTile);
    hipLaunchKernelGGL((gbuffer_kernel<kDynamicFlags>), dim3(tiles), dim3(kWave), stack_bytes(S, kWave, false), s, S,
                       P, gbuf, out);
    return hipGetLastError();
}

hipError_t launch_rays(const DevScene& S, const RayArgs& a, bool shadow, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    const uint32_t blocks = (a.n + kWave - 1) / kWave;
    const size_t lds = stack_bytes(S, kWave, a.wide != kTravPairs);
    if (a.wide == kTravQuant)
        hipLaunchKernelGGL((rays_kernel<kDynamicFlags, kTravQuant>), dim3(blocks), dim3(kWave), lds, s, S, a,
                           shadow ? 1 : 0);
    else if (a.wide == kTravWide)
        hipLaunchKernelGGL((rays_kernel<kDynamicFlags, kTravWide>), dim3(blocks), dim3(kWave), lds, s, S, a,
                           shadow ? 1 : 0);
    else
        hipLaunchKernelGGL((rays_kernel<kDynamicFlags, kTravPairs>), dim3(blocks), dim3(kWave), lds, s, S, a,
                           shadow ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_denoise(const DenoiseArgs& a, int stride_mode, hipStream_t s) {
    const dim3 blk(16, 16), grd((a.u.width + 15) / 16, (a.u.height + 15) / 16);
    hipLaunchKernelGGL(dn_guide, grd, blk, 0, s, a);
    hipLaunchKernelGGL(dn_prepass, grd, blk, 0, s, a);
    const uint32_t W = a.u.width, H = a.u.height;
    for (uint32_t i = 0; i < a.u.passes; ++i) {
        const uint32_t step = stride_mode == 0 ? 1u << i : 1u << (a.u.passes - 1);
        const float4* in = (i % 2 == 0) ? a.a : a.b;
        float4* outb = (i % 2 == 0) ? a.b : a.a;
        // sub-lattice tiles: phases present in the image x tiles of the largest sub-lattice
        const uint32_t phx = std::min(step, W), phy = std::min(step, H);
        const uint32_t tiles_x = ((W + step - 1) / step + kDnTx - 1) / kDnTx;
        const uint32_t tiles_y = ((H + step - 1) / step + kDnTy - 1) / kDnTy;
        const uint64_t blocks = uint64_t(tiles_x) * tiles_y * phx * phy;
        if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
        if (a.exact)
            hipLaunchKernelGGL(dn_atrous<true>, dim3(uint32_t(blocks)), dim3(256), 0, s, a, in, outb, int(step),
                               tiles_x, phx, phx * phy);
        else
            hipLaunchKernelGGL(dn_atrous<false>, dim3(uint32_t(blocks)), dim3(256), 0, s, a, in, outb, int(step),
                               tiles_x, phx, phx * phy);
    }
    hipLaunchKernelGGL(dn_composite, grd, blk, 0, s, a, (a.u.passes % 2 == 0) ? a.a : a.b);
    return hipGetLastError();
}

hipError_t launch_sky_cdf(const uint16_t* sky, uint32_t w, uint32_t h, float exposure, float* cdf, double* row_tmp,
                          double* total, hipStream_t s) {
    hipLaunchKernelGGL(cdf_rows, dim3(h), dim3(kCdfThreads), 0, s, sky, w, h, exposure, cdf, row_tmp);
    hipLaunchKernelGGL(cdf_marginal, dim3(1), dim3(kCdfThreads), 0, s, h, cdf, row_tmp, total);
    return hipGetLastError();
}

hipError_t launch_zero_u64(unsigned long long* p, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(zero_u64, dim3((n + 63) / 64), dim3(64), 0, s, p, n);
    return hipGetLastError();
}

}  // namespace pt
