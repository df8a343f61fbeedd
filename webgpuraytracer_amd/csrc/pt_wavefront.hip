// pt_wavefront.hip — the wavefront form of shader.wgsl `main` (2273-2492).
//
// A batch of paths (whole progressive passes over the section) lives in HBM
// as structure-of-arrays state; each bounce is three launches:
//   wf_trace<EXT>     closest hit for every path in the active queue —
//                     traversal only, at a fraction of the megakernel's
//                     registers
//   wf_shade          the bounce's shading (dev/path.h, shared with the
//                     megakernel); queues the path for the next bounce and its
//                     NEE shadow rays
//   wf_trace<SHADOW>  any-hit test of the queued shadow rays; adds the light
//                     NEE then the env NEE contribution of each path, in the
//                     reference's order
// The two traversal launches are persistent with dynamic fetch: a wave keeps
// its lanes busy by taking new queue entries for its idle lanes (one atomic
// per refill) instead of idling until its longest ray ends.
// The shading kernel appends each block's surviving paths and shadow
// queries with one atomic per block (one atomic per wave on one counter
// bounded that kernel at half its speed), grouped by the direction octant of
// the ray, so that a traversal wave gets rays from one image region that
// visit the tree's children in the same order.  Every
// kernel reads its queue length from device memory, so a whole batch is
// launched without a host round trip.  The per-path RNG state and draw order
// are the megakernel's, so both backends produce the same bits.
//
// Path state traffic: before the first bounce only the camera ray is stored
// (32 bytes per path); the first shading derives the rest of the path's
// state (PCG state, throughput, radiance) from the slot index again, as
// path_begin does; a path that ends writes only its radiance; a shadow query
// is 48 bytes (origin, direction, and the contribution already multiplied by
// the throughput), and which slots a path queued travels in its shadow-queue
// entry, so no slot is cleared or tested in memory.  (Computing the camera
// ray inside the first traversal launch instead took 8 more VGPRs there and a
// wave per SIMD on the small analytic scenes.)
#include <algorithm>

#include "pt_internal.h"
#include "dev/path.h"

namespace pt {

constexpr int kWfBlock = 256;
// counters: per extension queue q (0, 1) the pair {its length, the length of
// the shadow queue filled with it} at 2q, 2q + 1 (one 64-bit word), then the
// fetch cursors of the two traversal kernels
enum { kCntFetchExt = 4, kCntFetchShadow = 5, kCntWords = 6 };
__device__ __forceinline__ int cnt_ext(int q) { return 2 * q; }
__device__ __forceinline__ int cnt_shadow(int q) { return 2 * q + 1; }
// a shadow-queue entry: the path index and what its shading left pending
// (light slot, env slot, deferred emission); the pool is below 2^29 paths
constexpr uint32_t kSqSlot0 = 1u << 31, kSqSlot1 = 1u << 30, kSqEmit = 1u << 29, kSqPath = kSqEmit - 1u;

// x * 0 for each factor x of the throughput, as the shadow kernel must add
// it for an occluded slot (+0, -0, or NaN for an infinite or NaN factor), in
// two bits per component: bit 0 sign, bit 1 NaN.
__device__ __forceinline__ uint32_t zero_product_bits(v3 tp) {
    uint32_t b = 0;
    const float c[3] = {tp.x, tp.y, tp.z};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const float z = c[j] * 0.0f;
        b |= (z != z ? 2u : (__float_as_uint(z) >> 31)) << (2 * j);
    }
    return b;
}
__device__ __forceinline__ v3 zero_product(uint32_t b) {
    float c[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const uint32_t q = (b >> (2 * j)) & 3u;
        c[j] = q & 2u ? __uint_as_float(0x7fc00000u) : __uint_as_float(q << 31);
    }
    return mk(c[0], c[1], c[2]);
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// The shading kernel's block: its paths enter the next queues as one group
// per block (one atomic), ordered by direction octant.
constexpr int kShadeBlock = 256;
constexpr uint32_t kShadeWaves = uint32_t(kShadeBlock) / 64u;

// Appends the block's paths to the extension queue (pred pa, key ka) and to
// the shadow queue (pred pb, key kb) with one 64-bit atomic on the pair
// {extension length, shadow length}.  Each queue receives the block's
// entries contiguously, grouped by key (0 .. 7, key 0's first), each key's
// in thread order.  Every thread of the block must call it; sm holds 64
// words.
__device__ __forceinline__ void block_push2(unsigned long long* pair, bool pa, uint32_t ka, bool pb, uint32_t kb,
                                            uint32_t* sm, uint32_t& qa, uint32_t& qb) {
    static_assert(8u * kShadeWaves == 32u, "one scan lane per (queue, key, wave)");
    const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    uint32_t ra = 0, rb = 0;
#pragma unroll
    for (uint32_t k = 0; k < 8u; ++k) {
        const bool ia = pa && ka == k, ib = pb && kb == k;
        const unsigned long long ma = __ballot(ia), mb = __ballot(ib);
        if (lane == 0u) {  // [queue][key][wave]: the block's order
            sm[k * kShadeWaves + wv] = uint32_t(__popcll(ma));
            sm[32u + k * kShadeWaves + wv] = uint32_t(__popcll(mb));
        }
        if (ia) ra = uint32_t(__popcll(ma & below));
        if (ib) rb = uint32_t(__popcll(mb & below));
    }
    __syncthreads();
    if (wv == 0u) {  // one scan over both queues' counts, then the block's atomic
        const uint32_t v = sm[lane];
        uint32_t s = v;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t t = __shfl_up(s, off, 64);
            if (lane >= uint32_t(off)) s += t;
        }
        const uint32_t ta = __shfl(s, 31, 64), tb = __shfl(s, 63, 64) - ta;
        unsigned long long base = 0;
        if (lane == 0u && (ta | tb)) base = atomicAdd(pair, (static_cast<unsigned long long>(tb) << 32) | ta);
        base = __shfl(base, 0, 64);
        sm[lane] = lane < 32u ? uint32_t(base) + s - v : uint32_t(base >> 32) + s - v - ta;
    }
    __syncthreads();
    qa = pa ? sm[ka * kShadeWaves + wv] + ra : 0u;
    qb = pb ? sm[32u + kb * kShadeWaves + wv] + rb : 0u;
}

// Direction octant: the sign bits of d.
__device__ __forceinline__ uint32_t octant(v3 d) {
    return (__float_as_uint(d.x) >> 31) | ((__float_as_uint(d.y) >> 31) << 1) | ((__float_as_uint(d.z) >> 31) << 2);
}

__device__ __forceinline__ unsigned long long wsum64(uint32_t v) {
    unsigned long long s = v;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    return s;
}

__device__ __forceinline__ void add_stat(unsigned long long* stats, int slot, uint32_t v) {
    const unsigned long long s = wsum64(v);
    if (lane_id() == 0 && s) atomicAdd(stats + slot, s);
}

__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
    const uint32_t q = nb / 8u, r = nb % 8u;
    const uint32_t xcd = b % 8u, slot = b / 8u;
    return (xcd < r ? xcd * (q + 1u) : r * (q + 1u) + (xcd - r) * q) + slot;
}

// Pixel of section slot lp (a bijection onto the section): the largest
// 8-aligned block of the section in 8x8 tiles, so that a wave's 64 paths
// start from one tile (coherent primary rays, nearby secondary origins),
// then the right strip and the bottom strip row by row.
__device__ __forceinline__ void section_pixel(const WfArgs& A, uint32_t lp, uint32_t& px, uint32_t& py) {
    const uint32_t w8 = A.sec_w & ~7u, h8 = A.sec_h & ~7u;
    const uint32_t main = w8 * h8;
    uint32_t x, y;
    if (lp < main) {
        const uint32_t t = lp >> 6, in = lp & 63u, tx = w8 >> 3;
        x = (t % tx) * 8u + (in & 7u);
        y = (t / tx) * 8u + (in >> 3);
    } else {
        uint32_t r = lp - main;
        const uint32_t rw = A.sec_w - w8, right = rw * h8;
        if (r < right) {
            y = r / rw;
            x = w8 + (r - y * rw);
        } else {
            r -= right;
            y = h8 + r / A.sec_w;
            x = r - (y - h8) * A.sec_w;
        }
    }
    px = A.P.section[0] + x;
    py = A.P.section[1] + y;
}

__device__ __forceinline__ void slot_pixel(const WfArgs& A, uint32_t i, uint32_t& px, uint32_t& py, uint32_t& pass) {
    const uint32_t npix = A.sec_w * A.sec_h;
    pass = i / npix;
    section_pixel(A, i - pass * npix, px, py);
}

// Path of batch slot i at bounce 0: camera ray and PCG state as the
// reference seeds them (shader.wgsl:2280-2287).
__device__ __forceinline__ void wf_path_begin(const WfArgs& A, uint32_t i, PathState& ps) {
    uint32_t px, py, pass;
    slot_pixel(A, i, px, py, pass);
    const uint32_t sn = A.P.sample_number + A.first_pass + pass;
    // the fallback as a register value: a select between the argument's and
    // the list's addresses would compile to a flat load, whose counter waits
    // then spread through the kernel
    const uint32_t fallback = __builtin_amdgcn_readfirstlane(A.P.seed);
    const uint32_t seed = A.seeds ? A.seeds[A.first_pass + pass] : fallback;
    path_begin(A.P, px, py, sn, seed, ps);
}

// The batch's camera rays (the first shading recomputes the rest of each
// path's start); queue lengths and fetch cursors start at zero (bounce 0 reads
// no queue: its paths are the batch slots 0 .. n_items - 1).
__global__ void __launch_bounds__(kWfBlock) wf_raygen(WfArgs A) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < uint32_t(kCntWords)) A.counters[i] = 0u;
    if (i >= A.n_items) return;
    PathState ps;
    wf_path_begin(A, i, ps);
    A.o_pdf[i] = make_float4(ps.ro.x, ps.ro.y, ps.ro.z, ps.last_pdf);
    A.d_wsum[i] = make_float4(ps.rd.x, ps.rd.y, ps.rd.z, ps.last_wsum);
}

// Persistent traversal of a queue: the extension rays of queue[cur]
// (SHADOW = false; writes each path's hit) or the shadow slots of the paths in
// the shadow queue (SHADOW = true; traces slot 0 then slot 1 of a path and
// adds their contributions to its radiance in that order).  Each wave runs
// until the queue is drained; whenever at least A.refill of its lanes are
// idle it takes that many entries with one atomic on the fetch cursor
// (dynamic fetch), so short and long rays share the wave's lanes.
// Occupancy of the traversal launches: capping them at 80 VGPRs (6 waves per
// SIMD) measured 3% faster than the uncapped ~95 (5 waves); 64 VGPRs spills
// in the traversal loop and is 7% slower.
#ifndef PT_WF_TRACE_WAVES_PER_EU
#define PT_WF_TRACE_WAVES_PER_EU 6
#endif
template <uint32_t F, bool CN, int W, bool SHADOW>
__global__ void __launch_bounds__(kWfBlock) __attribute__((amdgpu_waves_per_eu(PT_WF_TRACE_WAVES_PER_EU)))
wf_trace(DevScene S, WfArgs A, int cur, int bounce) {
    extern __shared__ uint32_t lds_stack[];
    const Flags<F> fl{S.flags};
    const bool first = !SHADOW && bounce == 0;  // the batch's camera rays, slot by slot
    // the counter is read even when unused so that n stays a scalar load (a
    // select between the argument and the counter's address compiles to a
    // flat load whose wait spreads through the kernel)
    const uint32_t qn = A.counters[SHADOW ? cnt_shadow(cur ^ 1) : cnt_ext(cur)];
    const uint32_t n = first ? A.n_items : qn;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // counters this launch does not read, reset for the launches after it
        if (SHADOW) {
            A.counters[kCntFetchExt] = 0u;
        } else {
            A.counters[cnt_ext(cur ^ 1)] = 0u;
            A.counters[cnt_shadow(cur ^ 1)] = 0u;
            A.counters[kCntFetchShadow] = 0u;
        }
    }
    uint32_t* fetch = &A.counters[SHADOW ? kCntFetchShadow : kCntFetchExt];
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = lane_id();
    LdsStack st{lds_column(lds_stack + threadIdx.x), blockDim.x, A.stack_lds, 0, A.stack_spill + gid, A.spill_stride,
                W ? S.wstack_cap : S.stack_cap};
    Counters c;
    c.clear();
    Trav<F, CN, W> q;
    HitRec h;
    uint32_t p = 0, pend = 0;  // path, and (shadow) its pending kSq* bits
    int k = 0;
    v3 rad = splat(0.0f);
    bool busy = false, drained = false;
    auto store_hit = [&]() {
        A.hit_a[p] = make_int4(__float_as_int(h.t), (h.kind << 30) | (h.idx & 0x3fffffff), h.prim, int(h.count));
        A.hit_b[p] = make_float2(h.bu, h.bv);
    };
    auto apply_slot = [&]() {
        const uint32_t sl = 2u * p + uint32_t(k);
        const float4 ct = A.sh_c[sl];
        if (fl.has(kTranslucentBit)) {  // ct holds the unclamped radiance and the bounce
            const float4 tp = A.sh_tp[sl];
            const v3 add = q.blocked || q.trans < 0.001f ? splat(0.0f)
                                                         : firefly(mk(ct.x, ct.y, ct.z) * q.trans, int(ct.w));
            rad = rad + mk(tp.x, tp.y, tp.z) * add;
        } else {  // ct holds throughput x clamped radiance; occluded: throughput x 0
            rad = rad + (q.blocked ? zero_product(__float_as_uint(ct.w)) : mk(ct.x, ct.y, ct.z));
        }
        ++k;
    };
    // starts the next shadow slot of path p, or stores its radiance when none is left
    auto next_slot = [&]() {
        for (;;) {
            while (k < 2 && !(pend & (k == 0 ? kSqSlot0 : kSqSlot1))) ++k;
            if (k >= 2) {
                if (pend & kSqEmit) {
                    const float4 e = A.sh_em[p];
                    rad = rad + mk(e.x, e.y, e.z);
                }
                A.rad[p] = make_float4(rad.x, rad.y, rad.z, 0.0f);
                busy = false;
                return;
            }
            const uint32_t sl = 2u * p + uint32_t(k);
            const float4 o = A.sh_o[sl], d = A.sh_d[sl];
            c.shadow++;
            if (q.begin_any(S, fl, mk(o.x, o.y, o.z), mk(d.x, d.y, d.z), __float_as_int(d.w), o.w, st, h)) {
                busy = true;
                return;
            }
            q.finish(S, fl, st, h, c);
            apply_slot();
        }
    };
    for (;;) {
        const unsigned long long idle = __ballot(!busy);
        if (!drained && uint32_t(__popcll(idle)) >= A.refill) {
            const uint32_t cnt = uint32_t(__popcll(idle));
            const uint32_t leader = uint32_t(__ffsll(static_cast<long long>(idle)) - 1);
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(fetch, cnt);
            base = __shfl(base, int(leader), 64);
            if (base + cnt >= n) drained = true;
            if (!busy) {
                const uint32_t i = base + uint32_t(__popcll(idle & ((1ull << lane) - 1ull)));
                if (i < n) {
                    if (SHADOW) {
                        const uint32_t e = A.shadow_queue[i];
                        p = e & kSqPath;
                        pend = e;
                        const float4 r = A.rad[p];
                        rad = mk(r.x, r.y, r.z);
                        k = 0;
                        next_slot();
                    } else {
                        p = first ? i : A.queue[cur][i];
                        const float4 o = A.o_pdf[p], d = A.d_wsum[p];
                        c.ext++;
                        busy = q.begin_closest(S, fl, mk(o.x, o.y, o.z), mk(d.x, d.y, d.z), st, h);
                        if (!busy) {
                            q.finish(S, fl, st, h, c);
                            store_hit();
                        }
                    }
                }
            }
        }
        if (drained && __ballot(busy) == 0ull) break;
        if (busy) {
            busy = q.step(S, fl, st, h, c);
            if (!busy) {
                q.finish(S, fl, st, h, c);
                if (SHADOW) {
                    apply_slot();
                    next_slot();
                } else {
                    store_hit();
                }
            }
        }
    }
    if (A.stats) {
        add_stat(A.stats, SHADOW ? kStShadow : kStExt, SHADOW ? c.shadow : c.ext);
        if (bounce < kStatBounces) add_stat(A.stats, (SHADOW ? kStShadowB0 : kStExtB0) + bounce, SHADOW ? c.shadow : c.ext);
        if (CN) {
            add_stat(A.stats, kStPops, c.pops);
            add_stat(A.stats, kStInner, c.inner);
            add_stat(A.stats, kStObjs, c.objs);
            add_stat(A.stats, kStTris, c.tris);
        }
    }
}

// Queues the NEE shadow query of slot k of path p (0: light, 1: env); `pend`
// collects the kSq* bits of the path's shadow-queue entry.
struct QueuedNee {
    const WfArgs& A;
    uint32_t p;
    uint32_t pend;
    bool att;          // translucent occluders possible
    uint32_t key = 0;  // direction octant of the last queued ray
    __device__ void operator()(int k, bool need_ray, v3 o, v3 d, float dist, int ignore, v3 inc, v3 tp, int bounce,
                               v3& rad) {
        if (!need_ray) {
            rad = rad + tp * firefly(inc, bounce);
            return;
        }
        key = octant(d);
        const uint32_t s = 2u * p + uint32_t(k);
        A.sh_o[s] = make_float4(o.x, o.y, o.z, dist);
        A.sh_d[s] = make_float4(d.x, d.y, d.z, __int_as_float(ignore));
        if (att) {  // clamped after a translucent occluder scales it
            A.sh_c[s] = make_float4(inc.x, inc.y, inc.z, float(bounce));
            A.sh_tp[s] = make_float4(tp.x, tp.y, tp.z, 0.0f);
        } else {  // the product rad + tp * L would add: the same rounding, done here
            const v3 ct = tp * firefly(inc, bounce);
            A.sh_c[s] = make_float4(ct.x, ct.y, ct.z, __uint_as_float(zero_product_bits(tp)));
        }
        pend |= k == 0 ? kSqSlot0 : kSqSlot1;
    }
    // radiance that must follow the queued NEE contributions: deferred to the
    // shadow kernel while a slot is pending
    __device__ void emit(v3 add, v3& rad) {
        if (pend && A.sh_em) {
            A.sh_em[p] = make_float4(add.x, add.y, add.z, 0.0f);
            pend |= kSqEmit;
        } else {
            rad = rad + add;
        }
    }
};

template <uint32_t F, bool CN>
__global__ void __launch_bounds__(kShadeBlock) wf_shade(DevScene S, WfArgs A, int cur, int bounce) {
    const Flags<F> fl{S.flags};
    const uint32_t qn = A.counters[cnt_ext(cur)];  // read unconditionally: a scalar load (see wf_trace)
    const uint32_t n = bounce == 0 ? A.n_items : qn;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x * blockDim.x >= n) return;
    const Params P{A.P.sky_total_lum, A.P.sky_width, A.P.sky_height, A.P.sky_cdf_index};
    __shared__ uint32_t grp[64];
    bool alive = false;
    uint32_t p = 0, pend = 0, ext_key = 0, sh_key = 0;
    Counters c;
    c.clear();
    if (i < n) {
        PathState ps;
        uint32_t meta = 0;
        if (bounce == 0) {  // batch slot i, straight from the camera
            p = i;
            wf_path_begin(A, i, ps);
        } else {
            p = A.queue[cur][i];
            const float4 o = A.o_pdf[p], d = A.d_wsum[p], t = A.tp_rng[p], r = A.rad[p];
            ps.ro = mk(o.x, o.y, o.z);
            ps.last_pdf = o.w;
            ps.rd = mk(d.x, d.y, d.z);
            ps.last_wsum = d.w;
            ps.tp = mk(t.x, t.y, t.z);
            ps.rng.s = __float_as_uint(t.w);
            ps.rad = mk(r.x, r.y, r.z);
            meta = A.meta[p];
        }
        const int4 ha = A.hit_a[p];
        const float2 hb = A.hit_b[p];
        HitRec h;
        h.t = __int_as_float(ha.x);
        h.kind = (ha.y >> 30) & 3;
        h.idx = ha.y & 0x3fffffff;
        if (h.kind == kMiss) h.idx = -1;
        h.prim = ha.z;
        h.count = uint32_t(ha.w);
        h.bu = hb.x;
        h.bv = hb.y;
        Media md{A.media + p, A.pool, int(meta & 7u)};
        const SurfHit sh = resolve_for_shading<F>(S, fl, ps.ro, ps.rd, h);
        QueuedNee sink{A, p, 0u, fl.has(kTranslucentBit)};
        alive = shade_bounce<F, CN>(S, fl, P, bounce, h, sh, ps, md, c, sink);
        alive = alive && bounce + 1 < S.bounce_limit;
        pend = sink.pend;
        sh_key = sink.key;
        ext_key = octant(ps.rd);
        if (alive) {  // an ended path keeps only its radiance
            A.o_pdf[p] = make_float4(ps.ro.x, ps.ro.y, ps.ro.z, ps.last_pdf);
            A.d_wsum[p] = make_float4(ps.rd.x, ps.rd.y, ps.rd.z, ps.last_wsum);
            A.tp_rng[p] = make_float4(ps.tp.x, ps.tp.y, ps.tp.z, __uint_as_float(ps.rng.s));
            if (bounce == 0 || uint32_t(md.n) != meta) A.meta[p] = uint32_t(md.n);
        }
        A.rad[p] = make_float4(ps.rad.x, ps.rad.y, ps.rad.z, 0.0f);
    }
    uint32_t qa, qs;
    block_push2(reinterpret_cast<unsigned long long*>(A.counters + cnt_ext(cur ^ 1)), alive, ext_key, pend != 0u, sh_key,
                grp, qa, qs);
    if (alive) A.queue[cur ^ 1][qa] = p;
    if (pend) A.shadow_queue[qs] = p | pend;
    if (A.stats && CN) add_stat(A.stats, kStShade, c.shade);
}

// Running mean over the batch's passes in pass order (shader.wgsl:2482-2491).
__global__ void __launch_bounds__(kWfBlock) wf_accumulate(WfArgs A) {
    const uint32_t lp = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t npix = A.sec_w * A.sec_h;
    if (lp >= npix) return;
    uint32_t px, py;
    section_pixel(A, lp, px, py);
    const uint32_t idx = py * A.P.width + px;
    const float4 old = A.accum[idx];
    v3 acc = mk(old.x, old.y, old.z);
    for (uint32_t pass = 0; pass < A.batch_passes; ++pass) {
        const float4 r = A.rad[pass * npix + lp];
        const uint32_t sn = A.P.sample_number + A.first_pass + pass;
        acc = lerp(acc, mk(r.x, r.y, r.z), 1.0f / float(sn + 1u));
    }
    A.accum[idx] = make_float4(acc.x, acc.y, acc.z, 1.0f);
    if (A.write_output) {
        v3 o = acc * splat(A.P.exposure);
        o = o / (o + splat(0.3f));
        A.out[idx] = make_uchar4(unorm8(powf(o.x, 0.4545f)), unorm8(powf(o.y, 0.4545f)), unorm8(powf(o.z, 0.4545f)),
                                 255);
    }
    if (A.stats && lp == 0) atomicAdd(A.stats + kStPaths, (unsigned long long)A.n_items);
}

// bounce_ev (optional): bounce_limit + 1 events, recorded before each
// bounce's launches and after the last.
template <uint32_t F, bool CN, int W>
static void wf_bounces(const DevScene& S, const WfArgs& a0, uint32_t shade_blocks, hipStream_t s,
                       hipEvent_t* bounce_ev) {
    const size_t lds = size_t(a0.stack_lds) * kWfBlock * 4u;
    int cur = 0;
    for (int b = 0; b < S.bounce_limit; ++b) {
        if (bounce_ev) (void)hipEventRecord(bounce_ev[b], s);
        hipLaunchKernelGGL((wf_trace<F, CN, W, false>), dim3(a0.trace_grid), dim3(kWfBlock), lds, s, S, a0, cur, b);
        hipLaunchKernelGGL((wf_shade<F, CN>), dim3(shade_blocks), dim3(kShadeBlock), 0, s, S, a0, cur, b);
        hipLaunchKernelGGL((wf_trace<F, CN, W, true>), dim3(a0.trace_grid), dim3(kWfBlock), lds, s, S, a0, cur, b);
        cur ^= 1;
    }
    if (bounce_ev) (void)hipEventRecord(bounce_ev[S.bounce_limit], s);
}

// Work counting runs the reference-order traversal (its per-pop counts).
template <uint32_t F>
static hipError_t wf_run(const DevScene& S, const WfArgs& a0, bool cn, hipStream_t s, hipEvent_t* bounce_ev) {
    const uint32_t blocks = (a0.n_items + kWfBlock - 1) / kWfBlock;
    hipLaunchKernelGGL(wf_raygen, dim3(blocks), dim3(kWfBlock), 0, s, a0);
    // without a bounce no shading writes the radiance: the paths add zero
    if (S.bounce_limit <= 0) {
        const hipError_t e = launch_zero_u64(reinterpret_cast<unsigned long long*>(a0.rad), 2u * a0.n_items, s);
        if (e != hipSuccess) return e;
    }
    const uint32_t sb = (a0.n_items + uint32_t(kShadeBlock) - 1u) / uint32_t(kShadeBlock);
    if (cn)
        wf_bounces<F, true, kTravPairs>(S, a0, sb, s, bounce_ev);
    else if (a0.wide == kTravQuant)
        wf_bounces<F, false, kTravQuant>(S, a0, sb, s, bounce_ev);
    else if (a0.wide == kTravWide)
        wf_bounces<F, false, kTravWide>(S, a0, sb, s, bounce_ev);
    else
        wf_bounces<F, false, kTravPairs>(S, a0, sb, s, bounce_ev);
    const uint32_t npix = a0.sec_w * a0.sec_h;
    hipLaunchKernelGGL(wf_accumulate, dim3((npix + kWfBlock - 1) / kWfBlock), dim3(kWfBlock), 0, s, a0);
    return hipGetLastError();
}

template <uint32_t F, bool CN, int W>
static hipError_t wf_occupancy_w(size_t lds, int* bpc) {
    int a = 0, b = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, wf_trace<F, CN, W, false>, kWfBlock, lds);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, wf_trace<F, CN, W, true>, kWfBlock, lds);
    *bpc = std::max(a, b);
    return e;
}

template <uint32_t F>
static hipError_t wf_occupancy_f(bool cn, int wide, size_t lds, int* bpc) {
    if (cn) return wf_occupancy_w<F, true, kTravPairs>(lds, bpc);
    if (wide == kTravQuant) return wf_occupancy_w<F, false, kTravQuant>(lds, bpc);
    if (wide == kTravWide) return wf_occupancy_w<F, false, kTravWide>(lds, bpc);
    return wf_occupancy_w<F, false, kTravPairs>(lds, bpc);
}

hipError_t wf_trace_blocks_per_cu(const DevScene& S, bool count_nodes, int wide, uint32_t stack_lds,
                                  int* blocks_per_cu) {
    const size_t lds = size_t(stack_lds) * kWfBlock * 4u;
    if (S.flags == kFlagsC1) return wf_occupancy_f<kFlagsC1>(count_nodes, wide, lds, blocks_per_cu);
    if (S.flags == kFlagsC2) return wf_occupancy_f<kFlagsC2>(count_nodes, wide, lds, blocks_per_cu);
    if (S.flags == kFlagsC3) return wf_occupancy_f<kFlagsC3>(count_nodes, wide, lds, blocks_per_cu);
    if (S.flags == kFlagsC1T) return wf_occupancy_f<kFlagsC1T>(count_nodes, wide, lds, blocks_per_cu);
    if (S.flags == kFlagsC1H) return wf_occupancy_f<kFlagsC1H>(count_nodes, wide, lds, blocks_per_cu);
    if (S.flags == kFlagsC1TA) return wf_occupancy_f<kFlagsC1TA>(count_nodes, wide, lds, blocks_per_cu);
    if (S.flags == kFlagsC1HA) return wf_occupancy_f<kFlagsC1HA>(count_nodes, wide, lds, blocks_per_cu);
    return wf_occupancy_f<kDynamicFlags>(count_nodes, wide, lds, blocks_per_cu);
}

bool specialised_flag_set(uint32_t flags) {
    return flags == kFlagsC1 || flags == kFlagsC2 || flags == kFlagsC3 || flags == kFlagsC1T || flags == kFlagsC1H ||
           flags == kFlagsC1TA || flags == kFlagsC1HA;
}

hipError_t launch_wavefront_batch(const DevScene& S, const WfArgs& a, bool count_nodes, hipStream_t s,
                                  hipEvent_t* bounce_ev) {
    if (S.flags == kFlagsC1) return wf_run<kFlagsC1>(S, a, count_nodes, s, bounce_ev);
    if (S.flags == kFlagsC2) return wf_run<kFlagsC2>(S, a, count_nodes, s, bounce_ev);
    if (S.flags == kFlagsC3) return wf_run<kFlagsC3>(S, a, count_nodes, s, bounce_ev);
    if (S.flags == kFlagsC1T) return wf_run<kFlagsC1T>(S, a, count_nodes, s, bounce_ev);
    if (S.flags == kFlagsC1H) return wf_run<kFlagsC1H>(S, a, count_nodes, s, bounce_ev);
    if (S.flags == kFlagsC1TA) return wf_run<kFlagsC1TA>(S, a, count_nodes, s, bounce_ev);
    if (S.flags == kFlagsC1HA) return wf_run<kFlagsC1HA>(S, a, count_nodes, s, bounce_ev);
    return wf_run<kDynamicFlags>(S, a, count_nodes, s, bounce_ev);
}

}  // namespace pt
