// nn_kernels.hip — gfx950 kernels of the learned denoiser.
//
// 3x3 convolutions run as implicit GEMMs on v_mfma_f32_16x16x32_bf16
// (bf16 operands, fp32 accumulation).  Lane layout of one 16x16x32 product,
// checked on the device: lane l holds A[l % 16][(l / 16) * 8 + j] and
// B[(l / 16) * 8 + j][l % 16] for j = 0..7, and receives
// C[(l / 16) * 4 + r][l % 16] for r = 0..3.
//
//   conv3x3_k   forward (and, with transposed weights, the data gradient):
//               rows = 16 pixels per wave, columns = output channels, K = the
//               9 taps x input channels in chunks of 32; bias + ReLU fused
//   wgrad_k     weight gradient: rows = output channels, columns = input
//               channels, K = pixels; NHWC tiles staged in LDS, fragments
//               read with LDS transpose loads; split over pixel chunks,
//               partial sums reduced in a fixed order
//   the rest    pooling, upsampling, Adam, loss, packing (16-byte vectors)
#include "nn_internal.h"

#include <algorithm>

namespace pt {
namespace nn {

typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef short i16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr uint32_t kBlock = 256;

__device__ __forceinline__ bf16x8 load8(const bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }
__device__ __forceinline__ void store8(bf16* p, bf16x8 v) { *reinterpret_cast<bf16x8*>(p) = v; }
__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, u32x4{0u, 0u, 0u, 0u}); }
__device__ __forceinline__ float f(bf16 x) { return float(x); }
__device__ __forceinline__ bf16 b(float x) { return bf16(x); }

__device__ __forceinline__ unsigned char unorm8n(float v) {
    const float c = fminf(fmaxf(v, 0.0f), 1.0f);
    return (unsigned char)(c * 255.0f + 0.5f);
}

// The denoised texel from the network's log-space prediction q[0..2]:
// radiance expm1(q) (at least 0) into *fin and, when ldr is set, the display
// tonemap (x e / (x e + 0.3))^0.4545 as rgba8.
__device__ __forceinline__ void frame_texel(const float* q, float exposure, float4* fin, uchar4* ldr) {
    const float r = fmaxf(expm1f(q[0]), 0.0f), g = fmaxf(expm1f(q[1]), 0.0f), bl = fmaxf(expm1f(q[2]), 0.0f);
    *fin = make_float4(r, g, bl, 1.0f);
    if (ldr) {
        float o[3] = {r * exposure, g * exposure, bl * exposure};
#pragma unroll
        for (int k = 0; k < 3; ++k) o[k] = powf(o[k] / (o[k] + 0.3f), 0.4545f);
        *ldr = make_uchar4(unorm8n(o[0]), unorm8n(o[1]), unorm8n(o[2]), 255);
    }
}

// Four floats as OCP e4m3 bytes, clamped to +-448 first (v_cvt_pk_fp8_f32
// turns larger magnitudes into NaN rather than saturating), rounded to
// nearest even.
__device__ __forceinline__ uint32_t e4m3x4(float a, float b, float c, float d) {
    a = fminf(fmaxf(a, -448.0f), 448.0f);
    b = fminf(fmaxf(b, -448.0f), 448.0f);
    c = fminf(fmaxf(c, -448.0f), 448.0f);
    d = fminf(fmaxf(d, -448.0f), 448.0f);
    const int lo = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
    return uint32_t(__builtin_amdgcn_cvt_pk_fp8_f32(c, d, lo, true));
}

// ------------------------------------------------------------- conv fwd --
// A block computes a 16 x 16 pixel tile of one image for a group of up to
// NCT x 16 output channels.  Per 32-channel slice of the input it stages the
// 18 x 18 halo and the slice's weights in LDS, both stored by 8-channel
// group ([group][pixel], [group][tap][co]) so the 16 lanes of a fragment row
// read 256 consecutive bytes.  Wave w owns tile rows 4w .. 4w + 3 (one
// 16-pixel MFMA row tile each), so per tap it reads 4 input fragments and NCT
// weight fragments from LDS for 4 x NCT products.
constexpr uint32_t kTile = 16, kHalo = 18, kHaloPix = kHalo * kHalo;  // 324
constexpr uint32_t kSlice = 32;                                         // input channels per stage
constexpr uint32_t kHaloPieces = kHaloPix * (kSlice / 8u);              // 16-byte pieces: 1296

template <int NCT>
struct ConvSmem {
    bf16x8 halo[4 * kHaloPix];    // [8-channel group][pixel]
    bf16x8 w[4 * 9 * NCT * 16];   // [8-channel group][tap][co]
};

// FULL: every block has all NCT output tiles (cout a multiple of NCT x 16),
// so the products need no guards; the bf16 output then goes through LDS
// ([pixel][channel], rows padded by 16 bytes against bank conflicts) and
// leaves as 16-byte stores, those of pixels past the image's edge dropped.
// With a FULL bf16 output the epilogue can also write (fz) the 2x2 average
// pool of the output tile and its nearest 2x upsampling; with a FULL fp32
// output, the denoised frame texels.  A slice is fetched, staged and
// multiplied in turn: the staging registers are free during the products, so
// the kernel needs half the VGPRs of one that fetches the next slice while
// multiplying (twice the resident waves, which overlap one another's loads;
// measured 4% faster per 1080p frame than prefetching where cin > 32).  UP:
// input channels below fz.cup are read from the half-resolution fz.up at the
// nearest coarse pixel, the others from `in` (a decoder level's first
// convolution over [upsampled | skip] without the upsampled copy; `in`
// then holds only the skip channels).  PF
// (launches of few blocks over several slices, where the extra registers
// cost no occupancy): the next slice is fetched during this slice's products.
// OUT8 (FULL bf16 output, no pooling or upsampling): the staged bf16 tile
// leaves as e4m3 bytes quantized with fz.q_out[co] (the next layer's fp8
// input scales; strides and offsets in bytes).
template <int NCT, bool RELU, bool F32OUT, bool FULL, bool UP = false, bool PF = false, bool OUT8 = false>
__global__ void __launch_bounds__(kBlock) conv3x3_k(Tens in, uint32_t cin, const bf16* __restrict__ w,
                                                    const float* __restrict__ bias, Tens out, uint32_t cout, Dims d,
                                                    float* __restrict__ of32, uint32_t f32c, ConvFuse fz) {
    __shared__ ConvSmem<NCT> sm;
    constexpr uint32_t kWPieces = 9u * NCT * 16u * 4u;
    constexpr uint32_t kHaloPer = (kHaloPieces + kBlock - 1u) / kBlock;  // 6
    constexpr uint32_t kWPer = (kWPieces + kBlock - 1u) / kBlock;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wv = tid >> 6, kq = lane >> 4, ln = lane & 15u;
    const uint32_t tiles_x = (d.w + kTile - 1u) / kTile, tiles_y = (d.h + kTile - 1u) / kTile;
    const uint32_t co0 = blockIdx.y * uint32_t(NCT) * 16u;
    const uint32_t nct = FULL ? uint32_t(NCT) : min(uint32_t(NCT), (cout - co0) / 16u);
    // This thread's halo pieces are the same in every tile: 8-channel group
    // tid % 4 of halo pixels tid / 4 + 64 k (row << 8 | column, ~0u past the halo).
    const uint32_t hc = (tid & 3u) * 8u;
    uint32_t hpos[kHaloPer];
#pragma unroll
    for (uint32_t k = 0; k < kHaloPer; ++k) {
        const uint32_t j = (tid >> 2) + k * (kBlock / 4u);
        hpos[k] = j < kHaloPix ? ((j / kHalo) << 8) | (j % kHalo) : ~0u;
    }
    auto fetch_w = [&](uint32_t c0, bf16x8* wr) {
#pragma unroll
        for (uint32_t k = 0; k < kWPer; ++k) {
            const uint32_t piece = tid + k * kBlock;
            const uint32_t c8 = piece % 4u, co = (piece / 4u) % (NCT * 16u), tap = piece / (4u * NCT * 16u);
            const uint32_t c = c0 + c8 * 8u;
            const bool ok = piece < kWPieces && co0 + co < cout && c < cin;
            wr[k] = ok ? load8(w + ((tap * cout + co0 + co) * cin + c)) : zero8();
        }
    };
    auto stash_w = [&](const bf16x8* wr) {
#pragma unroll
        for (uint32_t k = 0; k < kWPer; ++k) {
            const uint32_t piece = tid + k * kBlock;
            if (piece < kWPieces) sm.w[(piece % 4u) * (kWPieces / 4u) + piece / 4u] = wr[k];
        }
    };
    // The products run as C[co][pixel] = W[co][ci] X[ci][pixel] (weights as
    // the A operand), so lane (kq, ln) receives output channels 4 kq .. +3
    // of pixel ln of a 16-pixel row: 4 consecutive channels, one 8-byte piece.
    // The accumulators start at the bias.
    float bv[NCT][4];
#pragma unroll
    for (int t = 0; t < NCT; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) bv[t][r] = 0.0f;
    if (bias) {
#pragma unroll
        for (int t = 0; t < NCT; ++t)
            if (uint32_t(t) < nct)
#pragma unroll
                for (int r = 0; r < 4; ++r) bv[t][r] = bias[co0 + uint32_t(t) * 16u + kq * 4u + uint32_t(r)];
    }
    const uint32_t img = blockIdx.x / (tiles_x * tiles_y), trem = blockIdx.x % (tiles_x * tiles_y);
    const uint32_t ty0 = trem / tiles_x * kTile, tx0 = trem % tiles_x * kTile;
    const uint32_t ibase = img * d.h * d.w;  // element offsets fit 32 bits (checked by the host)
    uint32_t hsrc[kHaloPer];
    uint32_t cpix[UP ? kHaloPer : 1];
    bool hval[kHaloPer];
#pragma unroll
    for (uint32_t k = 0; k < kHaloPer; ++k) {
        const uint32_t gy = ty0 + (hpos[k] >> 8) - 1u, gx = tx0 + (hpos[k] & 255u) - 1u;  // wraps below 0
        hval[k] = hpos[k] != ~0u && gy < d.h && gx < d.w;
        const uint32_t sp = hval[k] ? gy * d.w + gx : 0u;
        hsrc[k] = (ibase + sp) * in.stride + in.off + hc;
        if (UP) cpix[UP ? k : 0] = hval[k] ? (img * (d.h / 2u) + gy / 2u) * (d.w / 2u) + gx / 2u : 0u;
    }
    bf16x8 hreg[kHaloPer], wreg[kWPer];
    auto fetch = [&](uint32_t c0) {
        const uint32_t c = c0 + hc;
#pragma unroll
        for (uint32_t k = 0; k < kHaloPer; ++k) {
            if (UP && c < fz.cup)
                hreg[k] = hval[k] ? load8(fz.up.p + (cpix[UP ? k : 0] * fz.up.stride + fz.up.off + c)) : zero8();
            else
                hreg[k] = (hval[k] && c < cin) ? load8(in.p + (hsrc[k] + c0 - (UP ? fz.cup : 0u))) : zero8();
        }
        fetch_w(c0, wreg);
    };
    auto stash = [&]() {
#pragma unroll
        for (uint32_t k = 0; k < kHaloPer; ++k)
            if (hpos[k] != ~0u) sm.halo[(tid & 3u) * kHaloPix + (tid >> 2) + k * (kBlock / 4u)] = hreg[k];
        stash_w(wreg);
    };
    f32x4 acc[4][NCT];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int t = 0; t < NCT; ++t) acc[m][t] = f32x4{bv[t][0], bv[t][1], bv[t][2], bv[t][3]};
    if (PF) fetch(0u);
    for (uint32_t c0 = 0; c0 < cin; c0 += kSlice) {
        if (!PF) fetch(c0);
        __syncthreads();  // the previous slice's reads are done
        stash();
        __syncthreads();
        if (PF && c0 + kSlice < cin) fetch(c0 + kSlice);
        // (reading the wave's 6 halo rows once per horizontal offset instead of
        // 4 rows per tap measured ~5% slower: the loads then wait together)
#pragma unroll
        for (uint32_t tap = 0; tap < 9u; ++tap) {
            const uint32_t dy = tap / 3u, dx = tap % 3u;
            bf16x8 a[4], bw[NCT];
#pragma unroll
            for (int m = 0; m < 4; ++m) a[m] = sm.halo[kq * kHaloPix + (wv * 4u + uint32_t(m) + dy) * kHalo + ln + dx];
#pragma unroll
            for (int t = 0; t < NCT; ++t) bw[t] = sm.w[kq * (kWPieces / 4u) + tap * NCT * 16u + uint32_t(t) * 16u + ln];
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int t = 0; t < NCT; ++t)
                    if (uint32_t(t) < nct)
                        acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bw[t], a[m], acc[m][t], 0, 0, 0);
        }
    }
    // acc[m][t][r]: channel co0 + 16 t + 4 kq + r of tile pixel (4 wv + m, ln)
    if (FULL && F32OUT) {
        // the first f32c (<= 4) channels as fp32 [pixel][f32c]: a tile row is
        // 16 * f32c consecutive floats, written as 16-byte pieces
        float* stage = reinterpret_cast<float*>(&sm);
        static_assert(kTile * kTile * 4u * 4u <= sizeof(sm.halo), "fp32 output stage fits the halo area");
        __syncthreads();  // every wave is done with the last slice
        if (kq * 4u < f32c) {
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const uint32_t ch = kq * 4u + uint32_t(r);
                    float v = acc[m][0][r];
                    if (RELU) v = fmaxf(v, 0.0f);
                    if (ch < f32c) stage[((wv * 4u + uint32_t(m)) * kTile + ln) * f32c + ch] = v;
                }
        }
        __syncthreads();
        if (fz.fin) {  // the frame's denoised texels straight from the stage (f32c = 3, one image)
            const uint32_t x = tx0 + tid % kTile, y = ty0 + tid / kTile;
            if (x < fz.fw && y < fz.fh) {
                const size_t i = size_t(y) * fz.fw + x;
                frame_texel(stage + tid * f32c, fz.exposure, fz.fin + i, fz.ldr ? fz.ldr + i : nullptr);
            }
            return;
        }
        const uint32_t per_row = kTile * f32c / 4u;  // 16-byte pieces per tile row
        for (uint32_t i = tid; i < kTile * per_row; i += kBlock) {
            const uint32_t row = i / per_row, k = i % per_row;
            const uint32_t op = (img * d.h + ty0 + row) * d.w + tx0;
            *reinterpret_cast<float4*>(of32 + (op * f32c + k * 4u)) =
                *reinterpret_cast<const float4*>(stage + row * kTile * f32c + k * 4u);
        }
        return;
    }
    if (FULL && !F32OUT) {
        constexpr uint32_t kCo = uint32_t(NCT) * 16u, kPitch = kCo * 2u + 16u;  // bytes per staged pixel
        static_assert(kTile * kTile * kPitch <= sizeof(ConvSmem<NCT>), "output stage fits the LDS tile");
        char* stage = reinterpret_cast<char*>(&sm);
        __syncthreads();  // every wave is done with the last slice
#pragma unroll
        for (int t = 0; t < NCT; ++t)
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                bf16x4 q;
#pragma unroll
                for (int r = 0; r < 4; ++r) q[r] = b(acc[m][t][r]);
                // ReLU on the rounded values as signed 16-bit integers: a bf16
                // is negative exactly when its bit pattern is (max with +0)
                i16x4 qi = __builtin_bit_cast(i16x4, q);
                if (RELU) qi = __builtin_elementwise_max(qi, i16x4{0, 0, 0, 0});
                const uint32_t p = (wv * 4u + uint32_t(m)) * kTile + ln;
                *reinterpret_cast<i16x4*>(stage + p * kPitch + (uint32_t(t) * 16u + kq * 4u) * 2u) = qi;
            }
        __syncthreads();
        const uint32_t obase = (img * d.h + ty0) * d.w + tx0;
        if (OUT8) {  // 16 channels of a pixel per 16-byte store
            constexpr uint32_t kC16 = kCo / 16u;
#pragma unroll
            for (uint32_t i = tid; i < kTile * kTile * kC16; i += kBlock) {
                const uint32_t p = i / kC16, c16 = i % kC16;
                if (ty0 + p / kTile >= d.h || tx0 + p % kTile >= d.w) continue;
                const uint32_t op = obase + (p / kTile) * d.w + p % kTile;
                const bf16x8 v0 = *reinterpret_cast<const bf16x8*>(stage + p * kPitch + c16 * 32u);
                const bf16x8 v1 = *reinterpret_cast<const bf16x8*>(stage + p * kPitch + c16 * 32u + 16u);
                const f32x4* sc = reinterpret_cast<const f32x4*>(fz.q_out + co0 + c16 * 16u);
                const f32x4 s0 = sc[0], s1 = sc[1], s2 = sc[2], s3 = sc[3];
                const u32x4 q{e4m3x4(f(v0[0]) * s0[0], f(v0[1]) * s0[1], f(v0[2]) * s0[2], f(v0[3]) * s0[3]),
                              e4m3x4(f(v0[4]) * s1[0], f(v0[5]) * s1[1], f(v0[6]) * s1[2], f(v0[7]) * s1[3]),
                              e4m3x4(f(v1[0]) * s2[0], f(v1[1]) * s2[1], f(v1[2]) * s2[2], f(v1[3]) * s2[3]),
                              e4m3x4(f(v1[4]) * s3[0], f(v1[5]) * s3[1], f(v1[6]) * s3[2], f(v1[7]) * s3[3])};
                *reinterpret_cast<u32x4*>(reinterpret_cast<uint8_t*>(out.p) + op * out.stride + out.off + co0 +
                                          c16 * 16u) = q;
            }
            return;
        }
        constexpr uint32_t kChunks = kTile * kTile * kCo / 8u;  // 16-byte pieces of the tile
#pragma unroll
        for (uint32_t i = tid; i < kChunks; i += kBlock) {
            const uint32_t p = i / (kCo / 8u), c8 = i % (kCo / 8u);
            if (ty0 + p / kTile >= d.h || tx0 + p % kTile >= d.w) continue;
            const uint32_t op = obase + (p / kTile) * d.w + p % kTile;
            store8(out.p + (op * out.stride + out.off + co0 + c8 * 8u),
                   *reinterpret_cast<const bf16x8*>(stage + p * kPitch + c8 * 16u));
        }
        if (fz.pool.p) {  // the 8 x 8 pooled tile, summed in pool2's order
            constexpr uint32_t kQ = kTile / 2u;
            for (uint32_t i = tid; i < kQ * kQ * (kCo / 8u); i += kBlock) {
                const uint32_t q = i / (kCo / 8u), c8 = i % (kCo / 8u), qy = q / kQ, qx = q % kQ;
                if (ty0 + 2u * qy >= d.h || tx0 + 2u * qx >= d.w) continue;  // (d.h and d.w are even)
                const uint32_t p00 = 2u * qy * kTile + 2u * qx;
                const char* sb = stage + c8 * 16u;
                const bf16x8 v0 = *reinterpret_cast<const bf16x8*>(sb + p00 * kPitch);
                const bf16x8 v1 = *reinterpret_cast<const bf16x8*>(sb + (p00 + 1u) * kPitch);
                const bf16x8 v2 = *reinterpret_cast<const bf16x8*>(sb + (p00 + kTile) * kPitch);
                const bf16x8 v3 = *reinterpret_cast<const bf16x8*>(sb + (p00 + kTile + 1u) * kPitch);
                bf16x8 o;
#pragma unroll
                for (int k = 0; k < 8; ++k) o[k] = b(0.25f * (f(v0[k]) + f(v1[k]) + f(v2[k]) + f(v3[k])));
                const uint32_t cp = (img * (d.h / 2u) + ty0 / 2u + qy) * (d.w / 2u) + tx0 / 2u + qx;
                store8(fz.pool.p + (cp * fz.pool.stride + fz.pool.off + co0 + c8 * 8u), o);
            }
        }
        if (fz.upw.p) {  // each output pixel copied to its 2 x 2 block of the fine tensor (as up2)
            const uint32_t W2 = 2u * d.w;
            for (uint32_t i = tid; i < kTile * kTile * 4u * (kCo / 8u); i += kBlock) {
                const uint32_t c8 = i % (kCo / 8u), f = i / (kCo / 8u);  // f: fine pixel of the 32 x 32 block
                const uint32_t fy = f / (2u * kTile), fx = f % (2u * kTile);
                if (ty0 + fy / 2u >= d.h || tx0 + fx / 2u >= d.w) continue;
                const uint32_t p = (fy / 2u) * kTile + fx / 2u;
                const uint32_t fp = (img * 2u * d.h + 2u * ty0 + fy) * W2 + 2u * tx0 + fx;
                store8(fz.upw.p + (fp * fz.upw.stride + fz.upw.off + co0 + c8 * 8u),
                       *reinterpret_cast<const bf16x8*>(stage + p * kPitch + c8 * 16u));
            }
        }
        return;
    }
#pragma unroll
    for (int t = 0; t < NCT; ++t) {
        if (uint32_t(t) >= nct) continue;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const uint32_t y = ty0 + wv * 4u + uint32_t(m), x = tx0 + ln;
            if (y >= d.h || x >= d.w) continue;
            const uint32_t op = (img * d.h + y) * d.w + x;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t co = co0 + uint32_t(t) * 16u + kq * 4u + uint32_t(r);
                float v = acc[m][t][r];
                if (RELU) v = fmaxf(v, 0.0f);
                if (F32OUT) {
                    if (co < f32c) of32[op * f32c + co] = v;
                } else {
                    out.p[op * out.stride + out.off + co] = b(v);
                }
            }
        }
    }
}

// Launches of fewer blocks than this prefetch the next input slice.
constexpr uint32_t kConvPfBlocks = 1024;

template <int NCT>
static void conv_launch(const Tens& in, uint32_t cin, const bf16* w, const float* bias, const Tens& out,
                        uint32_t cout, Dims d, bool relu, float* of32, uint32_t f32c, const ConvFuse& fz,
                        hipStream_t s) {
    const uint32_t tiles = d.n * ((d.h + kTile - 1u) / kTile) * ((d.w + kTile - 1u) / kTile);
    const uint32_t groups = (cout / 16u + NCT - 1u) / NCT;
    const dim3 grid(tiles, groups);
    const bool full = cout % (uint32_t(NCT) * 16u) == 0 && out.stride % 8u == 0 && out.off % 8u == 0;
    // fp32 output through the LDS stage: one output group, whole 16-byte rows
    const bool f32_full = NCT == 1 && cout == 16u && f32c <= 4u && d.h % kTile == 0 && d.w % kTile == 0 &&
                          (kTile * f32c) % 4u == 0 && reinterpret_cast<uintptr_t>(of32) % 16u == 0;
#define PT_CONV(...) hipLaunchKernelGGL((conv3x3_k<__VA_ARGS__>), grid, dim3(kBlock), 0, s, in, cin, w, bias, out, \
                                        cout, d, of32, f32c, fz)
#define PT_CONV_PF(RELU_, FULL_, UP_)                 \
    do {                                              \
        if (pf)                                       \
            PT_CONV(NCT, RELU_, false, FULL_, UP_, true); \
        else                                          \
            PT_CONV(NCT, RELU_, false, FULL_, UP_);   \
    } while (0)
    const bool pf = cin > kSlice && tiles * groups < kConvPfBlocks;
    if (fz.q_out)  // (checked by the caller: a full ReLU layer without coarse reads, pooling or upsampling)
        PT_CONV(NCT, true, false, true, false, false, true);
    else if (of32 && f32_full)
        PT_CONV(NCT, false, true, true);
    else if (of32)
        PT_CONV(NCT, false, true, false);
    else if (fz.up.p && full)
        PT_CONV_PF(true, true, true);
    else if (fz.up.p)
        PT_CONV_PF(true, false, true);  // the decoders' first convolutions have ReLU
    else if (relu && full)
        PT_CONV_PF(true, true, false);
    else if (relu)
        PT_CONV_PF(true, false, false);
    else if (full)
        PT_CONV_PF(false, true, false);
    else
        PT_CONV_PF(false, false, false);
#undef PT_CONV_PF
#undef PT_CONV
}

// ------------------------------------------------ conv over [up2(C) | S] --
// A decoder level's first convolution over [nearest 2x upsampling of the
// coarse map C (channels [0, cup), cup = 8 GC) | skip S (channels [cup, cin),
// 8 GS of them, at channels [0, 8 GS) of `in`)] for NCT x 16 output channels,
// 16 x 16 tiles of full size.  Wave w computes the output pixels of parity
// class (py, px) = (w / 2, w % 2): coarse positions Y, X = 0..7 of the tile,
// as 4 MFMA tiles of 16 (tile m: Y = 2m + ln / 8, X = ln % 8).  The coarse
// part takes the 4 merged taps of its class over a 10 x 10 coarse halo, the
// skip part the 9 taps over the 18 x 18 halo at the class's pixels, both into
// the same accumulators (started at the bias).  One staging round: both halos
// go to LDS together; the merged weights come from global memory already in
// MFMA fragment order (see up_merge_k), each wave reading its class's, and the
// skip weights from global memory as well.  Both parts run K over (tap,
// 8-channel group) pieces: K-step s gives lane group kq piece q = 4s + kq, tap
// q / G, group q % G (G = GC or GS), so no step carries empty channels but the
// skip part's last (9 GS a multiple of 4 or not).
constexpr uint32_t kUpHalo = 10, kUpHaloPix = kUpHalo * kUpHalo;

template <uint32_t GC, uint32_t GS, int NCT>
union UpSmem {
    struct {
        bf16x8 hc[GC * kUpHaloPix];  // [8-channel group][coarse pixel]
        bf16x8 hs[GS * kHaloPix];    // [8-channel group][pixel]
    } m;
    char stage[kTile * kTile * (uint32_t(NCT) * 32u + 16u)];
};

template <uint32_t GC, uint32_t GS, int NCT>
__global__ void __launch_bounds__(kBlock) conv_up_k(Tens in, uint32_t cin, const bf16* __restrict__ w,
                                                    const bf16x8* __restrict__ wup, const float* __restrict__ bias,
                                                    Tens out, Dims d, Tens up) {
    __shared__ UpSmem<GC, GS, NCT> sm;
    constexpr uint32_t cup = GC * 8u, kCo = uint32_t(NCT) * 16u;
    constexpr uint32_t kCP = GC * kUpHaloPix, kCPer = (kCP + kBlock - 1u) / kBlock;
    constexpr uint32_t kSP = GS * kHaloPix, kSPer = (kSP + kBlock - 1u) / kBlock;
    constexpr uint32_t kSkipSteps = (9u * GS + 3u) / 4u;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wv = tid >> 6, kq = lane >> 4, ln = lane & 15u;
    const uint32_t py = wv >> 1, px = wv & 1u;
    const uint32_t tiles_x = d.w / kTile, tiles_y = d.h / kTile;
    const uint32_t img = blockIdx.x / (tiles_x * tiles_y), trem = blockIdx.x % (tiles_x * tiles_y);
    const uint32_t ty0 = trem / tiles_x * kTile, tx0 = trem % tiles_x * kTile;
    const uint32_t H2 = d.h / 2u, W2 = d.w / 2u, Y0 = ty0 / 2u, X0 = tx0 / 2u;
    // every halo read of the block, issued together
    bf16x8 hcr[kCPer], hsr[kSPer];
#pragma unroll
    for (uint32_t k = 0; k < kCPer; ++k) {
        const uint32_t piece = tid + k * kBlock, cp = piece / GC, g = piece % GC;
        const uint32_t gy = Y0 + cp / kUpHalo - 1u, gx = X0 + cp % kUpHalo - 1u;  // wraps below 0
        hcr[k] = piece < kCP && gy < H2 && gx < W2
                     ? load8(up.p + (((img * H2 + gy) * W2 + gx) * up.stride + up.off + g * 8u))
                     : zero8();
    }
#pragma unroll
    for (uint32_t k = 0; k < kSPer; ++k) {
        const uint32_t piece = tid + k * kBlock, sp = piece / GS, g = piece % GS;
        const uint32_t gy = ty0 + sp / kHalo - 1u, gx = tx0 + sp % kHalo - 1u;
        hsr[k] = piece < kSP && gy < d.h && gx < d.w
                     ? load8(in.p + (((img * d.h + gy) * d.w + gx) * in.stride + in.off + g * 8u))
                     : zero8();
    }
    f32x4 acc[4][NCT];
#pragma unroll
    for (int t = 0; t < NCT; ++t) {
        float bv[4] = {0.f, 0.f, 0.f, 0.f};
        if (bias)
#pragma unroll
            for (int r = 0; r < 4; ++r) bv[r] = bias[uint32_t(t) * 16u + kq * 4u + uint32_t(r)];
#pragma unroll
        for (int m = 0; m < 4; ++m) acc[m][t] = f32x4{bv[0], bv[1], bv[2], bv[3]};
    }
#pragma unroll
    for (uint32_t k = 0; k < kCPer; ++k) {
        const uint32_t piece = tid + k * kBlock;
        if (piece < kCP) sm.m.hc[(piece % GC) * kUpHaloPix + piece / GC] = hcr[k];
    }
#pragma unroll
    for (uint32_t k = 0; k < kSPer; ++k) {
        const uint32_t piece = tid + k * kBlock;
        if (piece < kSP) sm.m.hs[(piece % GS) * kHaloPix + piece / GS] = hsr[k];
    }
    __syncthreads();
    // coarse part: the class's merged taps
    const bf16x8* wf = wup + size_t(wv) * GC * uint32_t(NCT) * 64u + lane;
#pragma unroll
    for (uint32_t s = 0; s < GC; ++s) {
        const uint32_t q = 4u * s + kq, ab = q / GC, g = q % GC, a = ab >> 1, b2 = ab & 1u;
        bf16x8 cf[4], bw[NCT];
#pragma unroll
        for (int t = 0; t < NCT; ++t) bw[t] = wf[(s * uint32_t(NCT) + uint32_t(t)) * 64u];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const uint32_t Y = 2u * uint32_t(m) + (ln >> 3), X = ln & 7u;
            cf[m] = sm.m.hc[g * kUpHaloPix + (Y + a + py) * kUpHalo + X + b2 + px];
        }
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int t = 0; t < NCT; ++t)
                acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bw[t], cf[m], acc[m][t], 0, 0, 0);
    }
    // skip part: 9 taps over the full-resolution halo, at the class's pixels
#pragma unroll
    for (uint32_t s = 0; s < kSkipSteps; ++s) {
        const uint32_t q = 4u * s + kq;
        const bool valid = q < 9u * GS;
        const uint32_t qa = valid ? q : 0u, tap = qa / GS, g = qa % GS, dy = tap / 3u, dx = tap % 3u;
        bf16x8 a[4], bw[NCT];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const uint32_t Y = 2u * uint32_t(m) + (ln >> 3), X = ln & 7u;
            a[m] = valid ? sm.m.hs[g * kHaloPix + (2u * Y + py + dy) * kHalo + 2u * X + px + dx] : zero8();
        }
#pragma unroll
        for (int t = 0; t < NCT; ++t)
            bw[t] = valid ? load8(w + ((tap * kCo + uint32_t(t) * 16u + ln) * cin + cup + g * 8u)) : zero8();
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int t = 0; t < NCT; ++t)
                acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bw[t], a[m], acc[m][t], 0, 0, 0);
    }
    // epilogue: ReLU, bf16, staged [pixel][channel] for 16-byte stores
    constexpr uint32_t kPitch = kCo * 2u + 16u;
    char* stage = sm.stage;
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NCT; ++t)
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            bf16x4 q;
#pragma unroll
            for (int r = 0; r < 4; ++r) q[r] = b(acc[m][t][r]);
            const i16x4 qi = __builtin_elementwise_max(__builtin_bit_cast(i16x4, q), i16x4{0, 0, 0, 0});
            const uint32_t Y = 2u * uint32_t(m) + (ln >> 3), X = ln & 7u;
            const uint32_t p = (2u * Y + py) * kTile + 2u * X + px;
            *reinterpret_cast<i16x4*>(stage + p * kPitch + (uint32_t(t) * 16u + kq * 4u) * 2u) = qi;
        }
    __syncthreads();
    constexpr uint32_t kChunks = kTile * kTile * kCo / 8u;
    const uint32_t obase = (img * d.h + ty0) * d.w + tx0;
#pragma unroll
    for (uint32_t i = tid; i < kChunks; i += kBlock) {
        const uint32_t p = i / (kCo / 8u), c8 = i % (kCo / 8u);
        const uint32_t op = obase + (p / kTile) * d.w + p % kTile;
        store8(out.p + (op * out.stride + out.off + c8 * 8u),
               *reinterpret_cast<const bf16x8*>(stage + p * kPitch + c8 * 16u));
    }
}

// The merged weights of the coarse part, for each parity class (py, px) and
// each of its 2 x 2 coarse taps (a, b): the sum of the 3 x 3 taps (dy, dx)
// that land on coarse offset (a + py - 1, b + px - 1) (floor((py + dy - 1) / 2)
// = a + py - 1, and likewise in x), summed in fp32 from the bf16 weights.
// Stored in MFMA fragment order: [class][K-step s][column tile t][lane][8],
// lane (kq, ln) holding output channel 16 t + ln, input channels 8 (q % G) ..
// + 7 of merged tap q / G, q = 4 s + kq, G = cup / 8; a wave reads one
// 16-byte piece per lane and step.
__global__ void __launch_bounds__(kBlock) up_merge_k(const bf16* __restrict__ w, uint32_t cout, uint32_t cin,
                                                     uint32_t cup, bf16* __restrict__ wup) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= 16u * cout * cup) return;
    const uint32_t G = cup / 8u, nt = cout / 16u;
    const uint32_t j = i % 8u, lane = (i / 8u) % 64u, t = (i / 512u) % nt, st = (i / (512u * nt)) % G;
    const uint32_t cls = i / (512u * nt * G);
    const uint32_t q = 4u * st + (lane >> 4), ab = q / G, g = q % G;
    const int py = int(cls >> 1), px = int(cls & 1u), a = int(ab >> 1), bq = int(ab & 1u);
    const uint32_t co = t * 16u + (lane & 15u), ci = g * 8u + j;
    float s = 0.0f;
    for (int dy = 0; dy < 3; ++dy) {
        if (((py + dy + 1) >> 1) - 1 != a + py - 1) continue;  // floor((py + dy - 1) / 2)
        for (int dx = 0; dx < 3; ++dx) {
            if (((px + dx + 1) >> 1) - 1 != bq + px - 1) continue;
            s += f(w[(uint32_t(dy * 3 + dx) * cout + co) * cin + ci]);
        }
    }
    wup[i] = b(s);
}

size_t up_merged_elems(uint32_t cout, uint32_t cup) { return size_t(16) * cout * cup; }

hipError_t up_merge_weights(const bf16* w, uint32_t cout, uint32_t cin, uint32_t cup, bf16* wup, hipStream_t s) {
    if (cup == 0 || cup > cin || cup % 8u || cout == 0 || cout % 16u) return hipErrorInvalidValue;
    const uint32_t n = 16u * cout * cup;
    hipLaunchKernelGGL(up_merge_k, dim3((n + kBlock - 1u) / kBlock), dim3(kBlock), 0, s, w, cout, cin, cup, wup);
    return hipGetLastError();
}

// The (coarse, skip, output) channel counts conv_up_k is built for: the
// level-0 and level-1 decoders' first convolutions (and a 32 + 32 form).
static bool conv_up_shape(uint32_t cup, uint32_t cskip, uint32_t cout) {
    return (cup == 48u && cskip == 32u && cout == 32u) || (cup == 32u && cskip == 32u && cout == 32u) ||
           (cup == 64u && cskip == 48u && cout == 48u);
}

static void conv_up_launch(const Tens& in, uint32_t cin, const bf16* w, const bf16* wup, const float* bias,
                           const Tens& out, uint32_t cout, Dims d, const Tens& up, uint32_t cup, hipStream_t s) {
    const dim3 grid(d.n * (d.h / kTile) * (d.w / kTile));
    const bf16x8* wf = reinterpret_cast<const bf16x8*>(wup);
    if (cup == 48u)
        hipLaunchKernelGGL((conv_up_k<6, 4, 2>), grid, dim3(kBlock), 0, s, in, cin, w, wf, bias, out, d, up);
    else if (cup == 32u)
        hipLaunchKernelGGL((conv_up_k<4, 4, 2>), grid, dim3(kBlock), 0, s, in, cin, w, wf, bias, out, d, up);
    else
        hipLaunchKernelGGL((conv_up_k<8, 6, 3>), grid, dim3(kBlock), 0, s, in, cin, w, wf, bias, out, d, up);
    (void)cout;
}

hipError_t conv3x3(const Tens& in, uint32_t cin, const bf16* w, const float* bias, const Tens& out, uint32_t cout,
                   Dims d, bool relu, float* out_f32, uint32_t f32_channels, hipStream_t s, const ConvFuse* fuse) {
    if (cin == 0 || cin % 16u || cout % 16u || cout == 0 || cout > kMaxCout || in.stride % 8u || in.off % 8u)
        return hipErrorInvalidValue;
    if (d.pixels() == 0) return hipSuccess;
    ConvFuse fz{};
    if (fuse) fz = *fuse;
    // the kernels index every tensor with 32-bit element offsets
    const uint64_t P = d.pixels(), lim = 1ull << 31;
    if (P * (in.stride + 8u) >= lim || P * (out.stride + 8u) >= lim || P * (f32_channels + 4u) >= lim ||
        4u * P * (fz.upw.stride + 8u) >= lim || P * (fz.up.stride + 8u) >= lim || P * (fz.pool.stride + 8u) >= lim)
        return hipErrorInvalidValue;
    // output-channel tiles per block: all of them up to 4, else groups of 3 or 4
    const uint32_t ct = cout / 16u;
    // (every output group restages the halo, so more, smaller groups pay only
    // where the launch has fewer blocks than CUs: then they occupy idle CUs)
    uint32_t nct = ct <= 4u ? ct : (ct % 4u == 0 ? 4u : (ct % 3u == 0 ? 3u : 4u));
    const uint32_t tiles = d.n * ((d.h + kTile - 1u) / kTile) * ((d.w + kTile - 1u) / kTile);
    while (nct > 1u && tiles * ((ct + nct - 1u) / nct) < 256u) {
        uint32_t n = nct - 1u;
        while (n > 1u && ct % n) --n;
        nct = n;
    }
    if (fz.fin && (!out_f32 || f32_channels != 3u || d.n != 1u || fz.fw > d.w || fz.fh > d.h))
        return hipErrorInvalidValue;
    // the frame texels come from the fp32 stage when the output conv takes that path
    const bool texel_fused = fz.fin && nct == 1u && cout == 16u && d.h % kTile == 0 && d.w % kTile == 0 &&
                             reinterpret_cast<uintptr_t>(out_f32) % 16u == 0;
    const ConvFuse frame = fz;
    if (!texel_fused) fz.fin = nullptr;
    // pooling and upsampling fuse into the staged bf16 epilogue (full output
    // groups); otherwise pool2 / up2 follow the convolution
    const bool staged = !out_f32 && cout % (nct * 16u) == 0 && out.stride % 8u == 0 && out.off % 8u == 0;
    if ((fz.pool.p && (d.h % 2u || d.w % 2u || fz.pool.stride % 8u || fz.pool.off % 8u)) ||
        (fz.upw.p && (out_f32 || fz.upw.stride % 8u || fz.upw.off % 8u)) ||
        (fz.up.p && (d.h % 2u || d.w % 2u || !relu || out_f32 || fz.cup % 8u || fz.cup > cin ||
                     fz.up.stride % 8u || fz.up.off % 8u)))
        return hipErrorInvalidValue;
    // e4m3 output (for an fp8 layer reading it): a full ReLU layer only
    if (fz.q_out && (!staged || !relu || fz.pool.p || fz.upw.p || fz.up.p || fz.in8 || out.stride % 16u ||
                     out.off % 16u || reinterpret_cast<uintptr_t>(fz.q_out) % 16u))
        return hipErrorInvalidValue;
    const Tens pool = fz.pool, upw = fz.upw;
    if (!staged) fz.pool = fz.upw = Tens{nullptr, 0, 0};
    // a decoder level's first convolution with merged weights: the coarse
    // part on the coarse map (full tiles, the shapes conv_up_shape lists)
    if (fz.up.p && fz.wup && staged && !pool.p && !upw.p && conv_up_shape(fz.cup, cin - fz.cup, cout) &&
        d.h % kTile == 0 && d.w % kTile == 0) {
        conv_up_launch(in, cin, w, fz.wup, bias, out, cout, d, fz.up, fz.cup, s);
        return hipGetLastError();
    }
    switch (nct) {
        case 1: conv_launch<1>(in, cin, w, bias, out, cout, d, relu, out_f32, f32_channels, fz, s); break;
        case 2: conv_launch<2>(in, cin, w, bias, out, cout, d, relu, out_f32, f32_channels, fz, s); break;
        case 3: conv_launch<3>(in, cin, w, bias, out, cout, d, relu, out_f32, f32_channels, fz, s); break;
        default: conv_launch<4>(in, cin, w, bias, out, cout, d, relu, out_f32, f32_channels, fz, s); break;
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && pool.p && !staged) e = pool2(out, cout, d, pool, s);
    if (e == hipSuccess && upw.p && !staged) e = up2(out, cout, d, upw, s);
    if (e == hipSuccess && frame.fin && !texel_fused)
        e = unpack_frame(out_f32, d, frame.fw, frame.fh, frame.exposure, frame.fin, frame.ldr, s);
    return e;
}

// ------------------------------------------------------------- conv fp8 --
// The inference variant on v_mfma_scale_f32_16x16x128_f8f6f4 (OCP e4m3
// operands, unit block scales, fp32 accumulation).  The input stays bf16 in
// memory and is quantized while it is staged: x * inv_scale[c], clamped to
// +-448 (v_cvt_pk_fp8_f32 turns larger magnitudes into NaN rather than
// saturating), rounded to nearest even.  The weights arrive quantized
// ([tap][cout][cin] e4m3 with the input scales folded in) with one scale per
// output channel, applied in the epilogue with the bias.
//
// Staged per slice of SL = 32 or 64 input channels (32 when the layer has
// no more): the halo as [16-channel group][pixel] pieces of 16 bytes, the
// weights as [slot][half][co].  A slot is one tap of one 32-channel group of
// the slice (9 or 18 per slice); a K-step of 128 is four slots, one per
// 16-lane group, so lane group g of step s holds slot 4s + g: 32 bytes = 32
// channels of that tap, in the same order for both operands (the instruction
// pairs A's and B's K positions the same way, checked on the device).  Three
// or five steps per slice, the last one partly empty (zero operands).

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

template <int NCT, int SL>
struct ConvSmem8 {
    i32x4 halo[(SL / 16) * kHaloPix];   // [16-channel group][pixel]
    i32x4 w[9 * (SL / 16) * NCT * 16];  // [slot][half][co]
};
// the same LDS reused for the output tile: [pixel][channel] bf16, rows padded by 16 bytes
template <int NCT, int SL>
union ConvLds8 {
    ConvSmem8<NCT, SL> m;
    char stage[kTile * kTile * (NCT * 32 + 16)];
};


// FULL (all NCT output tiles present) stages the bf16 output tile for 16-byte
// stores, those past the image's edge dropped, and can write its 2x2 average
// pool (fz.pool, as conv3x3_k).  UP: input channels below fz.cup come from
// the half-resolution fz.up at the nearest coarse pixel, the others from `in`
// (channels [0, cin - cup)), as in conv3x3_k.  IN8: `in` (and fz.up) hold
// e4m3 bytes already quantized with inv_scale (channel strides and offsets in
// bytes); the halo is staged as read.  OUT8 (FULL, bf16 output): the staged
// bf16 tile leaves as e4m3 bytes quantized with fz.q_out[co] (the consuming
// layer's inverse input scales), the pooled tile likewise with fz.q_pool when
// set: the bytes the consumer would compute from the bf16 values.
template <int NCT, int SL, bool RELU, bool F32OUT, bool FULL, bool UP = false, bool IN8 = false, bool OUT8 = false>
__global__ void __launch_bounds__(kBlock) conv3x3_f8_k(Tens in, uint32_t cin, const uint8_t* __restrict__ wq,
                                                       const float* __restrict__ inv_scale,
                                                       const float* __restrict__ wscale,
                                                       const float* __restrict__ bias, Tens out, uint32_t cout, Dims d,
                                                       float* __restrict__ of32, uint32_t f32c, ConvFuse fz) {
    __shared__ ConvLds8<NCT, SL> lds;
    ConvSmem8<NCT, SL>& sm = lds.m;
    constexpr uint32_t kG16 = uint32_t(SL) / 16u, kG8 = uint32_t(SL) / 8u;  // 16- and 8-channel groups per slice
    constexpr uint32_t kSlots = 9u * (uint32_t(SL) / 32u), kSteps = (kSlots + 3u) / 4u;
    constexpr uint32_t kHalo8Pieces = kHaloPix * kG8;                        // 8-channel pieces of the halo
    constexpr uint32_t kCo = uint32_t(NCT) * 16u;
    constexpr uint32_t kWPieces = 9u * kCo * kG16;                           // 16-byte pieces: (tap, co, 16-ch group)
    constexpr uint32_t kHaloPer = (kHalo8Pieces + kBlock - 1u) / kBlock;
    constexpr uint32_t kWPer = (kWPieces + kBlock - 1u) / kBlock;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wv = tid >> 6, kq = lane >> 4;
    const uint32_t tiles_x = (d.w + kTile - 1u) / kTile, tiles_y = (d.h + kTile - 1u) / kTile;
    const uint32_t img = blockIdx.x / (tiles_x * tiles_y), trem = blockIdx.x % (tiles_x * tiles_y);
    const int ty0 = int(trem / tiles_x * kTile), tx0 = int(trem % tiles_x * kTile);
    const uint32_t co0 = blockIdx.y * kCo;
    const uint32_t nct = FULL ? uint32_t(NCT) : min(uint32_t(NCT), (cout - co0) / 16u);
    // halo piece k of this thread: pixel piece / 8, channels (piece % 8) * 8 .. + 7 of the slice
    constexpr uint32_t kE = IN8 ? 1u : 2u;  // bytes per staged input element in memory
    // byte offsets of this thread's halo pieces (32-bit: checked by the host)
    const char* inb = reinterpret_cast<const char*>(in.p);
    const char* upb = reinterpret_cast<const char*>(fz.up.p);
    uint32_t hsrc[kHaloPer];
    uint32_t csrc[UP ? kHaloPer : 1];
    bool hval[kHaloPer];
#pragma unroll
    for (uint32_t k = 0; k < kHaloPer; ++k) {
        const uint32_t piece = tid + k * kBlock;
        const uint32_t j = piece / kG8;
        const int gy = ty0 + int(j / kHalo) - 1, gx = tx0 + int(j % kHalo) - 1;
        hval[k] = piece < kHalo8Pieces && gy >= 0 && gy < int(d.h) && gx >= 0 && gx < int(d.w);
        const uint32_t sp = hval[k] ? (img * d.h + uint32_t(gy)) * d.w + uint32_t(gx) : 0u;
        hsrc[k] = (sp * in.stride + in.off + (piece % kG8) * 8u) * kE;
        if (UP) {
            const uint32_t cp = hval[k] ? (img * (d.h / 2u) + uint32_t(gy) / 2u) * (d.w / 2u) + uint32_t(gx) / 2u : 0u;
            csrc[UP ? k : 0] = (cp * fz.up.stride + fz.up.off + (piece % kG8) * 8u) * kE;
        }
    }
    u32x2 hreg[kHaloPer];
    i32x4 wreg[kWPer];
    auto fetch = [&](uint32_t c0) {
#pragma unroll
        for (uint32_t k = 0; k < kHaloPer; ++k) {
            const uint32_t c = c0 + ((tid + k * kBlock) % kG8) * 8u;
            hreg[k] = u32x2{0u, 0u};
            if (hval[k] && c < cin) {
                // (the skip offset wraps modulo 2^32 to the right channel)
                const char* src = UP && c < fz.cup ? upb + (csrc[UP ? k : 0] + c0 * kE)
                                                   : inb + (hsrc[k] + (c0 - (UP ? fz.cup : 0u)) * kE);
                if (IN8) {
                    hreg[k] = *reinterpret_cast<const u32x2*>(src);
                    continue;
                }
                const bf16x8 v = load8(reinterpret_cast<const bf16*>(src));
                const f32x4 s0 = *reinterpret_cast<const f32x4*>(inv_scale + c);
                const f32x4 s1 = *reinterpret_cast<const f32x4*>(inv_scale + c + 4u);
                hreg[k] = u32x2{e4m3x4(f(v[0]) * s0[0], f(v[1]) * s0[1], f(v[2]) * s0[2], f(v[3]) * s0[3]),
                                e4m3x4(f(v[4]) * s1[0], f(v[5]) * s1[1], f(v[6]) * s1[2], f(v[7]) * s1[3])};
            }
        }
#pragma unroll
        for (uint32_t k = 0; k < kWPer; ++k) {
            const uint32_t piece = tid + k * kBlock;
            const uint32_t g16 = piece % kG16, co = (piece / kG16) % kCo, tap = piece / (kG16 * kCo);
            const uint32_t c = c0 + g16 * 16u;
            const bool ok = piece < kWPieces && co0 + co < cout && c < cin;
            wreg[k] = ok ? *reinterpret_cast<const i32x4*>(wq + (size_t(tap) * cout + co0 + co) * cin + c)
                         : i32x4{0, 0, 0, 0};
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (uint32_t k = 0; k < kHaloPer; ++k) {
            const uint32_t piece = tid + k * kBlock;
            if (piece < kHalo8Pieces) {
                const uint32_t e = piece % kG8;  // 8-channel group: 16-channel group e / 2, half e % 2
                u32x2* dst = reinterpret_cast<u32x2*>(&sm.halo[(e / 2u) * kHaloPix + piece / kG8]) + (e % 2u);
                *dst = hreg[k];
            }
        }
#pragma unroll
        for (uint32_t k = 0; k < kWPer; ++k) {
            const uint32_t piece = tid + k * kBlock;
            if (piece < kWPieces) {
                const uint32_t g16 = piece % kG16, co = (piece / kG16) % kCo, tap = piece / (kG16 * kCo);
                const uint32_t slot = tap * (kG16 / 2u) + g16 / 2u;
                sm.w[(slot * 2u + g16 % 2u) * kCo + co] = wreg[k];
            }
        }
    };
    f32x4 acc[4][NCT];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int t = 0; t < NCT; ++t) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const i32x8 zero = {0, 0, 0, 0, 0, 0, 0, 0};
    // slices are fetched, staged and multiplied in turn (as conv3x3_k: the
    // staging registers are free during the products)
    for (uint32_t c0 = 0; c0 < cin; c0 += uint32_t(SL)) {
        fetch(c0);
        __syncthreads();  // the previous slice's reads are done
        stash();
        __syncthreads();
#pragma unroll
        for (uint32_t s = 0; s < kSteps; ++s) {
            const uint32_t q = s * 4u + kq;                 // this lane group's slot
            const bool valid = q < kSlots;
            const uint32_t qa = valid ? q : 0u;             // address a real slot either way
            const uint32_t tap = qa / (kG16 / 2u), g32 = qa % (kG16 / 2u);
            const uint32_t dy = tap / 3u, dx = tap % 3u;
            i32x8 a[4], bw[NCT];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const uint32_t hp = (wv * 4u + uint32_t(m) + dy) * kHalo + (lane & 15u) + dx;
                const i32x4 lo = sm.halo[(2u * g32) * kHaloPix + hp], hi = sm.halo[(2u * g32 + 1u) * kHaloPix + hp];
                a[m] = valid ? __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7) : zero;
            }
#pragma unroll
            for (int t = 0; t < NCT; ++t) {
                const uint32_t co = uint32_t(t) * 16u + (lane & 15u);
                const i32x4 lo = sm.w[(qa * 2u) * kCo + co], hi = sm.w[(qa * 2u + 1u) * kCo + co];
                bw[t] = valid ? __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7) : zero;
            }
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int t = 0; t < NCT; ++t)
                    if (uint32_t(t) < nct)
                        acc[m][t] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a[m], bw[t], acc[m][t], 0, 0, 0,
                                                                                     127, 0, 127);
        }
    }
    if (FULL && !F32OUT) {  // the bf16 tile through LDS and out as 16-byte stores, as in conv3x3_k
        constexpr uint32_t kPitch = kCo * 2u + 16u;
        static_assert(kTile * kTile * kPitch <= sizeof(lds.stage), "output stage fits the LDS tile");
        char* stage = lds.stage;
        __syncthreads();
#pragma unroll
        for (int t = 0; t < NCT; ++t) {
            const uint32_t ch = uint32_t(t) * 16u + (lane & 15u);
            const float sc = wscale[co0 + ch], bv = bias ? bias[co0 + ch] : 0.0f;
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    float v = fmaf(acc[m][t][r], sc, bv);
                    if (RELU) v = fmaxf(v, 0.0f);
                    const uint32_t p = (wv * 4u + uint32_t(m)) * kTile + kq * 4u + uint32_t(r);
                    *reinterpret_cast<bf16*>(stage + p * kPitch + ch * 2u) = b(v);
                }
        }
        __syncthreads();
        if (OUT8) {  // 16 channels of a pixel per 16-byte store
            constexpr uint32_t kC16 = kCo / 16u;
#pragma unroll
            for (uint32_t i = tid; i < kTile * kTile * kC16; i += kBlock) {
                const uint32_t p = i / kC16, c16 = i % kC16;
                if (uint32_t(ty0) + p / kTile >= d.h || uint32_t(tx0) + p % kTile >= d.w) continue;
                const size_t op = (size_t(img) * d.h + uint32_t(ty0) + p / kTile) * d.w + uint32_t(tx0) + p % kTile;
                const bf16x8 v0 = *reinterpret_cast<const bf16x8*>(stage + p * kPitch + c16 * 32u);
                const bf16x8 v1 = *reinterpret_cast<const bf16x8*>(stage + p * kPitch + c16 * 32u + 16u);
                const f32x4* sc = reinterpret_cast<const f32x4*>(fz.q_out + co0 + c16 * 16u);
                const f32x4 s0 = sc[0], s1 = sc[1], s2 = sc[2], s3 = sc[3];
                const u32x4 q{e4m3x4(f(v0[0]) * s0[0], f(v0[1]) * s0[1], f(v0[2]) * s0[2], f(v0[3]) * s0[3]),
                              e4m3x4(f(v0[4]) * s1[0], f(v0[5]) * s1[1], f(v0[6]) * s1[2], f(v0[7]) * s1[3]),
                              e4m3x4(f(v1[0]) * s2[0], f(v1[1]) * s2[1], f(v1[2]) * s2[2], f(v1[3]) * s2[3]),
                              e4m3x4(f(v1[4]) * s3[0], f(v1[5]) * s3[1], f(v1[6]) * s3[2], f(v1[7]) * s3[3])};
                *reinterpret_cast<u32x4*>(reinterpret_cast<uint8_t*>(out.p) + op * out.stride + out.off + co0 +
                                          c16 * 16u) = q;
            }
        } else {
            constexpr uint32_t kChunks = kTile * kTile * kCo / 8u;
#pragma unroll
            for (uint32_t i = tid; i < kChunks; i += kBlock) {
                const uint32_t p = i / (kCo / 8u), c8 = i % (kCo / 8u);
                if (uint32_t(ty0) + p / kTile >= d.h || uint32_t(tx0) + p % kTile >= d.w) continue;
                const size_t op = (size_t(img) * d.h + uint32_t(ty0) + p / kTile) * d.w + uint32_t(tx0) + p % kTile;
                store8(out.p + op * out.stride + out.off + co0 + c8 * 8u,
                       *reinterpret_cast<const bf16x8*>(stage + p * kPitch + c8 * 16u));
            }
        }
        if (fz.pool.p) {  // the 8 x 8 pooled tile, summed in pool2's order
            constexpr uint32_t kQ = kTile / 2u;
            for (uint32_t i = tid; i < kQ * kQ * (kCo / 8u); i += kBlock) {
                const uint32_t q = i / (kCo / 8u), c8 = i % (kCo / 8u), qy = q / kQ, qx = q % kQ;
                if (uint32_t(ty0) + 2u * qy >= d.h || uint32_t(tx0) + 2u * qx >= d.w) continue;  // (d.h, d.w even)
                const uint32_t p00 = 2u * qy * kTile + 2u * qx;
                const char* sb = stage + c8 * 16u;
                const bf16x8 v0 = *reinterpret_cast<const bf16x8*>(sb + p00 * kPitch);
                const bf16x8 v1 = *reinterpret_cast<const bf16x8*>(sb + (p00 + 1u) * kPitch);
                const bf16x8 v2 = *reinterpret_cast<const bf16x8*>(sb + (p00 + kTile) * kPitch);
                const bf16x8 v3 = *reinterpret_cast<const bf16x8*>(sb + (p00 + kTile + 1u) * kPitch);
                bf16x8 o;
#pragma unroll
                for (int k = 0; k < 8; ++k) o[k] = b(0.25f * (f(v0[k]) + f(v1[k]) + f(v2[k]) + f(v3[k])));
                const size_t cp = (size_t(img) * (d.h / 2u) + uint32_t(ty0) / 2u + qy) * (d.w / 2u) + uint32_t(tx0) / 2u + qx;
                if (OUT8 && fz.q_pool) {  // e4m3 with the pooled map's consumer's scales
                    const f32x4* sc = reinterpret_cast<const f32x4*>(fz.q_pool + co0 + c8 * 8u);
                    const f32x4 s0 = sc[0], s1 = sc[1];
                    const u32x2 q{e4m3x4(f(o[0]) * s0[0], f(o[1]) * s0[1], f(o[2]) * s0[2], f(o[3]) * s0[3]),
                                  e4m3x4(f(o[4]) * s1[0], f(o[5]) * s1[1], f(o[6]) * s1[2], f(o[7]) * s1[3])};
                    *reinterpret_cast<u32x2*>(reinterpret_cast<uint8_t*>(fz.pool.p) + cp * fz.pool.stride +
                                              fz.pool.off + co0 + c8 * 8u) = q;
                } else {
                    store8(fz.pool.p + (cp * fz.pool.stride + fz.pool.off + co0 + c8 * 8u), o);
                }
            }
        }
        return;
    }
#pragma unroll
    for (int t = 0; t < NCT; ++t) {
        if (uint32_t(t) >= nct) continue;
        const uint32_t co = co0 + uint32_t(t) * 16u + (lane & 15u);
        const float sc = wscale[co], bv = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int y = ty0 + int(wv * 4u) + m;
            if (y >= int(d.h)) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int x = tx0 + int(kq * 4u) + r;
                if (x >= int(d.w)) continue;
                const size_t op = (size_t(img) * d.h + uint32_t(y)) * d.w + uint32_t(x);
                float v = fmaf(acc[m][t][r], sc, bv);
                if (RELU) v = fmaxf(v, 0.0f);
                if (F32OUT) {
                    if (co < f32c) of32[op * f32c + co] = v;
                } else {
                    out.p[op * out.stride + out.off + co] = b(v);
                }
            }
        }
    }
}

template <int NCT, int SL>
static void conv8_launch_sl(const Tens& in, uint32_t cin, const uint8_t* w, const float* inv_scale,
                            const float* wscale, const float* bias, const Tens& out, uint32_t cout, Dims d, bool relu,
                            float* of32, uint32_t f32c, bool full, const ConvFuse& fz, hipStream_t s) {
    const uint32_t tiles = d.n * ((d.h + kTile - 1u) / kTile) * ((d.w + kTile - 1u) / kTile);
    const uint32_t groups = (cout / 16u + NCT - 1u) / NCT;
    const dim3 grid(tiles, groups);
#define PT_CONV8(RELU_, F32_, FULL_, UP_, ...)                                                                   \
    hipLaunchKernelGGL((conv3x3_f8_k<NCT, SL, RELU_, F32_, FULL_, UP_, ##__VA_ARGS__>), grid, dim3(kBlock), 0, s, in, \
                       cin, w, inv_scale, wscale, bias, out, cout, d, of32, f32c, fz)
    // the e4m3-in-memory forms (full ReLU layers): e4m3 in and out; e4m3 in
    // with coarse reads and bf16 out (the level-0 decoder); bf16 in, e4m3 out
    if (fz.in8 || fz.q_out) {
        if (fz.up.p && fz.q_out)
            PT_CONV8(true, false, true, true, true, true);
        else if (fz.up.p)
            PT_CONV8(true, false, true, true, true, false);
        else if (fz.in8)
            PT_CONV8(true, false, true, false, true, true);
        else
            PT_CONV8(true, false, true, false, false, true);
        return;
    }
    if (of32)
        PT_CONV8(false, true, false, false);
    else if (fz.up.p && full)
        PT_CONV8(true, false, true, true);
    else if (fz.up.p)
        PT_CONV8(true, false, false, true);
    else if (relu && full)
        PT_CONV8(true, false, true, false);
    else if (relu)
        PT_CONV8(true, false, false, false);
    else if (full)
        PT_CONV8(false, false, true, false);
    else
        PT_CONV8(false, false, false, false);
#undef PT_CONV8
}

template <int NCT>
static void conv8_launch(const Tens& in, uint32_t cin, const uint8_t* w, const float* inv_scale, const float* wscale,
                         const float* bias, const Tens& out, uint32_t cout, Dims d, bool relu, float* of32,
                         uint32_t f32c, bool full, const ConvFuse& fz, hipStream_t s) {
    // 32-channel slices also where 64 would leave most of the last slice
    // empty (80 = 32 + 32 + 16 channels: 9 K-steps of 128 instead of 10)
    if (cin <= 32u || cin % 64u == 16u)
        conv8_launch_sl<NCT, 32>(in, cin, w, inv_scale, wscale, bias, out, cout, d, relu, of32, f32c, full, fz, s);
    else
        conv8_launch_sl<NCT, 64>(in, cin, w, inv_scale, wscale, bias, out, cout, d, relu, of32, f32c, full, fz, s);
}

hipError_t conv3x3_fp8(const Tens& in, uint32_t cin, const uint8_t* w, const float* inv_scale, const float* wscale,
                       const float* bias, const Tens& out, uint32_t cout, Dims d, bool relu, float* out_f32,
                       uint32_t f32_channels, hipStream_t s, const ConvFuse* fuse) {
    if (cin == 0 || cin % 16u || cout % 16u || cout == 0 || cout > kMaxCout || in.stride % 8u || in.off % 8u ||
        !w || !inv_scale || !wscale || (reinterpret_cast<uintptr_t>(w) | reinterpret_cast<uintptr_t>(inv_scale)) % 16u)
        return hipErrorInvalidValue;
    ConvFuse fz{};
    if (fuse) fz = *fuse;
    fz.fin = nullptr;  // (no frame texels: the output convolution stays bf16)
    if (fz.up.p && (out_f32 || !relu || fz.cup == 0 || fz.cup % 16u || fz.cup >= cin || fz.up.stride % 8u ||
                    fz.up.off % 8u || d.h % 2u || d.w % 2u))
        return hipErrorInvalidValue;
    // the e4m3 forms: ReLU layers with bf16-free full output groups; an e4m3
    // input without coarse reads also has e4m3 output, e4m3 pooling needs
    // e4m3 output
    const bool e8 = fz.in8 || fz.q_out;
    if (e8 && (out_f32 || !relu || fz.upw.p || (fz.in8 && !fz.up.p && !fz.q_out) || (fz.up.p && !fz.in8) ||
               (fz.q_pool && !fz.q_out) || (fz.pool.p && fz.q_out && !fz.q_pool) ||
               (fz.q_out && (out.stride % 16u || out.off % 16u || reinterpret_cast<uintptr_t>(fz.q_out) % 16u)) ||
               (fz.q_pool && reinterpret_cast<uintptr_t>(fz.q_pool) % 16u)))
        return hipErrorInvalidValue;
    if (d.pixels() == 0) return hipSuccess;
    // the kernel's halo offsets are 32-bit bytes
    const uint64_t P = d.pixels(), lim = 1ull << 31;
    if (2u * P * (in.stride + 8u) >= lim || 2u * P * (fz.up.stride + 8u) >= lim) return hipErrorInvalidValue;
    const uint32_t ct = cout / 16u;
    uint32_t nct = ct <= 4u ? ct : (ct % 4u == 0 ? 4u : (ct % 3u == 0 ? 3u : 4u));
    // a launch of fewer blocks than CUs takes smaller output-channel groups
    const uint32_t tiles = d.n * ((d.h + kTile - 1u) / kTile) * ((d.w + kTile - 1u) / kTile);
    while (nct > 1u && tiles * ((ct + nct - 1u) / nct) < 256u) {
        uint32_t n = nct - 1u;
        while (n > 1u && ct % n) --n;
        nct = n;
    }
    const bool full = !out_f32 && cout % (nct * 16u) == 0 && out.stride % 8u == 0 && out.off % 8u == 0;
    if (e8 && (!full || (fz.pool.p && (d.h % 2u || d.w % 2u || fz.pool.stride % 8u || fz.pool.off % 8u))))
        return hipErrorInvalidValue;  // (no separate pass can take e4m3 data)
    const Tens pool = fz.pool;
    const Tens upw = fz.upw;
    if (!full || d.h % 2u || d.w % 2u || pool.stride % 8u || pool.off % 8u) fz.pool = Tens{nullptr, 0, 0};
    fz.upw = Tens{nullptr, 0, 0};
    switch (nct) {
        case 1: conv8_launch<1>(in, cin, w, inv_scale, wscale, bias, out, cout, d, relu, out_f32, f32_channels, full, fz, s); break;
        case 2: conv8_launch<2>(in, cin, w, inv_scale, wscale, bias, out, cout, d, relu, out_f32, f32_channels, full, fz, s); break;
        case 3: conv8_launch<3>(in, cin, w, inv_scale, wscale, bias, out, cout, d, relu, out_f32, f32_channels, full, fz, s); break;
        default: conv8_launch<4>(in, cin, w, inv_scale, wscale, bias, out, cout, d, relu, out_f32, f32_channels, full, fz, s); break;
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && pool.p && !fz.pool.p) e = pool2(out, cout, d, pool, s);
    if (e == hipSuccess && upw.p) e = up2(out, cout, d, upw, s);
    return e;
}

// amax[ch] = max(amax[ch], max over pixels of |x[p][ch]|), as float bits (a
// non-negative float orders like its bit pattern); per block in LDS, then one
// device atomic per channel.
constexpr uint32_t kMaxAmaxC = 256;

__global__ void __launch_bounds__(kBlock) chan_amax_k(Tens x, uint32_t c, uint32_t P, uint32_t* __restrict__ amax) {
    __shared__ uint32_t sm[kMaxAmaxC];
    for (uint32_t i = threadIdx.x; i < c; i += kBlock) sm[i] = 0u;
    __syncthreads();
    const uint32_t groups = c / 8u;
    const size_t n = size_t(P) * groups;
    for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += size_t(gridDim.x) * kBlock) {
        const size_t p = i / groups;
        const uint32_t g = uint32_t(i % groups);
        const bf16x8 v = load8(x.p + p * x.stride + x.off + g * 8u);
#pragma unroll
        for (int j = 0; j < 8; ++j) atomicMax(&sm[g * 8u + uint32_t(j)], __float_as_uint(fabsf(f(v[j]))));
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < c; i += kBlock) atomicMax(&amax[i], sm[i]);
}

hipError_t chan_amax(const Tens& x, uint32_t c, uint32_t P, uint32_t* amax, hipStream_t s) {
    if (c == 0 || c % 8u || c > kMaxAmaxC || x.stride % 8u || x.off % 8u) return hipErrorInvalidValue;
    if (P == 0) return hipSuccess;
    const size_t n = size_t(P) * (c / 8u);
    const uint32_t blocks = uint32_t(std::min<size_t>(1024, (n + kBlock - 1u) / kBlock));
    hipLaunchKernelGGL(chan_amax_k, dim3(blocks), dim3(kBlock), 0, s, x, c, P, amax);
    return hipGetLastError();
}

// Per-channel histograms of |x| over [0, amax[ch]] in kHistBins equal bins
// (values at amax land in the last bin): one workgroup per 8-channel group
// and pixel chunk, privatised in LDS, then added to hist[ch][bin].
__global__ void __launch_bounds__(kBlock) chan_hist_k(Tens x, uint32_t P, uint32_t chunk,
                                                      const uint32_t* __restrict__ amax, uint32_t* __restrict__ hist) {
    __shared__ uint32_t sm[8 * kHistBins];
    for (uint32_t i = threadIdx.x; i < 8u * kHistBins; i += kBlock) sm[i] = 0u;
    __syncthreads();
    const uint32_t g = blockIdx.y;
    float scale[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const float a = __uint_as_float(amax[g * 8u + uint32_t(j)]);
        scale[j] = a > 0.0f ? float(kHistBins) / a : 0.0f;
    }
    const uint32_t p0 = blockIdx.x * chunk, p1 = min(P, p0 + chunk);
    for (uint32_t p = p0 + threadIdx.x; p < p1; p += kBlock) {
        const bf16x8 v = load8(x.p + size_t(p) * x.stride + x.off + g * 8u);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uint32_t b = min(kHistBins - 1u, uint32_t(fabsf(f(v[j])) * scale[j]));
            atomicAdd(&sm[uint32_t(j) * kHistBins + b], 1u);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < 8u * kHistBins; i += kBlock)
        if (sm[i]) atomicAdd(&hist[size_t(g * 8u + i / kHistBins) * kHistBins + i % kHistBins], sm[i]);
}

hipError_t chan_hist(const Tens& x, uint32_t c, uint32_t P, const uint32_t* amax, uint32_t* hist, hipStream_t s) {
    if (c == 0 || c % 8u || x.stride % 8u || x.off % 8u) return hipErrorInvalidValue;
    if (P == 0) return hipSuccess;
    const uint32_t chunk = 8192;
    hipLaunchKernelGGL(chan_hist_k, dim3((P + chunk - 1) / chunk, c / 8u), dim3(kBlock), 0, s, x, P, chunk, amax, hist);
    return hipGetLastError();
}

// ------------------------------------------------------------ reductions --
// Deterministic sums over chunks: out[row * out_cols + col_off + col] =
// sum over k < chunks of partial[k * n + row * cols + col] for the first nw
// of the n elements, out2[i - nw] for the rest.  Pass 1 sums segments of at
// most kRedSeg chunks into tmp[segment][n]; pass 2 sums the segments.
constexpr uint32_t kRedSeg = 32;

__global__ void __launch_bounds__(kBlock) reduce_seg_k(const float* __restrict__ partial, uint32_t chunks, uint32_t n,
                                                       float* __restrict__ tmp) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t k0 = blockIdx.y * kRedSeg, k1 = min(k0 + kRedSeg, chunks);
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
    uint32_t k = k0;
    for (; k + 4u <= k1; k += 4u) {
        s0 += partial[size_t(k) * n + i];
        s1 += partial[size_t(k + 1u) * n + i];
        s2 += partial[size_t(k + 2u) * n + i];
        s3 += partial[size_t(k + 3u) * n + i];
    }
    for (; k < k1; ++k) s0 += partial[size_t(k) * n + i];
    tmp[size_t(blockIdx.y) * n + i] = (s0 + s1) + (s2 + s3);
}

__global__ void __launch_bounds__(kBlock) reduce_fin_k(const float* __restrict__ tmp, uint32_t segs, uint32_t n,
                                                       uint32_t nw, uint32_t cols, uint32_t out_cols, uint32_t col_off,
                                                       float* __restrict__ out, float* __restrict__ out2) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
    uint32_t k = 0;
    for (; k + 4u <= segs; k += 4u) {
        s0 += tmp[size_t(k) * n + i];
        s1 += tmp[size_t(k + 1u) * n + i];
        s2 += tmp[size_t(k + 2u) * n + i];
        s3 += tmp[size_t(k + 3u) * n + i];
    }
    for (; k < segs; ++k) s0 += tmp[size_t(k) * n + i];
    const float v = (s0 + s1) + (s2 + s3);
    if (i >= nw) {
        out2[i - nw] = v;
        return;
    }
    const uint32_t row = i / cols, col = i - row * cols;
    out[size_t(row) * out_cols + col_off + col] = v;
}

static uint32_t red_segs(uint32_t chunks) { return (chunks + kRedSeg - 1u) / kRedSeg; }

static hipError_t reduce_chunks(const float* partial, uint32_t chunks, uint32_t n, uint32_t nw, uint32_t cols,
                                uint32_t out_cols, uint32_t col_off, float* tmp, float* out, float* out2,
                                hipStream_t s) {
    const uint32_t segs = red_segs(chunks);
    hipLaunchKernelGGL(reduce_seg_k, dim3((n + kBlock - 1u) / kBlock, segs), dim3(kBlock), 0, s, partial, chunks, n,
                       tmp);
    hipLaunchKernelGGL(reduce_fin_k, dim3((n + kBlock - 1u) / kBlock), dim3(kBlock), 0, s, tmp, segs, n, nw, cols,
                       out_cols, col_off, out, out2);
    return hipGetLastError();
}

// ---------------------------------------------------------------- wgrad --
// Backward of a 3x3 convolution except the data gradient, in one kernel:
// dz = ga * (a > 0) (ReLU layers) or ga, dW[tap][co][ci] = sum_p dz[p][co] *
// x[p + offset(tap)][ci] (a GEMM over pixels), db[co] = sum_p dz[p][co], and
// dz written NHWC for the data gradient.  All operands are read NHWC.
//
// A block walks tiles of 32 * steps pixels (th rows x tw columns of one
// image, tw = 8, 16 or 32).  Per tile it stages dz (pixel-major, cout
// channels contiguous) and the (th + 2) x (tw + 2) halo of x (the block's
// input-channel span contiguous, zero outside the image) in LDS: every
// thread issues all its global loads of the tile before its LDS stores (a
// thread's pieces, <= kWgPieces per operand, are mapped once per launch).
// The MFMA fragments come from ds_read_b64_tr_b16.  That instruction,
// checked on the device: in each 16-lane group, lane 4e + g (e, g = 0..3)
// reads 4 bf16 at its own address and lane 4g + r receives element r of
// lanes 4e + g for e = 0..3.  So when lane L reads channels 4 (L % 4) .. +3
// of pixel 16h + L / 4, lane 16b + i holds channel i of pixels 16h + 4b + e,
// e = 0..3: two reads (h = 0, 1) give the 8 k-values 8b + 4h + e of an
// operand fragment, with k <-> pixel 32s + 16h + 4b + e (K step s) the same
// for dz and x.  A tap's shift is an address offset into the halo.  The
// bias sums are the product with a fragment of ones.
//
// Fragments: B fragment f = 9 * (16-channel input tile) + tap.  Block group
// blockIdx.y owns fragments [f0, f1) (a contiguous span of input tiles is
// staged) and wave w of it f0 + w + 4j; every wave takes all cout / 16
// output tiles; group 0 also writes dz and its wave 0 sums the bias.
// blockIdx.x is a chunk of consecutive tiles (down a column of tiles, so
// halo rows come again from the cache); its partial sums go to
// partial[chunk] and reduce_chunks adds the chunks in order.  LDS pixel
// pitches are odd multiples of 32 bytes: the 4 pixels of a 16-lane group then
// fall in distinct banks.
typedef __attribute__((address_space(3))) char lds_char;
typedef __attribute__((address_space(3))) i16x4 lds_i16x4;
typedef __attribute__((address_space(3))) bf16x8 lds_bf16x8;

constexpr uint32_t kWgLds = 65536;   // LDS per block (2 blocks per CU)
constexpr uint32_t kWgBlocks = 512;  // blocks per launch to aim for
constexpr int kWgPieces = 8;         // 16-byte pieces per thread and operand per tile
constexpr int kWgBatch = 4;          // pieces per thread in flight at once

struct WgGeom {
    uint32_t tw, th, steps;            // tile: th x tw pixels = 32 * steps
    uint32_t tiles_x, tiles_y, tiles;  // per image row, per image column, total
    uint32_t per_chunk;                // tiles per block
    uint32_t frags, per_group;         // B fragments in all, per block group
    uint32_t dpitch;                   // dz tile pixel pitch, bytes
};

__device__ __forceinline__ bf16x8 tr_frag(lds_char* p, uint32_t second) {
    const i16x4 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_i16x4*)p);
    const i16x4 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_i16x4*)(p + second));
    return __builtin_bit_cast(bf16x8, __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7));
}

// Piece k of thread t in a [rows][cols][c8] tile: (row, col, group) packed
// in bytes 0, 1, 2; ~0u past the end.
__device__ __forceinline__ uint32_t piece_of(uint32_t i, uint32_t rows, uint32_t cols, uint32_t c8) {
    if (i >= rows * cols * c8) return ~0u;
    const uint32_t q = i / c8, cg = i - q * c8;
    const uint32_t r = q / cols, col = q - r * cols;
    return r | (col << 8) | (cg << 16);
}

template <int MTX, int NB>
__global__ void __launch_bounds__(kBlock) wgrad_k(Tens ga, Tens a, uint32_t cout, Tens x, uint32_t cin, Dims d,
                                                  WgGeom g, bf16* __restrict__ dz_out, bool bias,
                                                  float* __restrict__ partial) {
    extern __shared__ __align__(16) char wg_smem[];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const uint32_t mt = cout / 16u;
    const uint32_t f0 = blockIdx.y * g.per_group, f1 = min(f0 + g.per_group, g.frags);
    const uint32_t ct0 = f0 / 9u, span = (f1 - 1u) / 9u - ct0 + 1u;
    const uint32_t xpitch = 32u * (span | 1u);
    const uint32_t hw = g.tw + 2u, hpix = (g.th + 2u) * hw;
    lds_char* xs = (lds_char*)wg_smem;
    lds_char* dzs = xs + hpix * xpitch;
    const uint32_t nb = f0 + wv < f1 ? (f1 - f0 - wv + 3u) / 4u : 0u;
    const bool lead = blockIdx.y == 0;
    const bool do_bias = bias && lead && wv == 0;
    uint32_t boff[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const uint32_t fj = f0 + wv + 4u * uint32_t(j), tap = fj % 9u;
        boff[j] = ((tap / 3u) * hw + tap % 3u) * xpitch + (fj / 9u - ct0) * 32u;
    }
    const uint32_t ql = lane >> 2, cq = (lane & 3u) * 8u;
    const uint32_t twq = min(g.tw, 16u);
    const uint32_t lx = ((ql / twq) * hw + ql % twq) * xpitch + cq;
    const uint32_t ld = ql * g.dpitch + cq;
    const uint32_t xhalf = ((16u / g.tw) * hw + 16u % g.tw) * xpitch;  // pixel + 16 in the halo
    const uint32_t xstep = (32u / g.tw) * hw * xpitch;                  // pixel + 32
    const uint32_t c8x = span * 2u, c8d = cout / 8u;
    uint32_t px[kWgPieces], pd[kWgPieces];
#pragma unroll
    for (int k = 0; k < kWgPieces; ++k) {
        px[k] = piece_of(threadIdx.x + uint32_t(k) * kBlock, g.th + 2u, hw, c8x);
        pd[k] = piece_of(threadIdx.x + uint32_t(k) * kBlock, g.th, g.tw, c8d);
    }
    f32x4 acc[NB][MTX], bacc[MTX];
#pragma unroll
    for (int m = 0; m < MTX; ++m) {
        bacc[m] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[j][m] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    const bf16x8 ones = __builtin_bit_cast(bf16x8, u32x4{0x3f803f80u, 0x3f803f80u, 0x3f803f80u, 0x3f803f80u});
    const uint32_t t0 = blockIdx.x * g.per_chunk, t1 = min(t0 + g.per_chunk, g.tiles);
    for (uint32_t t = t0; t < t1; ++t) {
        const uint32_t tyi = t % g.tiles_y, rest = t / g.tiles_y;
        const uint32_t txi = rest % g.tiles_x, img = rest / g.tiles_x;
        const int y0 = int(tyi * g.th), x0 = int(txi * g.tw);
        const size_t pix0 = size_t(img) * d.h * d.w;
        // the piece maps are decoded again per tile rather than held decoded
        // in registers across the loop
#pragma unroll
        for (int k = 0; k < kWgPieces; ++k) asm volatile("" : "+v"(px[k]), "+v"(pd[k]));
        __syncthreads();  // the previous tile's fragments are read
        // global loads in batches of kWgBatch pieces per thread, each batch
        // issued before its LDS stores
#pragma unroll
        for (int k0 = 0; k0 < kWgPieces; k0 += kWgBatch) {
            bf16x8 v[kWgBatch];
#pragma unroll
            for (int k = 0; k < kWgBatch; ++k) {
                const uint32_t pc = px[k0 + k];
                v[k] = zero8();
                if (pc != ~0u) {
                    const int iy = y0 + int(pc & 255u) - 1, ix = x0 + int((pc >> 8) & 255u) - 1;
                    if (iy >= 0 && iy < int(d.h) && ix >= 0 && ix < int(d.w))
                        v[k] = load8(x.p + (pix0 + uint32_t(iy) * d.w + uint32_t(ix)) * x.stride + x.off +
                                     ct0 * 16u + (pc >> 16) * 8u);
                }
            }
#pragma unroll
            for (int k = 0; k < kWgBatch; ++k) {
                const uint32_t pc = px[k0 + k];
                if (pc != ~0u)
                    *(lds_bf16x8*)(xs + (((pc & 255u) * hw + ((pc >> 8) & 255u)) * xpitch + (pc >> 16) * 16u)) = v[k];
            }
        }
#pragma unroll
        for (int k0 = 0; k0 < kWgPieces; k0 += kWgBatch) {
            bf16x8 v[kWgBatch], m[kWgBatch];
#pragma unroll
            for (int k = 0; k < kWgBatch; ++k) {
                const uint32_t pc = pd[k0 + k];
                v[k] = zero8();
                m[k] = zero8();
                if (pc != ~0u) {
                    const uint32_t iy = uint32_t(y0) + (pc & 255u), ix = uint32_t(x0) + ((pc >> 8) & 255u);
                    if (iy < d.h && ix < d.w) {
                        const size_t p = pix0 + iy * d.w + ix;
                        const uint32_t ch = (pc >> 16) * 8u;
                        v[k] = load8(ga.p + p * ga.stride + ga.off + ch);
                        if (a.p) m[k] = load8(a.p + p * a.stride + a.off + ch);
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < kWgBatch; ++k) {
                const uint32_t pc = pd[k0 + k];
                if (pc == ~0u) continue;
                if (a.p) {
#pragma unroll
                    for (int i = 0; i < 8; ++i)
                        if (!(f(m[k][i]) > 0.0f)) v[k][i] = b(0.0f);
                }
                const uint32_t ty = pc & 255u, tx = (pc >> 8) & 255u, ch = (pc >> 16) * 8u;
                *(lds_bf16x8*)(dzs + (ty * g.tw + tx) * g.dpitch + ch * 2u) = v[k];
                const uint32_t iy = uint32_t(y0) + ty, ix = uint32_t(x0) + tx;
                if (dz_out && lead && iy < d.h && ix < d.w) store8(dz_out + (pix0 + iy * d.w + ix) * cout + ch, v[k]);
            }
        }
        __syncthreads();
        for (uint32_t s = 0; s < g.steps; ++s) {
            bf16x8 af[MTX], bx[NB];
#pragma unroll
            for (int m = 0; m < MTX; ++m)
                if (uint32_t(m) < mt) af[m] = tr_frag(dzs + ld + s * 32u * g.dpitch + uint32_t(m) * 32u, 16u * g.dpitch);
#pragma unroll
            for (int j = 0; j < NB; ++j)
                if (uint32_t(j) < nb) bx[j] = tr_frag(xs + lx + s * xstep + boff[j], xhalf);
#pragma unroll
            for (int j = 0; j < NB; ++j)
#pragma unroll
                for (int m = 0; m < MTX; ++m)
                    if (uint32_t(j) < nb && uint32_t(m) < mt)
                        acc[j][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[m], bx[j], acc[j][m], 0, 0, 0);
            if (do_bias) {
#pragma unroll
                for (int m = 0; m < MTX; ++m)
                    if (uint32_t(m) < mt) bacc[m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[m], ones, bacc[m], 0, 0, 0);
            }
        }
    }
    const uint32_t nw = 9u * cout * cin;
    float* out = partial + size_t(blockIdx.x) * (nw + (bias ? cout : 0u));
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        if (uint32_t(j) >= nb) continue;
        const uint32_t fj = f0 + wv + 4u * uint32_t(j), tap = fj % 9u;
        const uint32_t ci = (fj / 9u) * 16u + (lane & 15u);
#pragma unroll
        for (int m = 0; m < MTX; ++m) {
            if (uint32_t(m) >= mt) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t co = uint32_t(m) * 16u + (lane >> 4) * 4u + uint32_t(r);
                out[(size_t(tap) * cout + co) * cin + ci] = acc[j][m][r];
            }
        }
    }
    if (do_bias && (lane & 15u) == 0u) {
#pragma unroll
        for (int m = 0; m < MTX; ++m) {
            if (uint32_t(m) >= mt) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) out[nw + uint32_t(m) * 16u + (lane >> 4) * 4u + uint32_t(r)] = bacc[m][r];
        }
    }
}

struct WgPlan {
    WgGeom g;
    uint32_t mtx, nb;  // template: output tiles (at least cout / 16), fragments per wave
    uint32_t groups, chunks;
    size_t lds;
};

static uint32_t pieces(uint32_t rows, uint32_t cols, uint32_t c8) { return (rows * cols * c8 + kBlock - 1u) / kBlock; }

static bool wgrad_plan(uint32_t cin, uint32_t cout, Dims d, WgPlan& p) {
    p = WgPlan{};
    const uint32_t mt = cout / 16u;
    p.mtx = mt <= 4u ? mt : (mt <= 6u ? 6u : 8u);
    p.nb = p.mtx <= 3u ? 8u : p.mtx == 4u ? 6u : 2u;  // <= 96 accumulators
    WgGeom& g = p.g;
    g.frags = 9u * (cin / 16u);
    g.tw = d.w >= 32u ? 32u : (d.w >= 16u ? 16u : 8u);
    g.dpitch = 32u * (mt | 1u);
    const uint32_t gmin = (g.frags + 4u * p.nb - 1u) / (4u * p.nb);
    for (uint32_t steps = 8; steps >= 2u; steps /= 2u) {
        g.steps = steps;
        g.th = 32u * steps / g.tw;
        if (pieces(g.th, g.tw, cout / 8u) > uint32_t(kWgPieces)) continue;
        for (uint32_t groups = gmin; groups <= g.frags; ++groups) {
            const uint32_t per = (g.frags + groups - 1u) / groups;
            uint32_t span = 0;
            for (uint32_t f0 = 0; f0 < g.frags; f0 += per)
                span = std::max(span, (std::min(f0 + per, g.frags) - 1u) / 9u - f0 / 9u + 1u);
            const size_t lds = size_t(g.th + 2u) * (g.tw + 2u) * 32u * (span | 1u) + size_t(32u * steps) * g.dpitch;
            if (lds <= kWgLds && pieces(g.th + 2u, g.tw + 2u, 2u * span) <= uint32_t(kWgPieces)) {
                g.per_group = per;
                p.groups = (g.frags + per - 1u) / per;
                p.lds = lds;
                g.tiles_x = (d.w + g.tw - 1u) / g.tw;
                g.tiles_y = (d.h + g.th - 1u) / g.th;
                g.tiles = d.n * g.tiles_x * g.tiles_y;
                const uint32_t want = std::max(1u, std::min(g.tiles, kWgBlocks / p.groups));
                g.per_chunk = (g.tiles + want - 1u) / want;
                p.chunks = (g.tiles + g.per_chunk - 1u) / g.per_chunk;
                return true;
            }
        }
    }
    return false;
}

size_t wgrad_partial_floats(uint32_t cin, uint32_t cout, Dims d) {
    WgPlan p;
    if (!wgrad_plan(cin, cout, d, p)) return 0;
    const size_t n = size_t(9) * cout * cin + cout;
    return size_t(p.chunks) * n + size_t(red_segs(p.chunks)) * n;
}

hipError_t conv3x3_wgrad(const Tens& ga, const Tens* a, uint32_t cout, const Tens& x, uint32_t cin, Dims d,
                         bf16* dz_out, float* partial, float* dw, uint32_t ci_off, uint32_t cin_total, float* db,
                         hipStream_t s) {
    if (cin == 0 || cin % 16u || cin > kMaxCout || cout % 16u || cout == 0 || cout > kMaxCout ||
        ci_off + cin > cin_total || d.pixels() == 0 || x.stride % 8u || x.off % 8u || ga.stride % 8u ||
        ga.off % 8u || (a && (a->stride % 8u || a->off % 8u)) || d.h > 65535u || d.w > 65535u)
        return hipErrorInvalidValue;
    WgPlan p;
    if (!wgrad_plan(cin, cout, d, p)) return hipErrorInvalidValue;
    const dim3 grid(p.chunks, p.groups);
    const Tens an = a ? *a : Tens{nullptr, 0u, 0u};
    const bool bias = db != nullptr;
#define PT_WGRAD_LAUNCH(MT_, NB_)                                                                                   \
    hipLaunchKernelGGL((wgrad_k<MT_, NB_>), grid, dim3(kBlock), p.lds, s, ga, an, cout, x, cin, d, p.g, dz_out, bias, \
                       partial)
    switch (p.mtx) {
        case 1: PT_WGRAD_LAUNCH(1, 8); break;
        case 2: PT_WGRAD_LAUNCH(2, 8); break;
        case 3: PT_WGRAD_LAUNCH(3, 8); break;
        case 4: PT_WGRAD_LAUNCH(4, 6); break;
        case 6: PT_WGRAD_LAUNCH(6, 2); break;
        default: PT_WGRAD_LAUNCH(8, 2); break;
    }
#undef PT_WGRAD_LAUNCH
    const uint32_t nw = 9u * cout * cin, n = nw + (bias ? cout : 0u);
    return reduce_chunks(partial, p.chunks, n, nw, cin, cin_total, ci_off, partial + size_t(p.chunks) * n, dw, db, s);
}

// --------------------------------------------------------- elementwise --
// 8 channels (16 bytes) per thread; c, strides and offsets are multiples of 8.
// d: the fine (input) dims
__global__ void __launch_bounds__(kBlock) pool2_k(Tens in, uint32_t c, Dims d, Tens out) {
    const uint32_t ow = d.w / 2u, oh = d.h / 2u, c8 = c / 8u;
    const uint32_t n = d.n * oh * ow * c8;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t ch = (i % c8) * 8u, op = i / c8;
    const uint32_t x = op % ow, y = (op / ow) % oh, img = op / (ow * oh);
    const size_t p00 = (size_t(img) * d.h + 2u * y) * d.w + 2u * x;
    const bf16* q = in.p + in.off + ch;
    const bf16x8 v0 = load8(q + p00 * in.stride), v1 = load8(q + (p00 + 1u) * in.stride);
    const bf16x8 v2 = load8(q + (p00 + d.w) * in.stride), v3 = load8(q + (p00 + d.w + 1u) * in.stride);
    bf16x8 o;
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = b(0.25f * (f(v0[k]) + f(v1[k]) + f(v2[k]) + f(v3[k])));
    store8(out.p + size_t(op) * out.stride + out.off + ch, o);
}

static bool views_ok(const Tens& a, const Tens& b2, uint32_t c) {
    return c % 8u == 0 && a.stride % 8u == 0 && a.off % 8u == 0 && b2.stride % 8u == 0 && b2.off % 8u == 0;
}

hipError_t pool2(const Tens& in, uint32_t c, Dims d, const Tens& out, hipStream_t s) {
    if (d.h % 2u || d.w % 2u || !views_ok(in, out, c)) return hipErrorInvalidValue;
    const uint32_t n = d.n * (d.h / 2u) * (d.w / 2u) * (c / 8u);
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(pool2_k, dim3((n + kBlock - 1u) / kBlock), dim3(kBlock), 0, s, in, c, d, out);
    return hipGetLastError();
}

// d: the fine dims; din (fine) += dout (coarse) / 4
__global__ void __launch_bounds__(kBlock) pool2_back_k(Tens dout, uint32_t c, Dims d, Tens din) {
    const uint32_t c8 = c / 8u;
    const uint32_t n = d.pixels() * c8;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t ch = (i % c8) * 8u, p = i / c8;
    const uint32_t x = p % d.w, y = (p / d.w) % d.h, img = p / (d.w * d.h);
    const size_t cp = (size_t(img) * (d.h / 2u) + y / 2u) * (d.w / 2u) + x / 2u;
    bf16* q = din.p + size_t(p) * din.stride + din.off + ch;
    const bf16x8 g = load8(dout.p + cp * dout.stride + dout.off + ch), v = load8(q);
    bf16x8 o;
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = b(f(v[k]) + 0.25f * f(g[k]));
    store8(q, o);
}

hipError_t pool2_back_add(const Tens& dout, uint32_t c, Dims d, const Tens& din, hipStream_t s) {
    if (!views_ok(dout, din, c)) return hipErrorInvalidValue;
    const uint32_t n = d.pixels() * (c / 8u);
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(pool2_back_k, dim3((n + kBlock - 1u) / kBlock), dim3(kBlock), 0, s, dout, c, d, din);
    return hipGetLastError();
}

// d: the coarse (input) dims
__global__ void __launch_bounds__(kBlock) up2_k(Tens in, uint32_t c, Dims d, Tens out) {
    const uint32_t W = d.w * 2u, H = d.h * 2u, c8 = c / 8u;
    const uint32_t n = d.n * H * W * c8;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t ch = (i % c8) * 8u, p = i / c8;
    const uint32_t x = p % W, y = (p / W) % H, img = p / (W * H);
    const size_t cp = (size_t(img) * d.h + y / 2u) * d.w + x / 2u;
    store8(out.p + size_t(p) * out.stride + out.off + ch, load8(in.p + cp * in.stride + in.off + ch));
}

hipError_t up2(const Tens& in, uint32_t c, Dims d, const Tens& out, hipStream_t s) {
    if (!views_ok(in, out, c)) return hipErrorInvalidValue;
    const uint32_t n = d.n * d.h * d.w * 4u * (c / 8u);
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(up2_k, dim3((n + kBlock - 1u) / kBlock), dim3(kBlock), 0, s, in, c, d, out);
    return hipGetLastError();
}

// d: the coarse (output) dims
__global__ void __launch_bounds__(kBlock) up2_back_k(Tens dout, uint32_t c, Dims d, Tens din) {
    const uint32_t c8 = c / 8u;
    const uint32_t n = d.pixels() * c8;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t ch = (i % c8) * 8u, p = i / c8;
    const uint32_t x = p % d.w, y = (p / d.w) % d.h, img = p / (d.w * d.h);
    const uint32_t W = d.w * 2u;
    const size_t f00 = (size_t(img) * d.h * 2u + 2u * y) * W + 2u * x;
    const bf16* q = dout.p + dout.off + ch;
    const bf16x8 v0 = load8(q + f00 * dout.stride), v1 = load8(q + (f00 + 1u) * dout.stride);
    const bf16x8 v2 = load8(q + (f00 + W) * dout.stride), v3 = load8(q + (f00 + W + 1u) * dout.stride);
    bf16x8 o;
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = b(f(v0[k]) + f(v1[k]) + f(v2[k]) + f(v3[k]));
    store8(din.p + size_t(p) * din.stride + din.off + ch, o);
}

hipError_t up2_back(const Tens& dout, uint32_t c, Dims d, const Tens& din, hipStream_t s) {
    if (!views_ok(dout, din, c)) return hipErrorInvalidValue;
    const uint32_t n = d.pixels() * (c / 8u);
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(up2_back_k, dim3((n + kBlock - 1u) / kBlock), dim3(kBlock), 0, s, dout, c, d, din);
    return hipGetLastError();
}

__global__ void __launch_bounds__(kBlock) transpose_w_k(const bf16* __restrict__ w, uint32_t cout, uint32_t cin,
                                                        uint32_t ci0, uint32_t nci, bf16* __restrict__ wd) {
    const uint32_t n = 9u * cout * nci;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t co = i % cout, ci = (i / cout) % nci, tap = i / (cout * nci);
    wd[(size_t(8u - tap) * nci + ci) * cout + co] = w[(size_t(tap) * cout + co) * cin + ci0 + ci];
}

hipError_t transpose_weights(const bf16* w, uint32_t cout, uint32_t cin, uint32_t ci_begin, uint32_t ci_count,
                             bf16* wd, hipStream_t s) {
    if (ci_begin + ci_count > cin) return hipErrorInvalidValue;
    const uint32_t n = 9u * cout * ci_count;
    hipLaunchKernelGGL(transpose_w_k, dim3((n + kBlock - 1u) / kBlock), dim3(kBlock), 0, s, w, cout, cin, ci_begin,
                       ci_count, wd);
    return hipGetLastError();
}

__global__ void __launch_bounds__(kBlock) to_bf16_k(const float* __restrict__ src, bf16* __restrict__ dst, size_t n) {
    const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) dst[i] = b(src[i]);
}
hipError_t to_bf16(const float* src, bf16* dst, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(to_bf16_k, dim3(uint32_t((n + kBlock - 1u) / kBlock)), dim3(kBlock), 0, s, src, dst, n);
    return hipGetLastError();
}

__global__ void __launch_bounds__(kBlock) zero_bf16_k(bf16* p, size_t n) {
    const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) p[i] = b(0.0f);
}
__global__ void __launch_bounds__(kBlock) zero_f32_k(float* p, size_t n) {
    const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}
hipError_t fill_zero_bf16(bf16* p, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(zero_bf16_k, dim3(uint32_t((n + kBlock - 1u) / kBlock)), dim3(kBlock), 0, s, p, n);
    return hipGetLastError();
}
hipError_t fill_zero_f32(float* p, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(zero_f32_k, dim3(uint32_t((n + kBlock - 1u) / kBlock)), dim3(kBlock), 0, s, p, n);
    return hipGetLastError();
}

__global__ void __launch_bounds__(kBlock) adam_k(float* __restrict__ w, const float* __restrict__ g,
                                                 float* __restrict__ m, float* __restrict__ v, size_t n, float lr,
                                                 float b1, float b2, float eps, float c1, float c2) {
    const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

hipError_t adam(float* w, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2, float eps,
                uint32_t t, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const float c1 = 1.0f - powf(b1, float(t)), c2 = 1.0f - powf(b2, float(t));
    hipLaunchKernelGGL(adam_k, dim3(uint32_t((n + kBlock - 1u) / kBlock)), dim3(kBlock), 0, s, w, g, m, v, n, lr, b1,
                       b2, eps, c1, c2);
    return hipGetLastError();
}

// ------------------------------------------------------- input / output --
// The 16 packed input channels of a pixel (o 16-byte aligned), written as
// two 16-byte stores.
__device__ __forceinline__ void write_features(bf16* o, float r, float g, float bl, float ar, float ag, float ab,
                                               float nx, float ny, float nz, float dep) {
    bf16x8 lo, hi = zero8();
    lo[0] = b(log1pf(fmaxf(r, 0.0f)));
    lo[1] = b(log1pf(fmaxf(g, 0.0f)));
    lo[2] = b(log1pf(fmaxf(bl, 0.0f)));
    lo[3] = b(ar);
    lo[4] = b(ag);
    lo[5] = b(ab);
    lo[6] = b(nx);
    lo[7] = b(ny);
    hi[0] = b(nz);
    const float dd = fmaxf(dep, 0.0f);
    hi[1] = b(dd / (dd + 1.0f));
    store8(o, lo);
    store8(o + 8, hi);
}

__global__ void __launch_bounds__(kBlock) pack_input_k(const float* __restrict__ feat, uint32_t n,
                                                       bf16* __restrict__ out) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    const float* q = feat + size_t(p) * 10u;
    write_features(out + size_t(p) * 16u, q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], q[8], q[9]);
}

hipError_t pack_input(const float* feat, Dims d, bf16* out, hipStream_t s) {
    const uint32_t n = d.pixels();
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(pack_input_k, dim3((n + kBlock - 1u) / kBlock), dim3(kBlock), 0, s, feat, n, out);
    return hipGetLastError();
}

struct GTexel {  // pt_gbuffer_texel
    float normal[3];
    float roughness;
    float albedo[3];
    float metallic;
    float motion[2];
    float transmission;
    float dist;
};

__global__ void __launch_bounds__(kBlock) pack_frame_k(const float4* __restrict__ accum, const GTexel* __restrict__ g,
                                                       uint32_t fw, uint32_t fh, uint32_t x0, uint32_t y0, Dims d,
                                                       bf16* __restrict__ out) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= d.pixels()) return;
    const uint32_t x = x0 + p % d.w, y = y0 + p / d.w;
    bf16* o = out + size_t(p) * 16u;
    if (x >= fw || y >= fh) {
        store8(o, zero8());
        store8(o + 8, zero8());
        return;
    }
    const uint32_t i = y * fw + x;
    const float4 a = accum[i];
    const GTexel& t = g[i];
    write_features(o, a.x, a.y, a.z, t.albedo[0], t.albedo[1], t.albedo[2], t.normal[0], t.normal[1], t.normal[2],
                   t.dist);
}

hipError_t pack_frame(const float4* accum, const void* gbuf, uint32_t fw, uint32_t fh, uint32_t x0, uint32_t y0,
                      Dims d, bf16* out, hipStream_t s) {
    const uint32_t n = d.pixels();
    if (n == 0 || d.n != 1u) return hipErrorInvalidValue;
    hipLaunchKernelGGL(pack_frame_k, dim3((n + kBlock - 1u) / kBlock), dim3(kBlock), 0, s, accum,
                       static_cast<const GTexel*>(gbuf), fw, fh, x0, y0, d, out);
    return hipGetLastError();
}

__global__ void __launch_bounds__(kBlock) crop_radiance_k(const float4* __restrict__ accum, uint32_t fw, uint32_t x0,
                                                          uint32_t y0, Dims d, float* __restrict__ out) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= d.pixels()) return;
    const float4 a = accum[size_t(y0 + p / d.w) * fw + x0 + p % d.w];
    out[size_t(p) * 3u] = a.x;
    out[size_t(p) * 3u + 1u] = a.y;
    out[size_t(p) * 3u + 2u] = a.z;
}

hipError_t crop_radiance(const float4* accum, uint32_t fw, uint32_t fh, uint32_t x0, uint32_t y0, Dims d, float* out,
                         hipStream_t s) {
    const uint32_t n = d.pixels();
    if (n == 0 || d.n != 1u || x0 + d.w > fw || y0 + d.h > fh) return hipErrorInvalidValue;
    hipLaunchKernelGGL(crop_radiance_k, dim3((n + kBlock - 1u) / kBlock), dim3(kBlock), 0, s, accum, fw, x0, y0, d,
                       out);
    return hipGetLastError();
}

__global__ void __launch_bounds__(kBlock) l1_log_loss_k(const float* __restrict__ pred,
                                                        const float* __restrict__ target, uint32_t n, float scale,
                                                        bf16* __restrict__ dpred, float* __restrict__ part) {
    __shared__ float red[kBlock];
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    float l = 0.0f;
    if (p < n) {
        bf16* o = dpred + size_t(p) * 16u;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float t = log1pf(fmaxf(target[size_t(p) * 3u + k], 0.0f));
            const float diff = pred[size_t(p) * 3u + k] - t;
            l += fabsf(diff);
            o[k] = b(diff > 0.0f ? scale : (diff < 0.0f ? -scale : 0.0f));
        }
#pragma unroll
        for (int k = 3; k < 16; ++k) o[k] = b(0.0f);
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (uint32_t off = kBlock / 2u; off > 0u; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// loss += the block partials, summed in a fixed order (the same bits on
// every run, which float atomics would not give)
__global__ void __launch_bounds__(kBlock) loss_sum_k(const float* __restrict__ part, uint32_t nb,
                                                     float* __restrict__ loss) {
    __shared__ float red[kBlock];
    float l = 0.0f;
    for (uint32_t i = threadIdx.x; i < nb; i += kBlock) l += part[i];
    red[threadIdx.x] = l;
    __syncthreads();
    for (uint32_t off = kBlock / 2u; off > 0u; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss += red[0];
}

uint32_t loss_partials(uint32_t pixels) { return (pixels + kBlock - 1u) / kBlock; }

hipError_t l1_log_loss(const float* pred, const float* target, Dims d, bf16* dpred, float* part, float* loss,
                       hipStream_t s) {
    const uint32_t n = d.pixels();
    if (n == 0) return hipSuccess;
    const float scale = 1.0f / (float(n) * 3.0f);
    const uint32_t nb = loss_partials(n);
    hipLaunchKernelGGL(l1_log_loss_k, dim3(nb), dim3(kBlock), 0, s, pred, target, n, scale, dpred, part);
    hipLaunchKernelGGL(loss_sum_k, dim3(1), dim3(kBlock), 0, s, part, nb, loss);
    return hipGetLastError();
}

__global__ void __launch_bounds__(kBlock) unpack_frame_k(const float* __restrict__ pred, Dims d, uint32_t fw,
                                                         uint32_t fh, float exposure, float4* __restrict__ fin,
                                                         uchar4* __restrict__ out) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= fw * fh) return;
    const uint32_t x = i % fw, y = i / fw;
    frame_texel(pred + (size_t(y) * d.w + x) * 3u, exposure, fin + i, out ? out + i : nullptr);
}

hipError_t unpack_frame(const float* pred, Dims d, uint32_t fw, uint32_t fh, float exposure, float4* fin,
                        uchar4* out, hipStream_t s) {
    const uint32_t n = fw * fh;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(unpack_frame_k, dim3((n + kBlock - 1u) / kBlock), dim3(kBlock), 0, s, pred, d, fw, fh,
                       exposure, fin, out);
    return hipGetLastError();
}

}  // namespace nn
}  // namespace pt
