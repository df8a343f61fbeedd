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
//               channels, K = pixels; dz and the tap-shifted input are staged
//               through LDS transposed (channel-major), so every fragment is
//               one 16-byte LDS read; split over pixel chunks, partial sums
//               reduced by wgrad_reduce_k (deterministic)
//   the rest    pooling, upsampling, ReLU gradient, Adam, loss, packing
#include "nn_internal.h"

namespace pt {
namespace nn {

typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr uint32_t kBlock = 256;

__device__ __forceinline__ bf16x8 load8(const bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }
__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, u32x4{0u, 0u, 0u, 0u}); }
__device__ __forceinline__ float f(bf16 x) { return float(x); }
__device__ __forceinline__ bf16 b(float x) { return bf16(x); }

// ------------------------------------------------------------- conv fwd --
template <int NCT, bool RELU, bool F32OUT>
__global__ void __launch_bounds__(kBlock) conv3x3_k(Tens in, uint32_t cin, const bf16* __restrict__ w,
                                                    const float* __restrict__ bias, Tens out, uint32_t cout, Dims d,
                                                    float* __restrict__ of32, uint32_t f32c) {
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const uint32_t npix = d.pixels();
    const uint32_t p0 = (blockIdx.x * 4u + wv) * 16u;
    if (p0 >= npix) return;  // the whole wave is past the end
    const uint32_t pix = p0 + (lane & 15u), kq = lane >> 4;
    const bool pv = pix < npix;
    const uint32_t x = pix % d.w, y = (pix / d.w) % d.h, img = pix / (d.w * d.h);
    f32x4 acc[NCT];
#pragma unroll
    for (int t = 0; t < NCT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (uint32_t tap = 0; tap < 9u; ++tap) {
        const int sy = int(y) + int(tap / 3u) - 1, sx = int(x) + int(tap % 3u) - 1;
        const bool valid = pv && sy >= 0 && sy < int(d.h) && sx >= 0 && sx < int(d.w);
        const size_t sp = valid ? (size_t(img) * d.h + uint32_t(sy)) * d.w + uint32_t(sx) : 0;
        const bf16* src = in.p + sp * in.stride + in.off;
        const bf16* wt = w + size_t(tap) * cout * cin;
        for (uint32_t c0 = 0; c0 < cin; c0 += 32u) {
            const uint32_t c = c0 + kq * 8u;
            const bool cv = c < cin;
            const bf16x8 a = (valid && cv) ? load8(src + c) : zero8();
#pragma unroll
            for (int t = 0; t < NCT; ++t) {
                const uint32_t co = uint32_t(t) * 16u + (lane & 15u);
                const bf16x8 bw = cv ? load8(wt + size_t(co) * cin + c) : zero8();
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw, acc[t], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < NCT; ++t) {
        const uint32_t co = uint32_t(t) * 16u + (lane & 15u);
        const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t op = p0 + kq * 4u + uint32_t(r);
            if (op >= npix) continue;
            float v = acc[t][r] + bv;
            if (RELU) v = fmaxf(v, 0.0f);
            if (F32OUT) {
                if (co < f32c) of32[size_t(op) * f32c + co] = v;
            } else {
                out.p[size_t(op) * out.stride + out.off + co] = b(v);
            }
        }
    }
}

template <int NCT>
static void conv_launch(const Tens& in, uint32_t cin, const bf16* w, const float* bias, const Tens& out,
                        uint32_t cout, Dims d, bool relu, float* of32, uint32_t f32c, hipStream_t s) {
    const uint32_t blocks = (d.pixels() + 63u) / 64u;
    if (of32)
        hipLaunchKernelGGL((conv3x3_k<NCT, false, true>), dim3(blocks), dim3(kBlock), 0, s, in, cin, w, bias, out,
                           cout, d, of32, f32c);
    else if (relu)
        hipLaunchKernelGGL((conv3x3_k<NCT, true, false>), dim3(blocks), dim3(kBlock), 0, s, in, cin, w, bias, out,
                           cout, d, of32, f32c);
    else
        hipLaunchKernelGGL((conv3x3_k<NCT, false, false>), dim3(blocks), dim3(kBlock), 0, s, in, cin, w, bias, out,
                           cout, d, of32, f32c);
}

hipError_t conv3x3(const Tens& in, uint32_t cin, const bf16* w, const float* bias, const Tens& out, uint32_t cout,
                   Dims d, bool relu, float* out_f32, uint32_t f32_channels, hipStream_t s) {
    if (cin == 0 || cin % 16u || cout % 16u || cout == 0 || cout > kMaxCout || in.stride % 8u || in.off % 8u)
        return hipErrorInvalidValue;
    if (d.pixels() == 0) return hipSuccess;
    switch (cout / 16u) {
        case 1: conv_launch<1>(in, cin, w, bias, out, cout, d, relu, out_f32, f32_channels, s); break;
        case 2: conv_launch<2>(in, cin, w, bias, out, cout, d, relu, out_f32, f32_channels, s); break;
        case 3: conv_launch<3>(in, cin, w, bias, out, cout, d, relu, out_f32, f32_channels, s); break;
        case 4: conv_launch<4>(in, cin, w, bias, out, cout, d, relu, out_f32, f32_channels, s); break;
        case 5: conv_launch<5>(in, cin, w, bias, out, cout, d, relu, out_f32, f32_channels, s); break;
        case 6: conv_launch<6>(in, cin, w, bias, out, cout, d, relu, out_f32, f32_channels, s); break;
        case 7: conv_launch<7>(in, cin, w, bias, out, cout, d, relu, out_f32, f32_channels, s); break;
        case 8: conv_launch<8>(in, cin, w, bias, out, cout, d, relu, out_f32, f32_channels, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------- wgrad --
constexpr uint32_t kWgStep = 32;     // pixels per K step
constexpr uint32_t kWgPad = 40;      // LDS row length (bf16): 80 B rows keep 16-byte alignment
constexpr uint32_t kWgBlocks = 2048; // target workgroups per launch

static uint32_t wgrad_chunks(uint32_t cout, Dims d) {
    const uint32_t groups = 9u * ((cout + 31u) / 32u);
    const uint32_t steps = (d.pixels() + kWgStep - 1u) / kWgStep;
    uint32_t chunks = kWgBlocks / groups;
    if (chunks < 1u) chunks = 1u;
    if (chunks > steps) chunks = steps;
    return chunks;
}

size_t wgrad_partial_floats(uint32_t cin, uint32_t cout, Dims d) {
    return size_t(wgrad_chunks(cout, d)) * 9u * cout * cin;
}

// grid: x = pixel chunk, y = tap, z = group of 32 output channels
__global__ void __launch_bounds__(kBlock) wgrad_k(Tens dz, uint32_t cout, Tens x, uint32_t cin, Dims d,
                                                  uint32_t chunk, float* __restrict__ partial) {
    __shared__ __attribute__((aligned(16))) bf16 sdz[32][kWgPad];    // [co][px]
    __shared__ __attribute__((aligned(16))) bf16 sx[kMaxCout][kWgPad];  // [ci][px]
    const uint32_t tap = blockIdx.y, co0 = blockIdx.z * 32u;
    const int dy = int(tap / 3u) - 1, dx = int(tap % 3u) - 1;
    const uint32_t npix = d.pixels();
    const uint32_t pb = blockIdx.x * chunk;
    const uint32_t pe = pb + chunk < npix ? pb + chunk : npix;
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6, kq = lane >> 4;
    const uint32_t cit = cin / 16u, ntiles = 2u * cit;
    f32x4 acc[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (uint32_t p = pb; p < pe; p += kWgStep) {
        // stage dz: 32 px x 32 co, 4 channels per thread
        {
            const uint32_t px = threadIdx.x >> 3, c4 = (threadIdx.x & 7u) * 4u;
            const uint32_t gp = p + px, co = co0 + c4;
            float v[4] = {0.f, 0.f, 0.f, 0.f};
            if (gp < pe && co < cout) {
                const bf16* q = dz.p + size_t(gp) * dz.stride + dz.off + co;
#pragma unroll
                for (int i = 0; i < 4; ++i) v[i] = f(q[i]);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) sdz[c4 + i][px] = b(v[i]);
        }
        // stage the tap-shifted input: 32 px x cin, 8 channels per piece
        for (uint32_t piece = threadIdx.x; piece < 32u * (cin / 8u); piece += kBlock) {
            const uint32_t px = piece / (cin / 8u), c8 = (piece % (cin / 8u)) * 8u;
            const uint32_t gp = p + px;
            bf16x8 v = zero8();
            if (gp < pe) {
                const uint32_t xx = gp % d.w, yy = (gp / d.w) % d.h, img = gp / (d.w * d.h);
                const int sy = int(yy) + dy, sx2 = int(xx) + dx;
                if (sy >= 0 && sy < int(d.h) && sx2 >= 0 && sx2 < int(d.w))
                    v = load8(x.p + ((size_t(img) * d.h + uint32_t(sy)) * d.w + uint32_t(sx2)) * x.stride + x.off + c8);
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) sx[c8 + i][px] = v[i];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t t = wv + 4u * uint32_t(i);
            if (t < ntiles) {
                const uint32_t cot = t & 1u, cti = t >> 1;
                const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sdz[cot * 16u + (lane & 15u)][kq * 8u]);
                const bf16x8 bb = *reinterpret_cast<const bf16x8*>(&sx[cti * 16u + (lane & 15u)][kq * 8u]);
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc[i], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    float* out = partial + size_t(blockIdx.x * 9u + tap) * cout * cin;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t t = wv + 4u * uint32_t(i);
        if (t >= ntiles) continue;
        const uint32_t cot = t & 1u, cti = t >> 1;
        const uint32_t ci = cti * 16u + (lane & 15u);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint32_t co = co0 + cot * 16u + kq * 4u + uint32_t(r);
            if (co < cout) out[size_t(co) * cin + ci] = acc[i][r];
        }
    }
}

// out[row][col_off + col] (row length out_cols) = sum over chunks of
// partial[chunk][row][col] (row length cols), for rows x cols = n entries
__global__ void __launch_bounds__(kBlock) reduce_chunks_k(const float* __restrict__ partial, uint32_t chunks,
                                                          uint32_t n, uint32_t cols, uint32_t out_cols,
                                                          uint32_t col_off, float* __restrict__ out) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (uint32_t c = 0; c < chunks; ++c) s += partial[size_t(c) * n + i];
    const uint32_t row = i / cols, col = i - row * cols;
    out[size_t(row) * out_cols + col_off + col] = s;
}

hipError_t conv3x3_wgrad(const Tens& dz, uint32_t cout, const Tens& x, uint32_t cin, Dims d, float* partial,
                         float* dw, uint32_t ci_off, uint32_t cin_total, hipStream_t s) {
    if (cin == 0 || cin % 16u || cin > kMaxCout || cout % 16u || cout == 0 || cout > kMaxCout || x.stride % 8u ||
        x.off % 8u || ci_off + cin > cin_total)
        return hipErrorInvalidValue;
    const uint32_t n = 9u * cout * cin;
    if (d.pixels() == 0) return hipErrorInvalidValue;
    const uint32_t chunks = wgrad_chunks(cout, d);
    const uint32_t steps = (d.pixels() + kWgStep - 1u) / kWgStep;
    const uint32_t chunk = ((steps + chunks - 1u) / chunks) * kWgStep;
    const uint32_t used = (d.pixels() + chunk - 1u) / chunk;
    hipLaunchKernelGGL(wgrad_k, dim3(used, 9, (cout + 31u) / 32u), dim3(kBlock), 0, s, dz, cout, x, cin, d, chunk,
                       partial);
    hipLaunchKernelGGL(reduce_chunks_k, dim3((n + kBlock - 1u) / kBlock), dim3(kBlock), 0, s, partial, used, n, cin,
                       cin_total, ci_off, dw);
    return hipGetLastError();
}

// ------------------------------------------------------------ bias grad --
constexpr uint32_t kBgChunks = 256;
__global__ void __launch_bounds__(kBlock) bias_grad_k(Tens dz, uint32_t cout, Dims d, uint32_t chunk,
                                                      float* __restrict__ partial) {
    __shared__ float red[kBlock];
    const uint32_t lanes = kBlock / cout;  // >= 2 for cout <= 128; threads past lanes * cout idle
    const uint32_t co = threadIdx.x % cout, ph = threadIdx.x / cout;
    const uint32_t npix = d.pixels();
    const uint32_t pb = blockIdx.x * chunk;
    const uint32_t pe = pb + chunk < npix ? pb + chunk : npix;
    float sum = 0.0f;
    if (ph < lanes)
        for (uint32_t p = pb + ph; p < pe; p += lanes) sum += f(dz.p[size_t(p) * dz.stride + dz.off + co]);
    red[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x < cout) {
        float t = 0.0f;
        for (uint32_t k = 0; k < lanes; ++k) t += red[k * cout + threadIdx.x];
        partial[size_t(blockIdx.x) * cout + threadIdx.x] = t;
    }
}

size_t bias_partial_floats(uint32_t cout) { return size_t(kBgChunks) * cout; }

hipError_t bias_grad(const Tens& dz, uint32_t cout, Dims d, float* partial, float* db, hipStream_t s) {
    if (cout == 0 || cout > kMaxCout) return hipErrorInvalidValue;
    const uint32_t npix = d.pixels();
    const uint32_t chunk = (npix + kBgChunks - 1u) / kBgChunks;
    const uint32_t used = chunk ? (npix + chunk - 1u) / chunk : 0u;
    if (used == 0) return fill_zero_f32(db, cout, s);
    hipLaunchKernelGGL(bias_grad_k, dim3(used), dim3(kBlock), 0, s, dz, cout, d, chunk, partial);
    hipLaunchKernelGGL(reduce_chunks_k, dim3((cout + kBlock - 1u) / kBlock), dim3(kBlock), 0, s, partial, used, cout,
                       cout, cout, 0u, db);
    return hipGetLastError();
}

// --------------------------------------------------------- elementwise --
__global__ void __launch_bounds__(kBlock) relu_grad_k(Tens da, Tens a, uint32_t c, uint32_t n, bf16* __restrict__ dz) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t p = i / c, ch = i - p * c;
    const float av = f(a.p[size_t(p) * a.stride + a.off + ch]);
    dz[i] = av > 0.0f ? da.p[size_t(p) * da.stride + da.off + ch] : b(0.0f);
}

hipError_t relu_grad(const Tens& da, const Tens& a, uint32_t c, Dims d, bf16* dz, hipStream_t s) {
    const uint32_t n = d.pixels() * c;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(relu_grad_k, dim3((n + kBlock - 1u) / kBlock), dim3(kBlock), 0, s, da, a, c, n, dz);
    return hipGetLastError();
}

// d: the fine (input) dims
__global__ void __launch_bounds__(kBlock) pool2_k(Tens in, uint32_t c, Dims d, Tens out) {
    const uint32_t ow = d.w / 2u, oh = d.h / 2u;
    const uint32_t n = d.n * oh * ow * c;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t ch = i % c, op = i / c;
    const uint32_t x = op % ow, y = (op / ow) % oh, img = op / (ow * oh);
    const size_t p00 = (size_t(img) * d.h + 2u * y) * d.w + 2u * x;
    const bf16* q = in.p + in.off + ch;
    const float s = f(q[p00 * in.stride]) + f(q[(p00 + 1u) * in.stride]) + f(q[(p00 + d.w) * in.stride]) +
                    f(q[(p00 + d.w + 1u) * in.stride]);
    out.p[size_t(op) * out.stride + out.off + ch] = b(0.25f * s);
}

hipError_t pool2(const Tens& in, uint32_t c, Dims d, const Tens& out, hipStream_t s) {
    if (d.h % 2u || d.w % 2u) return hipErrorInvalidValue;
    const uint32_t n = d.n * (d.h / 2u) * (d.w / 2u) * c;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(pool2_k, dim3((n + kBlock - 1u) / kBlock), dim3(kBlock), 0, s, in, c, d, out);
    return hipGetLastError();
}

// d: the fine dims; din (fine) += dout (coarse) / 4
__global__ void __launch_bounds__(kBlock) pool2_back_k(Tens dout, uint32_t c, Dims d, Tens din) {
    const uint32_t n = d.pixels() * c;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t ch = i % c, p = i / c;
    const uint32_t x = p % d.w, y = (p / d.w) % d.h, img = p / (d.w * d.h);
    const size_t cp = (size_t(img) * (d.h / 2u) + y / 2u) * (d.w / 2u) + x / 2u;
    bf16* q = din.p + size_t(p) * din.stride + din.off + ch;
    *q = b(f(*q) + 0.25f * f(dout.p[cp * dout.stride + dout.off + ch]));
}

hipError_t pool2_back_add(const Tens& dout, uint32_t c, Dims d, const Tens& din, hipStream_t s) {
    const uint32_t n = d.pixels() * c;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(pool2_back_k, dim3((n + kBlock - 1u) / kBlock), dim3(kBlock), 0, s, dout, c, d, din);
    return hipGetLastError();
}

// d: the coarse (input) dims
__global__ void __launch_bounds__(kBlock) up2_k(Tens in, uint32_t c, Dims d, Tens out) {
    const uint32_t W = d.w * 2u, H = d.h * 2u;
    const uint32_t n = d.n * H * W * c;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t ch = i % c, p = i / c;
    const uint32_t x = p % W, y = (p / W) % H, img = p / (W * H);
    const size_t cp = (size_t(img) * d.h + y / 2u) * d.w + x / 2u;
    out.p[size_t(p) * out.stride + out.off + ch] = in.p[cp * in.stride + in.off + ch];
}

hipError_t up2(const Tens& in, uint32_t c, Dims d, const Tens& out, hipStream_t s) {
    const uint32_t n = d.n * d.h * d.w * 4u * c;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(up2_k, dim3((n + kBlock - 1u) / kBlock), dim3(kBlock), 0, s, in, c, d, out);
    return hipGetLastError();
}

// d: the coarse (output) dims
__global__ void __launch_bounds__(kBlock) up2_back_k(Tens dout, uint32_t c, Dims d, Tens din) {
    const uint32_t n = d.pixels() * c;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t ch = i % c, p = i / c;
    const uint32_t x = p % d.w, y = (p / d.w) % d.h, img = p / (d.w * d.h);
    const uint32_t W = d.w * 2u;
    const size_t f00 = (size_t(img) * d.h * 2u + 2u * y) * W + 2u * x;
    const bf16* q = dout.p + dout.off + ch;
    const float s = f(q[f00 * dout.stride]) + f(q[(f00 + 1u) * dout.stride]) + f(q[(f00 + W) * dout.stride]) +
                    f(q[(f00 + W + 1u) * dout.stride]);
    din.p[size_t(p) * din.stride + din.off + ch] = b(s);
}

hipError_t up2_back(const Tens& dout, uint32_t c, Dims d, const Tens& din, hipStream_t s) {
    const uint32_t n = d.pixels() * c;
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
__device__ __forceinline__ void write_features(bf16* o, float r, float g, float bl, float ar, float ag, float ab,
                                               float nx, float ny, float nz, float dep) {
    o[0] = b(log1pf(fmaxf(r, 0.0f)));
    o[1] = b(log1pf(fmaxf(g, 0.0f)));
    o[2] = b(log1pf(fmaxf(bl, 0.0f)));
    o[3] = b(ar);
    o[4] = b(ag);
    o[5] = b(ab);
    o[6] = b(nx);
    o[7] = b(ny);
    o[8] = b(nz);
    const float dd = fmaxf(dep, 0.0f);
    o[9] = b(dd / (dd + 1.0f));
#pragma unroll
    for (int k = 10; k < 16; ++k) o[k] = b(0.0f);
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
                                                       uint32_t fw, uint32_t fh, Dims d, bf16* __restrict__ out) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= d.pixels()) return;
    const uint32_t x = p % d.w, y = p / d.w;
    bf16* o = out + size_t(p) * 16u;
    if (x >= fw || y >= fh) {
#pragma unroll
        for (int k = 0; k < 16; ++k) o[k] = b(0.0f);
        return;
    }
    const uint32_t i = y * fw + x;
    const float4 a = accum[i];
    const GTexel& t = g[i];
    write_features(o, a.x, a.y, a.z, t.albedo[0], t.albedo[1], t.albedo[2], t.normal[0], t.normal[1], t.normal[2],
                   t.dist);
}

hipError_t pack_frame(const float4* accum, const void* gbuf, uint32_t fw, uint32_t fh, Dims d, bf16* out,
                      hipStream_t s) {
    const uint32_t n = d.pixels();
    if (n == 0 || d.n != 1u || d.w < fw || d.h < fh) return hipErrorInvalidValue;
    hipLaunchKernelGGL(pack_frame_k, dim3((n + kBlock - 1u) / kBlock), dim3(kBlock), 0, s, accum,
                       static_cast<const GTexel*>(gbuf), fw, fh, d, out);
    return hipGetLastError();
}

__global__ void __launch_bounds__(kBlock) l1_log_loss_k(const float* __restrict__ pred,
                                                        const float* __restrict__ target, uint32_t n, float scale,
                                                        bf16* __restrict__ dpred, float* __restrict__ loss) {
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
    if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

hipError_t l1_log_loss(const float* pred, const float* target, Dims d, bf16* dpred, float* loss, hipStream_t s) {
    const uint32_t n = d.pixels();
    if (n == 0) return hipSuccess;
    const float scale = 1.0f / (float(n) * 3.0f);
    hipLaunchKernelGGL(l1_log_loss_k, dim3((n + kBlock - 1u) / kBlock), dim3(kBlock), 0, s, pred, target, n, scale,
                       dpred, loss);
    return hipGetLastError();
}

__device__ __forceinline__ unsigned char unorm8n(float v) {
    const float c = fminf(fmaxf(v, 0.0f), 1.0f);
    return (unsigned char)(c * 255.0f + 0.5f);
}

__global__ void __launch_bounds__(kBlock) unpack_frame_k(const float* __restrict__ pred, Dims d, uint32_t fw,
                                                         uint32_t fh, float exposure, float4* __restrict__ fin,
                                                         uchar4* __restrict__ out) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= fw * fh) return;
    const uint32_t x = i % fw, y = i / fw;
    const float* q = pred + (size_t(y) * d.w + x) * 3u;
    const float r = fmaxf(expm1f(q[0]), 0.0f), g = fmaxf(expm1f(q[1]), 0.0f), bl = fmaxf(expm1f(q[2]), 0.0f);
    fin[i] = make_float4(r, g, bl, 1.0f);
    if (out) {
        float o[3] = {r * exposure, g * exposure, bl * exposure};
#pragma unroll
        for (int k = 0; k < 3; ++k) o[k] = powf(o[k] / (o[k] + 0.3f), 0.4545f);
        out[i] = make_uchar4(unorm8n(o[0]), unorm8n(o[1]), unorm8n(o[2]), 255);
    }
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
