// nn_internal.h — kernels of the learned denoiser (nn_kernels.hip) as seen
// from the host code that runs the network (pt_unet.cpp).
//
// Activations and activation gradients are NHWC bf16 with a channel stride
// and offset per tensor (so a layer can read or write a slice of a
// concatenation buffer); weights are bf16 [tap][cout][cin] with fp32 master
// copies; every product accumulates in fp32 on MFMA.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pt {
namespace nn {

typedef __bf16 bf16;

// An NHWC view: element (pixel, c) at p[pixel * stride + off + c].
struct Tens {
    bf16* p;
    uint32_t stride;
    uint32_t off;
};

struct Dims {
    uint32_t n, h, w;
    __host__ __device__ uint32_t pixels() const { return n * h * w; }
};

// Largest output-channel count of a 3x3 convolution (per launch).
constexpr uint32_t kMaxCout = 128;

// out = act(conv3x3(in, w) + bias) with zero padding; cin and cout are
// multiples of 16 (cout <= kMaxCout); w is bf16 [9][cout][cin].  relu: ReLU;
// out_f32 (when non-null) receives the first `f32_channels` channels as fp32
// [pixel][f32_channels] instead of `out`.  bias may be null.
// Work done with a convolution's output (in its epilogue where the output
// tile is staged, else by the separate kernel after it):
//   pool.p: (bf16 output, d.h and d.w even) also the 2x2 average pool of the
//           output, as pool2 writes it;
//   upw.p:  (bf16 output) also its nearest 2x upsampling (dims 2d), as up2
//           writes it;
//   up.p:   (ReLU, d even) input channels [0, cup) are read from `up` (dims
//           d / 2) at the nearest coarse pixel, as up2 would have written
//           them; channels [cup, cin) from channels [0, cin - cup) of `in`;
//   fin:    (fp32 output of 3 channels, one image) the denoised frame texels
//           as unpack_frame writes them, for the fw x fh frame, instead of
//           out_f32 (which is still required, and written when the texels
//           cannot come from the kernel's stage);
//   wup:    (with up.p) the coarse channels' merged weights (up_merge_weights):
//           where the tiles are full, the coarse part runs as 2x2 products on
//           the coarse map instead of 3x3 products on its upsampling.
struct ConvFuse {
    Tens pool;
    Tens upw;
    Tens up;
    uint32_t cup;
    float4* fin;
    uchar4* ldr;
    uint32_t fw, fh;
    float exposure;
    const bf16* wup = nullptr;
    // (conv3x3_fp8 only) in8: `in` and up hold e4m3 bytes quantized with the
    // layer's inv_scale (strides and offsets in bytes); q_out / q_pool: write
    // the output / the pooled output as e4m3 bytes quantized with these
    // per-channel inverse scales (16-byte aligned; strides and offsets in bytes)
    bool in8 = false;
    const float* q_out = nullptr;
    const float* q_pool = nullptr;
};

// A 3x3 convolution over the nearest 2x upsampling of a coarse map, per
// output pixel parity (py, px): its taps land on 2x2 coarse pixels, coarse
// offset (a + py - 1, b + px - 1) for a, b in {0, 1}.  The merged weight of
// class 2 py + px, tap 2 a + b, [co][ci] (bf16, 16 x cout x cup in all) = the
// sum, in fp32, of the taps w[dy][dx][co][ci] (ci < cup of w [9][cout][cin])
// whose upsampled pixel lies in that coarse pixel.  (Equal to the 3x3 form up
// to the rounding of the sums; the image borders agree too, since d.h and d.w
// are even.)  Stored in the order the convolution's MFMA fragments read them:
// [class][K-step][16-channel output tile][lane][8] (cup a multiple of 8, cout
// of 16).
size_t up_merged_elems(uint32_t cout, uint32_t cup);
hipError_t up_merge_weights(const bf16* w, uint32_t cout, uint32_t cin, uint32_t cup, bf16* wup, hipStream_t s);
hipError_t conv3x3(const Tens& in, uint32_t cin, const bf16* w, const float* bias, const Tens& out, uint32_t cout,
                   Dims d, bool relu, float* out_f32, uint32_t f32_channels, hipStream_t s,
                   const ConvFuse* fuse = nullptr);

// The same convolution for inference on fp8 (OCP e4m3) operands with fp32
// accumulation: the input (bf16, as above) is quantized per input channel as
// clamp(x * inv_scale[ci], -448, 448) rounded to nearest even; w is e4m3
// [9][cout][cin] (cin-contiguous rows, 16-byte aligned, with the input
// scales folded in), and the accumulated product is multiplied by
// wscale[co] before the bias.  inv_scale has cin floats (16-byte aligned).
// fuse: pool and up/cup as for conv3x3 (up needs a ReLU layer), upw as a
// pass after the convolution; no frame texels.
hipError_t conv3x3_fp8(const Tens& in, uint32_t cin, const uint8_t* w, const float* inv_scale, const float* wscale,
                       const float* bias, const Tens& out, uint32_t cout, Dims d, bool relu, float* out_f32,
                       uint32_t f32_channels, hipStream_t s, const ConvFuse* fuse = nullptr);

// amax[ch] = max(amax[ch], max over P pixels of |x[p][ch]|) for ch < c
// (c a multiple of 8, at most 256), the maxima as float bit patterns.
hipError_t chan_amax(const Tens& x, uint32_t c, uint32_t P, uint32_t* amax, hipStream_t s);
// Per-channel histograms of |x| in kHistBins bins over [0, amax[ch]] (amax as
// float bits, from chan_amax), added to hist[ch * kHistBins + bin].
constexpr uint32_t kHistBins = 1024;
hipError_t chan_hist(const Tens& x, uint32_t c, uint32_t P, const uint32_t* amax, uint32_t* hist, hipStream_t s);

// Backward of a 3x3 convolution (cout output channels) over an
// input-channel slice, except the data gradient: with dz = ga * (a > 0)
// when a is given (ReLU output), else ga,
//   dw[tap][co][ci_off + ci] (fp32, rows of cin_total, overwritten) = sum over
//     pixels of dz[p][co] * x[p + offset(tap)][ci] for ci < cin (x the
//     slice's NHWC view, zero outside the image; cin <= kMaxCout),
//   db[co] (fp32, overwritten; when db is non-null) = sum over pixels of dz,
//   dz_out (when non-null) = dz, NHWC contiguous [pixel][cout].
// partial holds wgrad_partial_floats(cin, cout, d) floats of scratch.
size_t wgrad_partial_floats(uint32_t cin, uint32_t cout, Dims d);
hipError_t conv3x3_wgrad(const Tens& ga, const Tens* a, uint32_t cout, const Tens& x, uint32_t cin, Dims d,
                         bf16* dz_out, float* partial, float* dw, uint32_t ci_off, uint32_t cin_total, float* db,
                         hipStream_t s);

// 2x2 average pooling (d: the input's dims; h and w even).
hipError_t pool2(const Tens& in, uint32_t c, Dims d, const Tens& out, hipStream_t s);
// din += dout / 4 spread over each 2x2 block (accumulates into din).
hipError_t pool2_back_add(const Tens& dout, uint32_t c, Dims d, const Tens& din, hipStream_t s);
// 2x nearest upsampling (d: the input's dims).
hipError_t up2(const Tens& in, uint32_t c, Dims d, const Tens& out, hipStream_t s);
// din = sum of dout over each 2x2 block.
hipError_t up2_back(const Tens& dout, uint32_t c, Dims d, const Tens& din, hipStream_t s);

// The weights the data-gradient pass uses, for input channels
// [ci_begin, ci_begin + ci_count): wd[8 - tap][ci - ci_begin][co] = w[tap][co][ci].
hipError_t transpose_weights(const bf16* w, uint32_t cout, uint32_t cin, uint32_t ci_begin, uint32_t ci_count,
                             bf16* wd, hipStream_t s);

// fp32 master weights (layout [9][cout][cin]) to bf16.
hipError_t to_bf16(const float* src, bf16* dst, size_t n, hipStream_t s);
hipError_t fill_zero_bf16(bf16* p, size_t n, hipStream_t s);
hipError_t fill_zero_f32(float* p, size_t n, hipStream_t s);

// Adam over n parameters; step t >= 1.
hipError_t adam(float* w, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2, float eps,
                uint32_t t, hipStream_t s);

// Network input: 16 bf16 channels per pixel from fp32 features [pixel][10]
// (radiance rgb, albedo rgb, normal xyz, depth): log(1 + radiance), albedo,
// normal, depth / (depth + 1), then zeros.
hipError_t pack_input(const float* feat, Dims d, bf16* out, hipStream_t s);
// The same input straight from a rendered frame: accum (float4 per pixel),
// G-buffer (48 B texels), the d.w x d.h window at (x0, y0) of the fw x fh
// frame (d.n = 1; pixels outside the frame are zero, which pads a frame up
// to a multiple of 16).
hipError_t pack_frame(const float4* accum, const void* gbuf, uint32_t fw, uint32_t fh, uint32_t x0, uint32_t y0,
                      Dims d, bf16* out, hipStream_t s);
// Radiance of the d.w x d.h window at (x0, y0) (inside the frame) as fp32 [pixel][3].
hipError_t crop_radiance(const float4* accum, uint32_t fw, uint32_t fh, uint32_t x0, uint32_t y0, Dims d, float* out,
                         hipStream_t s);

// L1 loss on log(1 + x): pred [pixel][3] fp32 is in log space already, target
// [pixel][3] is linear radiance.  Writes d loss / d pred as bf16 [pixel][16]
// (channels 3..15 zero) and adds the batch loss sum into *loss (fp32).
// L1 loss of pred against log1p(max(target, 0)): its gradient sign into
// dpred, per-block partial sums into part (loss_partials(pixels) floats),
// and their sum added to *loss.
uint32_t loss_partials(uint32_t pixels);
hipError_t l1_log_loss(const float* pred, const float* target, Dims d, bf16* dpred, float* part, float* loss,
                       hipStream_t s);

// Denoised frame: exp(pred) - 1 (clamped at 0) into fp32 float4 and rgba8
// (tonemapped with `exposure`), cropping the padded frame to fw x fh.
hipError_t unpack_frame(const float* pred, Dims d, uint32_t fw, uint32_t fh, float exposure, float4* fin,
                        uchar4* out, hipStream_t s);

}  // namespace nn
}  // namespace pt
