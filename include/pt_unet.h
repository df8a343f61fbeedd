/* pt_unet.h — the learned denoiser of libpt: a 5-level U-Net (3x3 conv +
 * ReLU, 2x average pooling / nearest upsampling, skip concatenation;
 * channels 32/48/64/96/128) over noisy radiance plus albedo, normal and
 * depth, trained and run on the device in bf16 with fp32 accumulation.
 *
 * The reference has no counterpart: its denoiser is the a-trous filter
 * (denoiser.wgsl, Denoiser.execute tracer.js:2583-2637), which pt_denoise
 * replaces.  pt_unet_denoise is the learned alternative for the same frame
 * and writes the same outputs (pt_read_denoised, pt_read_output), so a caller
 * switches between them per frame.
 *
 * Tensors crossing this boundary are host fp32 NHWC: features
 * [n][h][w][10] = radiance rgb, albedo rgb, normal xyz, hit distance (as the
 * accumulation buffer and the G-buffer hold them); targets [n][h][w][3] =
 * reference radiance; predictions [n][h][w][3] in log(1 + radiance).  h and w
 * are multiples of 16.  Every call returns PT_OK or a negative PT_ERR_*. */
#ifndef PT_UNET_H
#define PT_UNET_H

#include <stdint.h>

#include "pt_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Creates (or re-creates) the network with He-initialised weights drawn from
 * `seed`, and clears the Adam state. */
int pt_unet_init(pt_ctx* ctx, uint64_t seed);

/* Parameters as one fp32 array: per layer the weights [9][cout][cin] then the
 * biases [cout], layers in network order (19 convolutions). */
int pt_unet_param_count(pt_ctx* ctx, uint64_t* count);
int pt_unet_get_params(pt_ctx* ctx, float* out, uint64_t count);
int pt_unet_set_params(pt_ctx* ctx, const float* in, uint64_t count); /* also clears the Adam state */

/* The Adam state, to resume training in another process: the first and
 * second moments (param_count floats each, laid out as the parameters) and
 * the number of steps taken.  pt_unet_set_adam after pt_unet_set_params
 * restores a run exactly: the next pt_unet_train gives the same bits as the
 * run that saved it. */
int pt_unet_get_adam(pt_ctx* ctx, float* m, float* v, uint64_t count, uint32_t* steps);
int pt_unet_set_adam(pt_ctx* ctx, const float* m, const float* v, uint64_t count, uint32_t steps);

/* Inference on host features: pred [n][h][w][3] (log space). */
int pt_unet_forward(pt_ctx* ctx, const float* feat, uint32_t n, uint32_t h, uint32_t w, float* pred);

/* L1 loss on log(1 + x) of one batch, mean over pixels and channels, and its
 * gradient with respect to every parameter (grad: param_count floats, or
 * NULL). */
int pt_unet_loss_grad(pt_ctx* ctx, const float* feat, const float* target, uint32_t n, uint32_t h, uint32_t w,
                      float* loss, float* grad);

/* Keeps a training set on the device: n crops of h x w. */
int pt_unet_set_dataset(pt_ctx* ctx, const float* feat, const float* target, uint32_t n, uint32_t h, uint32_t w);

/* The training set built on the device from rendered frames: room for n
 * crops of h x w (cleared), then one call per crop and kind copying the h x w
 * window at (x0, y0) of the current frame: PT_UNET_INPUT packs the
 * accumulated radiance and the G-buffer (as pt_unet_denoise reads them),
 * PT_UNET_TARGET the accumulated radiance.  Replaces a pt_unet_set_dataset
 * set. */
enum { PT_UNET_INPUT = 0, PT_UNET_TARGET = 1 };
int pt_unet_dataset_reserve(pt_ctx* ctx, uint32_t n, uint32_t h, uint32_t w);
int pt_unet_dataset_capture(pt_ctx* ctx, uint32_t index, int what, uint32_t x0, uint32_t y0);

/* `steps` Adam steps (lr, beta 0.9 / 0.999, eps 1e-8) on batches of `batch`
 * crops drawn from the dataset with `seed`; losses[i] (optional) receives
 * step i's batch loss. */
int pt_unet_train(pt_ctx* ctx, uint32_t steps, uint32_t batch, float lr, uint64_t seed, float* losses);

/* Denoises the context's current frame (accumulation buffer + G-buffer, as
 * pt_denoise reads them): the result goes to the buffers pt_read_denoised
 * (fp32) and pt_read_output (rgba8, tonemapped with `exposure`) read. */
int pt_unet_denoise(pt_ctx* ctx, float exposure);

/* Inference precision of pt_unet_denoise and pt_unet_forward.  PT_UNET_FP8
 * calibrates on the context's current frame (accumulation buffer + G-buffer,
 * as pt_unet_denoise reads it) and from then on runs 16 of the 19
 * convolutions on fp8 products (OCP e4m3, fp32 accumulation; the first, the
 * last decoder and the output convolution stay bf16): each layer input is
 * quantized per channel as x / s_a[ci], s_a = a / 224 where a is the calibration
 * frame's 99.995th percentile of |x| (two bf16 forward passes: the maxima,
 * then 1024-bin histograms up to them); one binade of headroom, larger values
 * saturate at 448; the weights are
 * stored in e4m3 as w * s_a[ci] / s_w[co], s_w = that row's max / 448.
 * Between two fp8 layers the activations are held as those e4m3 values (the
 * bf16 output quantized by its producer with the reader's s_a), elsewhere as
 * bf16.  Training (pt_unet_train,
 * pt_unet_loss_grad) always runs in bf16; pt_unet_init, pt_unet_set_params
 * and pt_unet_train return inference to PT_UNET_BF16. */
enum { PT_UNET_BF16 = 0, PT_UNET_FP8 = 1 };
int pt_unet_set_precision(pt_ctx* ctx, int precision);

/* The fp8 state (PT_UNET_FP8 only), all 19 layers in network order (the
 * entries of the three bf16 layers are computed but not used): the input
 * scales s_a (sum of the layers' input channels, 1472 floats), the weight
 * scales s_w (sum of output channels, 1232 floats) and the e4m3 weights
 * [9][cout][cin] per layer (1,020,672 bytes), with the padded channel counts
 * (input 16, output 16). */
int pt_unet_fp8_state(pt_ctx* ctx, float* act_scales, uint64_t n_act, float* w_scales, uint64_t n_w,
                      uint8_t* weights, uint64_t n_bytes);

#ifdef __cplusplus
}
#endif

#endif /* PT_UNET_H */
