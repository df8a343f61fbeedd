// pt_unet.cpp — the learned denoiser on the host side: the layer table,
// device buffers, forward and backward passes, Adam, and the C ABI of
// include/pt_unet.h.  The kernels are in nn_kernels.hip.
//
// Network (levels L = 0..4 at 1/2^L resolution, C = 32/48/64/96/128):
//   encoder L < 4:  x -> conv(C_L) -> conv(C_L) = skip_L -> pool -> level L+1
//   level 4:        conv(128) -> conv(128)
//   decoder L < 4:  up(level L+1) ++ skip_L -> conv(C_L) -> conv(C_L)
//   output:         conv(3), linear, in log(1 + radiance)
// skip_L is written straight into the decoder's concatenation buffer
// cat_L = [up | skip] (NHWC, so concatenation is a channel offset).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "fp8.h"
#include "nn_internal.h"
#include "pt_ctx.h"
#include "pt_unet.h"

using namespace pt;
using namespace pt::nn;

namespace pt {

namespace {

constexpr uint32_t kC[5] = {32, 48, 64, 96, 128};
constexpr uint32_t kInC = 16;     // packed input channels (10 used)
constexpr uint32_t kOutC = 16;    // output convolution channels (3 used)
constexpr uint32_t kFeat = 10;    // host feature channels
constexpr uint32_t kLayers = 19;

struct Layer {
    uint32_t cin, cout;            // padded (multiples of 16)
    uint32_t cin_real, cout_real;  // fan-in / used outputs for initialisation
    size_t w_off, b_off;           // into the flat fp32 parameter array
    size_t wbf_off;                // into the bf16 weight arrays (same size as the weights)
    size_t a_off, s_off;           // into the per-input-channel and per-output-channel fp8 scale arrays
    bool relu;
    uint32_t cup = 0;              // a decoder level's first convolution: its upsampled (coarse) channels
    size_t wup_off = 0;            // into the merged coarse weights (up_merge_weights)
};

enum {
    E0A, E0B, E1A, E1B, E2A, E2B, E3A, E3B, B4A, B4B, D3A, D3B, D2A, D2B, D1A, D1B, D0A, D0B, OUT
};

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double uniform01(uint64_t& s) { return double(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

struct UNet {
    Layer L[kLayers];
    size_t n_params = 0, n_wbf = 0;
    DevBuf params, grads, m, v;    // fp32, n_params
    DevBuf wbf, wdt;               // bf16 weights [9][cout][cin], transposed slices [9][ci][cout]
    size_t n_wup = 0;
    DevBuf wup;                    // the decoder-first layers' coarse weights merged per output parity (bf16)
    uint32_t adam_t = 0;
    // fp8 inference: e4m3 weights (layout of wbf), input-channel inverse scales, output-channel scales
    size_t n_act = 0, n_out = 0;   // sums of cin and cout over the layers
    bool fp8 = false;              // pt_unet_denoise / pt_unet_forward run on fp8
    uint32_t* calib = nullptr;     // set while a forward pass records per-channel input maxima
    uint32_t* calib_hist = nullptr;  // set while a forward pass histograms |x| up to those maxima
    DevBuf w8, inv_sa, sw, amax, hist;
    std::vector<float> h_sa, h_sw;
    std::vector<uint8_t> h_w8;
    // activations for the current dims
    Dims dims[5] = {};
    bool have_acts = false;
    DevBuf in0, tA[4], cat[4], pin[5], t4A, b4, dA[4], dd[4], pred;
    DevBuf g_tA[4], g_cat[4], g_pin[5], g_t4A, g_b4, g_dA[4], g_dd[4], g_pred, dz;
    DevBuf target, feat, loss, lpart, wpart;
    // dataset
    DevBuf ds_in, ds_tg;
    uint32_t ds_n = 0, ds_h = 0, ds_w = 0;

    std::vector<DevBuf*> all_bufs() {
        std::vector<DevBuf*> bufs{&params, &grads, &m, &v, &wbf, &wdt, &wup, &in0, &t4A, &b4, &pred, &g_t4A, &g_b4,
                                  &g_pred, &dz, &target, &feat, &loss, &wpart, &ds_in,
                                  &ds_tg, &w8, &inv_sa, &sw, &amax, &hist};
        for (int i = 0; i < 4; ++i) {
            for (DevBuf* b : {&tA[i], &cat[i], &dA[i], &dd[i], &g_tA[i], &g_cat[i], &g_dA[i], &g_dd[i]})
                bufs.push_back(b);
        }
        for (int i = 0; i < 5; ++i) {
            bufs.push_back(&pin[i]);
            bufs.push_back(&g_pin[i]);
        }
        return bufs;
    }
    void release_acts() {
        for (DevBuf* b : all_bufs())
            if (b != &params && b != &grads && b != &m && b != &v && b != &wbf && b != &wdt && b != &wup && b != &ds_in &&
                b != &ds_tg && b != &w8 && b != &inv_sa && b != &sw && b != &amax && b != &hist)
                b->release();
        have_acts = false;
    }
};

void unet_free(UNet* u) {
    if (!u) return;
    for (DevBuf* b : u->all_bufs()) b->release();
    delete u;
}

namespace {

bf16* bfp(DevBuf& b) { return static_cast<bf16*>(b.p); }
float* fp(DevBuf& b) { return static_cast<float*>(b.p); }
Tens T(DevBuf& b, uint32_t stride, uint32_t off = 0) { return Tens{bfp(b), stride, off}; }

void build_layers(UNet& u) {
    struct Spec { uint32_t cin, cout, cin_real, cout_real; bool relu; };
    const Spec spec[kLayers] = {
        {kInC, kC[0], kFeat, kC[0], true},  {kC[0], kC[0], kC[0], kC[0], true},
        {kC[0], kC[1], kC[0], kC[1], true}, {kC[1], kC[1], kC[1], kC[1], true},
        {kC[1], kC[2], kC[1], kC[2], true}, {kC[2], kC[2], kC[2], kC[2], true},
        {kC[2], kC[3], kC[2], kC[3], true}, {kC[3], kC[3], kC[3], kC[3], true},
        {kC[3], kC[4], kC[3], kC[4], true}, {kC[4], kC[4], kC[4], kC[4], true},
        {kC[4] + kC[3], kC[3], kC[4] + kC[3], kC[3], true}, {kC[3], kC[3], kC[3], kC[3], true},
        {kC[3] + kC[2], kC[2], kC[3] + kC[2], kC[2], true}, {kC[2], kC[2], kC[2], kC[2], true},
        {kC[2] + kC[1], kC[1], kC[2] + kC[1], kC[1], true}, {kC[1], kC[1], kC[1], kC[1], true},
        {kC[1] + kC[0], kC[0], kC[1] + kC[0], kC[0], true}, {kC[0], kC[0], kC[0], kC[0], true},
        {kC[0], kOutC, kC[0], 3, false},
    };
    size_t off = 0, woff = 0, aoff = 0, soff = 0;
    for (uint32_t i = 0; i < kLayers; ++i) {
        Layer& l = u.L[i];
        l.a_off = aoff;
        aoff += spec[i].cin;
        l.s_off = soff;
        soff += spec[i].cout;
        l.cin = spec[i].cin;
        l.cout = spec[i].cout;
        l.cin_real = spec[i].cin_real;
        l.cout_real = spec[i].cout_real;
        l.relu = spec[i].relu;
        l.w_off = off;
        off += size_t(9) * l.cout * l.cin;
        l.b_off = off;
        off += l.cout;
        l.wbf_off = woff;
        woff += size_t(9) * l.cout * l.cin;
    }
    // decoder level l's first convolution reads kC[l + 1] upsampled channels
    size_t uoff = 0;
    for (int lv = 3; lv >= 0; --lv) {
        Layer& l = u.L[D3A + 2 * (3 - lv)];
        l.cup = kC[lv + 1];
        l.wup_off = uoff;
        uoff += up_merged_elems(l.cout, l.cup);
    }
    u.n_wup = uoff;
    u.n_params = off;
    u.n_wbf = woff;
    u.n_act = aoff;
    u.n_out = soff;
}

// He initialisation; padded input channels and unused outputs stay zero.
void init_params(const UNet& u, uint64_t seed, std::vector<float>& p) {
    p.assign(u.n_params, 0.0f);
    uint64_t s = seed * 0x2545F4914F6CDD1Dull + 1u;
    for (uint32_t i = 0; i < kLayers; ++i) {
        const Layer& l = u.L[i];
        const double sd = std::sqrt(2.0 / (9.0 * l.cin_real));
        for (uint32_t tap = 0; tap < 9; ++tap)
            for (uint32_t co = 0; co < l.cout; ++co)
                for (uint32_t ci = 0; ci < l.cin; ++ci) {
                    const double u1 = std::max(uniform01(s), 1e-300), u2 = uniform01(s);
                    const double g = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
                    const bool used = ci < l.cin_real && co < l.cout_real;
                    p[l.w_off + (size_t(tap) * l.cout + co) * l.cin + ci] = used ? float(g * sd) : 0.0f;
                }
    }
}

// Number of 128-channel slices of a layer's input (the data-gradient and
// weight-gradient kernels produce at most 128 channels per launch).
uint32_t slices(const Layer& l) { return (l.cin + kMaxCout - 1u) / kMaxCout; }
uint32_t slice_begin(uint32_t k) { return k * kMaxCout; }
uint32_t slice_count(const Layer& l, uint32_t k) { return std::min(kMaxCout, l.cin - k * kMaxCout); }

int hip_rc(pt_ctx* c, hipError_t e, const char* what) { return e == hipSuccess ? PT_OK : c->hip_fail(e, what); }

// bf16 weights and their transposed slices from the fp32 master copy.
int refresh_weights(pt_ctx* c, UNet& u) {
    hipStream_t s = c->stream;
    for (uint32_t i = 0; i < kLayers; ++i) {
        const Layer& l = u.L[i];
        hipError_t e = to_bf16(fp(u.params) + l.w_off, bfp(u.wbf) + l.wbf_off, size_t(9) * l.cout * l.cin, s);
        if (e == hipSuccess && l.cup)
            e = up_merge_weights(bfp(u.wbf) + l.wbf_off, l.cout, l.cin, l.cup, bfp(u.wup) + l.wup_off, s);
        size_t so = l.wbf_off;
        for (uint32_t k = 0; e == hipSuccess && k < slices(l); ++k) {
            e = transpose_weights(bfp(u.wbf) + l.wbf_off, l.cout, l.cin, slice_begin(k), slice_count(l, k),
                                  bfp(u.wdt) + so, s);
            so += size_t(9) * slice_count(l, k) * l.cout;
        }
        if (e != hipSuccess) return c->hip_fail(e, "unet weights");
    }
    return PT_OK;
}

int ensure_net(pt_ctx* c) {
    if (c->unet) return PT_OK;
    return c->fail(PT_ERR_STATE, "the learned denoiser is not initialised (pt_unet_init)");
}

int alloc_acts(pt_ctx* c, UNet& u, Dims d0) {
    if (d0.h % 16u || d0.w % 16u || d0.n == 0 || d0.h == 0 || d0.w == 0)
        return c->fail(PT_ERR_ARG, "network input must be a positive multiple of 16 in both dimensions");
    if (u.have_acts && u.dims[0].n == d0.n && u.dims[0].h == d0.h && u.dims[0].w == d0.w) return PT_OK;
    u.release_acts();
    for (int l = 0; l < 5; ++l) u.dims[l] = Dims{d0.n, d0.h >> l, d0.w >> l};
    auto P = [&](int l) { return size_t(u.dims[l].pixels()); };
    int rc;
    auto A = [&](DevBuf& b, size_t elems, size_t each, const char* what) { return c->alloc(b, elems * each, what); };
    if ((rc = A(u.in0, P(0) * kInC, 2, "unet input"))) return rc;
    for (int l = 0; l < 4; ++l) {
        const size_t cc = kC[l + 1] + kC[l];
        if ((rc = A(u.tA[l], P(l) * kC[l], 2, "unet act")) || (rc = A(u.cat[l], P(l) * cc, 2, "unet cat")) ||
            (rc = A(u.dA[l], P(l) * kC[l], 2, "unet act")) || (rc = A(u.dd[l], P(l) * kC[l], 2, "unet act")) ||
            (rc = A(u.g_tA[l], P(l) * kC[l], 2, "unet grad")) || (rc = A(u.g_cat[l], P(l) * cc, 2, "unet grad")) ||
            (rc = A(u.g_dA[l], P(l) * kC[l], 2, "unet grad")) || (rc = A(u.g_dd[l], P(l) * kC[l], 2, "unet grad")))
            return rc;
        if ((rc = A(u.pin[l + 1], P(l + 1) * kC[l], 2, "unet pool")) ||
            (rc = A(u.g_pin[l + 1], P(l + 1) * kC[l], 2, "unet grad")))
            return rc;
    }
    if ((rc = A(u.t4A, P(4) * kC[4], 2, "unet act")) || (rc = A(u.b4, P(4) * kC[4], 2, "unet act")) ||
        (rc = A(u.g_t4A, P(4) * kC[4], 2, "unet grad")) || (rc = A(u.g_b4, P(4) * kC[4], 2, "unet grad")) ||
        (rc = A(u.pred, P(0) * 3, 4, "unet pred")) || (rc = A(u.g_pred, P(0) * kOutC, 2, "unet grad")) ||
        (rc = A(u.target, P(0) * 3, 4, "unet target")) || (rc = A(u.feat, P(0) * kFeat, 4, "unet features")))
        return rc;
    // backward scratch, sized for the largest layer: dz and the partial sums of
    // the weight and bias gradients
    size_t dz = 0, wp = 0;
    const int lvl[kLayers] = {0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 3, 3, 2, 2, 1, 1, 0, 0, 0};
    for (uint32_t i = 0; i < kLayers; ++i) {
        const Layer& l = u.L[i];
        const Dims d = u.dims[lvl[i]];
        dz = std::max(dz, size_t(d.pixels()) * l.cout);
        for (uint32_t k = 0; k < slices(l); ++k) wp = std::max(wp, wgrad_partial_floats(slice_count(l, k), l.cout, d));
    }
    if ((rc = A(u.dz, dz, 2, "unet dz")) ||
        (rc = A(u.wpart, wp, 4, "unet gradient scratch")) ||
        (rc = A(u.loss, 1024, 4, "unet loss")) ||
        (rc = A(u.lpart, loss_partials(uint32_t(P(0))), 4, "unet loss partials")))
        return rc;
    u.have_acts = true;
    return PT_OK;
}

// e8 (fp8 inference with the coarse-read layout): activations between two
// fp8 layers are e4m3 bytes in memory, quantized by the producing layer with
// the inverse input scales of the layer that reads them (the same bytes the
// reader would compute from the bf16 values); the buffers are the bf16 ones,
// used at one byte per channel.
int conv_fwd(pt_ctx* c, UNet& u, bool q8, uint32_t li, const Tens& in, const Tens& out, Dims d,
             float* of32 = nullptr, const ConvFuse* fuse = nullptr, bool e8 = false) {
    const Layer& l = u.L[li];
    if (u.calib) {
        const hipError_t e = chan_amax(in, l.cin, d.pixels(), u.calib + l.a_off, c->stream);
        if (e != hipSuccess) return c->hip_fail(e, "unet calibration");
    }
    if (u.calib_hist) {
        const hipError_t e = chan_hist(in, l.cin, d.pixels(), static_cast<const uint32_t*>(u.amax.p) + l.a_off,
                                       u.calib_hist + size_t(l.a_off) * kHistBins, c->stream);
        if (e != hipSuccess) return c->hip_fail(e, "unet calibration");
    }
    // the first convolution, the last decoder convolution and the output
    // convolution stay bf16: in e4m3 the input features, the full-resolution
    // features the output is read from and the output radiance lose the most
    // (first and output alone measured at +12% and +3..9% relMSE of the
    // denoised frame; with the weights trained on 4096-spp targets the last
    // decoder convolution in e4m3 took a 1080p frame from 1.13x to 1.93x the
    // bf16 relMSE)
    if (q8 && li != E0A && li != D0B && li != OUT) {
        ConvFuse f8{};
        if (fuse) f8 = *fuse;
        if (e8) {
            const bool enc_b = li <= E3B && li % 2u == 1u;  // its output is a skip; its pooled copy feeds li + 1
            const uint32_t lvl = (li - 1u) / 2u;
            const uint32_t next = enc_b ? D3A + 2u * (3u - lvl) : li + 1u;  // the layer reading the output
            f8.in8 = true;  // (the first layer writes e4m3 for the second, see below)
            if (next != D0B && next != OUT)
                f8.q_out = fp(u.inv_sa) + u.L[next].a_off + (enc_b ? kC[lvl + 1] : 0u);
            if (enc_b) f8.q_pool = fp(u.inv_sa) + u.L[li + 1u].a_off;
        }
        const hipError_t e = conv3x3_fp8(in, l.cin, static_cast<const uint8_t*>(u.w8.p) + l.wbf_off,
                                         fp(u.inv_sa) + l.a_off, fp(u.sw) + l.s_off, fp(u.params) + l.b_off, out,
                                         l.cout, d, l.relu, of32, of32 ? 3u : 0u, c->stream, &f8);
        return hip_rc(c, e, "unet conv fp8");
    }
    // the first convolution stays bf16 but writes its output as e4m3 for the
    // second, an fp8 layer
    ConvFuse fq{};
    if (e8 && li == E0A) {
        if (fuse) fq = *fuse;
        fq.q_out = fp(u.inv_sa) + u.L[E0B].a_off;
        fuse = &fq;
    }
    return hip_rc(c,
                  conv3x3(in, l.cin, bfp(u.wbf) + l.wbf_off, fp(u.params) + l.b_off, out, l.cout, d, l.relu, of32,
                          of32 ? 3u : 0u, c->stream, fuse),
                  "unet conv");
}

// The forward pass over u.in0; q8: the fp8 convolutions (inference only);
// infer: no backward follows.  The encoder's second convolutions also write
// the pooled copy of their output (in the convolution's epilogue where its
// output tile is staged).  For inference each decoder level's first
// convolution reads the upsampled half of its input straight from the coarse
// map; otherwise (training, calibration) the convolution below writes the
// upsampled copy into the level's concatenation buffer.  out_fuse: work
// for the output convolution (the frame texels of pt_unet_denoise).
int forward(pt_ctx* c, UNet& u, bool q8 = false, const ConvFuse* out_fuse = nullptr, bool infer = false) {
    const Dims* D = u.dims;
    int rc;
    const Tens none{nullptr, 0, 0};
    // inference (no backward, no calibration) skips the upsampled
    // copies: the decoders' first convolutions read the coarse maps themselves
    // (1080p frame 1.07 -> 1.02 ms against writing the copy in the epilogue of
    // the convolution below), and the encoders write their skip outputs densely
    // (into the idle gradient buffers) instead of into the concatenation
    const bool coarse = infer && !u.calib && !u.calib_hist;
    const bool e8 = q8 && coarse;
    // encoder
    Tens x = T(u.in0, kInC);
    for (int l = 0; l < 4; ++l) {
        const uint32_t cc = kC[l + 1] + kC[l];
        const Tens skip = coarse ? T(u.g_tA[l], kC[l]) : T(u.cat[l], cc, kC[l + 1]);
        const ConvFuse fz{T(u.pin[l + 1], kC[l]), none, none, 0u, nullptr, nullptr, 0u, 0u, 0.0f};
        if ((rc = conv_fwd(c, u, q8, E0A + 2 * l, x, T(u.tA[l], kC[l]), D[l], nullptr, nullptr, e8))) return rc;
        if ((rc = conv_fwd(c, u, q8, E0B + 2 * l, T(u.tA[l], kC[l]), skip, D[l], nullptr, &fz, e8))) return rc;
        x = fz.pool;
    }
    // the bottleneck's output goes up into cat[3], each decoder level's (but the last) into the next one's
    if ((rc = conv_fwd(c, u, q8, B4A, x, T(u.t4A, kC[4]), D[4], nullptr, nullptr, e8))) return rc;
    const ConvFuse up3{none, coarse ? none : T(u.cat[3], kC[4] + kC[3], 0), none, 0u, nullptr, nullptr, 0u, 0u, 0.0f};
    if ((rc = conv_fwd(c, u, q8, B4B, T(u.t4A, kC[4]), T(u.b4, kC[4]), D[4], nullptr, &up3, e8))) return rc;
    for (int l = 3; l >= 0; --l) {
        const uint32_t cc = kC[l + 1] + kC[l];
        const Tens low = l == 3 ? T(u.b4, kC[4]) : T(u.dd[l + 1], kC[l + 1]);
        const ConvFuse rd{none, none, low, kC[l + 1], nullptr, nullptr, 0u, 0u, 0.0f,
                          bfp(u.wup) + u.L[D3A + 2 * (3 - l)].wup_off};
        const ConvFuse upl{none, l > 0 && !coarse ? T(u.cat[l - 1], kC[l] + kC[l - 1], 0) : none, none, 0u, nullptr,
                           nullptr, 0u, 0u, 0.0f};
        if ((rc = conv_fwd(c, u, q8, D3A + 2 * (3 - l), coarse ? T(u.g_tA[l], kC[l]) : T(u.cat[l], cc),
                           T(u.dA[l], kC[l]), D[l], nullptr, coarse ? &rd : nullptr, e8)))
            return rc;
        if ((rc = conv_fwd(c, u, q8, D3B + 2 * (3 - l), T(u.dA[l], kC[l]), T(u.dd[l], kC[l]), D[l], nullptr,
                           l > 0 ? &upl : nullptr, e8)))
            return rc;
    }
    return conv_fwd(c, u, q8, OUT, T(u.dd[0], kC[0]), none, D[0], fp(u.pred), out_fuse);
}

// Backward through layer li: x its input, a its output (for the ReLU mask),
// ga the gradient at its output; writes the weight and bias gradients and,
// when gx.p is non-null, the gradient at its input.
int conv_back(pt_ctx* c, UNet& u, uint32_t li, const Tens& x, const Tens& a, const Tens& ga, const Tens& gx, Dims d) {
    const Layer& l = u.L[li];
    hipStream_t s = c->stream;
    // the masked gradient dz goes to u.dz for the data gradient (a linear
    // layer's is ga itself); the first slice's pass also sums the bias
    const Tens dz = l.relu ? T(u.dz, l.cout) : ga;
    hipError_t e = hipSuccess;
    size_t so = l.wbf_off;
    for (uint32_t k = 0; e == hipSuccess && k < slices(l); ++k) {
        const uint32_t c0 = slice_begin(k), nc = slice_count(l, k);
        e = conv3x3_wgrad(ga, l.relu ? &a : nullptr, l.cout, Tens{x.p, x.stride, x.off + c0}, nc, d,
                          k == 0 && gx.p && l.relu ? bfp(u.dz) : nullptr, fp(u.wpart), fp(u.grads) + l.w_off, c0,
                          l.cin, k == 0 ? fp(u.grads) + l.b_off : nullptr, s);
        if (e == hipSuccess && gx.p) {
            const Tens gxs{gx.p, gx.stride, gx.off + c0};
            e = conv3x3(dz, l.cout, bfp(u.wdt) + so, nullptr, gxs, nc, d, false, nullptr, 0, s);
        }
        so += size_t(9) * nc * l.cout;
    }
    return hip_rc(c, e, "unet backward");
}

int backward(pt_ctx* c, UNet& u) {
    const Dims* D = u.dims;
    hipStream_t s = c->stream;
    int rc;
    const Tens none{nullptr, 0, 0};
    // output layer (linear): dz = d loss / d pred, already [pixel][16]
    if ((rc = conv_back(c, u, OUT, T(u.dd[0], kC[0]), none, T(u.g_pred, kOutC), T(u.g_dd[0], kC[0]), D[0])))
        return rc;
    // decoder, level 0 up to 3
    for (int l = 0; l < 4; ++l) {
        const uint32_t cc = kC[l + 1] + kC[l];
        const uint32_t la = D3A + 2 * (3 - l), lb = la + 1;
        if ((rc = conv_back(c, u, lb, T(u.dA[l], kC[l]), T(u.dd[l], kC[l]), T(u.g_dd[l], kC[l]),
                            T(u.g_dA[l], kC[l]), D[l])))
            return rc;
        if ((rc = conv_back(c, u, la, T(u.cat[l], cc), T(u.dA[l], kC[l]), T(u.g_dA[l], kC[l]), T(u.g_cat[l], cc),
                            D[l])))
            return rc;
        // the upsampled half of the concatenation goes back to the level below
        const Tens g_low = l < 3 ? T(u.g_dd[l + 1], kC[l + 1]) : T(u.g_b4, kC[4]);
        if ((rc = hip_rc(c, up2_back(T(u.g_cat[l], cc, 0), kC[l + 1], D[l + 1], g_low, s), "unet up back")))
            return rc;
    }
    // bottleneck
    if ((rc = conv_back(c, u, B4B, T(u.t4A, kC[4]), T(u.b4, kC[4]), T(u.g_b4, kC[4]), T(u.g_t4A, kC[4]), D[4])))
        return rc;
    if ((rc = conv_back(c, u, B4A, T(u.pin[4], kC[3]), T(u.t4A, kC[4]), T(u.g_t4A, kC[4]), T(u.g_pin[4], kC[3]),
                        D[4])))
        return rc;
    // encoder, level 3 down to 0: the skip gradient is the concatenation's
    // skip half plus the pooled path's gradient
    for (int l = 3; l >= 0; --l) {
        const uint32_t cc = kC[l + 1] + kC[l];
        const Tens g_skip = T(u.g_cat[l], cc, kC[l + 1]);
        if ((rc = hip_rc(c, pool2_back_add(T(u.g_pin[l + 1], kC[l]), kC[l], D[l], g_skip, s), "unet pool back")))
            return rc;
        const Tens xin = l == 0 ? T(u.in0, kInC) : T(u.pin[l], kC[l - 1]);
        const Tens gxin = l == 0 ? none : T(u.g_pin[l], kC[l - 1]);
        if ((rc = conv_back(c, u, E0B + 2 * l, T(u.tA[l], kC[l]), T(u.cat[l], cc, kC[l + 1]), g_skip,
                            T(u.g_tA[l], kC[l]), D[l])))
            return rc;
        if ((rc = conv_back(c, u, E0A + 2 * l, xin, T(u.tA[l], kC[l]), T(u.g_tA[l], kC[l]), gxin, D[l]))) return rc;
    }
    return PT_OK;
}

// Loss of the batch in u.pred against u.target; adds its sum into loss[slot].
int loss_and_grad(pt_ctx* c, UNet& u, uint32_t slot) {
    return hip_rc(c, l1_log_loss(fp(u.pred), fp(u.target), u.dims[0], bfp(u.g_pred), fp(u.lpart), fp(u.loss) + slot,
                                   c->stream),
                  "unet loss");
}

int upload_batch_features(pt_ctx* c, UNet& u, const float* feat, const float* target) {
    const size_t P = u.dims[0].pixels();
    hipError_t e = hipMemcpyAsync(u.feat.p, feat, P * kFeat * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = pack_input(fp(u.feat), u.dims[0], bfp(u.in0), c->stream);
    if (e == hipSuccess && target) e = hipMemcpyAsync(u.target.p, target, P * 3 * 4, hipMemcpyHostToDevice, c->stream);
    return hip_rc(c, e, "unet batch upload");
}

// The context's current frame (accumulation buffer + G-buffer), padded to
// multiples of 16, through the network into u.pred.
int frame_forward(pt_ctx* c, UNet& u, bool q8, const ConvFuse* out_fuse = nullptr) {
    const Dims d{1u, (c->height + 15u) & ~15u, (c->width + 15u) & ~15u};
    int rc;
    if ((rc = alloc_acts(c, u, d))) return rc;
    const hipError_t e = pack_frame(static_cast<const float4*>(c->accum.p), c->gbuf.p, c->width, c->height, 0u, 0u,
                                    d, bfp(u.in0), c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "unet input");
    return forward(c, u, q8, out_fuse, true);
}

// Calibrated amax of a layer input maps to this e4m3 value: one binade of
// headroom below the largest finite 448 for frames brighter than the
// calibration frame (beyond it values saturate).
constexpr float kFp8ActTarget = 224.0f;
// The calibrated amax of a channel is the |x| that all but this fraction of
// the calibration frame's pixels stay below (the top bin edge of its
// histogram reaching it), not the maximum: on a full 1080p frame a few
// hundred outlier activations per channel set the maximum and coarsen every
// other value's quantisation (relMSE +36% against bf16; at the percentile,
// fp8 tracks bf16 as a 192x112 frame's maximum does).
constexpr double kFp8CalibTail = 5e-5;

// fp8 inference state from the calibrated per-channel input amax of two bf16
// forward passes (maxima, then histograms up to them): per layer
// s_a[ci] = amax[ci] / 224 (1 for a channel that stayed 0),
// w'[tap][co][ci] = w * s_a[ci], s_w[co] = max |w'| / 448 (1 for an all-zero
// row), weights e4m3(w' / s_w[co]); the kernels multiply the input by
// 1 / s_a and the product by s_w.
int quantize_fp8(pt_ctx* c, UNet& u) {
    std::vector<uint32_t> am(u.n_act), hist(size_t(u.n_act) * kHistBins);
    std::vector<float> p(u.n_params);
    int rc;
    if ((rc = hip_rc(c, hipMemcpy(am.data(), u.amax.p, am.size() * 4, hipMemcpyDeviceToHost), "unet calibration")) ||
        (rc = hip_rc(c, hipMemcpy(hist.data(), u.hist.p, hist.size() * 4, hipMemcpyDeviceToHost),
                     "unet calibration")) ||
        (rc = hip_rc(c, hipMemcpy(p.data(), u.params.p, p.size() * 4, hipMemcpyDeviceToHost), "unet params")))
        return rc;
    // the percentile: the smallest top edge of bin b with at most tail x count above it
    for (uint32_t ch = 0; ch < u.n_act; ++ch) {
        float a;
        std::memcpy(&a, &am[ch], 4);
        if (!(a > 0.0f) || !std::isfinite(a)) continue;
        const uint32_t* h = hist.data() + size_t(ch) * kHistBins;
        uint64_t total = 0;
        for (uint32_t b = 0; b < kHistBins; ++b) total += h[b];
        const uint64_t allowed = uint64_t(kFp8CalibTail * double(total));
        uint64_t above = total;
        uint32_t b = 0;
        for (; b < kHistBins; ++b) {
            above -= h[b];
            if (above <= allowed) break;
        }
        const float edge = a * float(std::min(b + 1u, kHistBins)) / float(kHistBins);
        std::memcpy(&am[ch], &edge, 4);
    }
    u.h_sa.assign(u.n_act, 1.0f);
    u.h_sw.assign(u.n_out, 1.0f);
    u.h_w8.assign(u.n_wbf, 0);
    std::vector<float> inv(u.n_act), wr;
    for (uint32_t i = 0; i < kLayers; ++i) {
        const Layer& l = u.L[i];
        float* sa = u.h_sa.data() + l.a_off;
        for (uint32_t ci = 0; ci < l.cin; ++ci) {
            float a;
            std::memcpy(&a, &am[l.a_off + ci], 4);
            sa[ci] = (a > 0.0f && std::isfinite(a)) ? a / kFp8ActTarget : 1.0f;
            inv[l.a_off + ci] = 1.0f / sa[ci];
        }
        wr.resize(size_t(9) * l.cin);
        for (uint32_t co = 0; co < l.cout; ++co) {
            float m = 0.0f;
            for (uint32_t tap = 0; tap < 9u; ++tap)
                for (uint32_t ci = 0; ci < l.cin; ++ci) {
                    const float w = p[l.w_off + (size_t(tap) * l.cout + co) * l.cin + ci] * sa[ci];
                    wr[size_t(tap) * l.cin + ci] = w;
                    m = std::max(m, std::fabs(w));
                }
            const float s = m > 0.0f ? m / 448.0f : 1.0f;
            u.h_sw[l.s_off + co] = s;
            for (uint32_t tap = 0; tap < 9u; ++tap)
                for (uint32_t ci = 0; ci < l.cin; ++ci)
                    u.h_w8[l.wbf_off + (size_t(tap) * l.cout + co) * l.cin + ci] =
                        e4m3_encode(wr[size_t(tap) * l.cin + ci] / s);
        }
    }
    if ((rc = c->upload(u.w8, u.h_w8.data(), u.h_w8.size(), "unet fp8 weights")) ||
        (rc = c->upload(u.inv_sa, inv.data(), inv.size() * 4, "unet fp8 scales")) ||
        (rc = c->upload(u.sw, u.h_sw.data(), u.h_sw.size() * 4, "unet fp8 scales")))
        return rc;
    return c->sync("unet fp8");
}

}  // namespace
}  // namespace pt

extern "C" {

int pt_unet_init(pt_ctx* c, uint64_t seed) {
    if (!c) return PT_ERR_ARG;
    (void)hipSetDevice(c->device);
    if (!c->unet) {
        c->unet = new (std::nothrow) UNet();
        if (!c->unet) return c->fail(PT_ERR_OOM, "unet");
        build_layers(*c->unet);
    }
    UNet& u = *c->unet;
    std::vector<float> p;
    init_params(u, seed, p);
    int rc;
    if ((rc = c->upload(u.params, p.data(), p.size() * 4, "unet params")) ||
        (rc = c->alloc(u.grads, u.n_params * 4, "unet grads")) || (rc = c->alloc(u.m, u.n_params * 4, "unet adam")) ||
        (rc = c->alloc(u.v, u.n_params * 4, "unet adam")) || (rc = c->alloc(u.wbf, u.n_wbf * 2, "unet bf16 weights")) ||
        (rc = c->alloc(u.wdt, u.n_wbf * 2, "unet bf16 weights")) ||
        (rc = c->alloc(u.wup, u.n_wup * 2, "unet merged weights")))
        return rc;
    hipError_t e = fill_zero_f32(fp(u.m), u.n_params, c->stream);
    if (e == hipSuccess) e = fill_zero_f32(fp(u.v), u.n_params, c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "unet init");
    u.adam_t = 0;
    u.fp8 = false;
    if ((rc = refresh_weights(c, u))) return rc;
    return c->sync("unet init");
}

int pt_unet_param_count(pt_ctx* c, uint64_t* count) {
    if (!c || !count) return PT_ERR_ARG;
    int rc = ensure_net(c);
    if (rc) return rc;
    *count = c->unet->n_params;
    return PT_OK;
}

int pt_unet_get_params(pt_ctx* c, float* out, uint64_t count) {
    if (!c) return PT_ERR_ARG;
    int rc = ensure_net(c);
    if (rc) return rc;
    if (!out || count != c->unet->n_params) return c->fail(PT_ERR_ARG, "expected %llu parameters",
                                                            (unsigned long long)c->unet->n_params);
    if ((rc = c->sync("unet params"))) return rc;
    return hip_rc(c, hipMemcpy(out, c->unet->params.p, count * 4, hipMemcpyDeviceToHost), "unet params");
}

int pt_unet_set_params(pt_ctx* c, const float* in, uint64_t count) {
    if (!c) return PT_ERR_ARG;
    int rc = ensure_net(c);
    if (rc) return rc;
    UNet& u = *c->unet;
    if (!in || count != u.n_params) return c->fail(PT_ERR_ARG, "expected %llu parameters", (unsigned long long)u.n_params);
    if ((rc = c->upload(u.params, in, count * 4, "unet params"))) return rc;
    hipError_t e = fill_zero_f32(fp(u.m), u.n_params, c->stream);
    if (e == hipSuccess) e = fill_zero_f32(fp(u.v), u.n_params, c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "unet params");
    u.adam_t = 0;
    u.fp8 = false;
    if ((rc = refresh_weights(c, u))) return rc;
    return c->sync("unet params");
}

int pt_unet_get_adam(pt_ctx* c, float* m, float* v, uint64_t count, uint32_t* steps) {
    if (!c) return PT_ERR_ARG;
    int rc = ensure_net(c);
    if (rc) return rc;
    UNet& u = *c->unet;
    if (!m || !v || !steps || count != u.n_params)
        return c->fail(PT_ERR_ARG, "expected two arrays of %llu floats and a step count", (unsigned long long)u.n_params);
    if ((rc = c->sync("unet adam"))) return rc;
    if ((rc = hip_rc(c, hipMemcpy(m, u.m.p, count * 4, hipMemcpyDeviceToHost), "unet adam m"))) return rc;
    if ((rc = hip_rc(c, hipMemcpy(v, u.v.p, count * 4, hipMemcpyDeviceToHost), "unet adam v"))) return rc;
    *steps = u.adam_t;
    return PT_OK;
}

int pt_unet_set_adam(pt_ctx* c, const float* m, const float* v, uint64_t count, uint32_t steps) {
    if (!c) return PT_ERR_ARG;
    int rc = ensure_net(c);
    if (rc) return rc;
    UNet& u = *c->unet;
    if (!m || !v || count != u.n_params)
        return c->fail(PT_ERR_ARG, "expected two arrays of %llu floats", (unsigned long long)u.n_params);
    if ((rc = c->upload(u.m, m, count * 4, "unet adam m")) || (rc = c->upload(u.v, v, count * 4, "unet adam v")))
        return rc;
    u.adam_t = steps;
    return c->sync("unet adam");
}

int pt_unet_forward(pt_ctx* c, const float* feat, uint32_t n, uint32_t h, uint32_t w, float* pred) {
    if (!c) return PT_ERR_ARG;
    int rc = ensure_net(c);
    if (rc) return rc;
    if (!feat || !pred) return c->fail(PT_ERR_ARG, "NULL tensor");
    (void)hipSetDevice(c->device);
    UNet& u = *c->unet;
    if ((rc = alloc_acts(c, u, Dims{n, h, w})) || (rc = upload_batch_features(c, u, feat, nullptr)) ||
        (rc = forward(c, u, u.fp8, nullptr, true)) || (rc = c->sync("unet forward")))
        return rc;
    return hip_rc(c, hipMemcpy(pred, u.pred.p, size_t(u.dims[0].pixels()) * 3 * 4, hipMemcpyDeviceToHost),
                  "unet forward readback");
}

int pt_unet_loss_grad(pt_ctx* c, const float* feat, const float* target, uint32_t n, uint32_t h, uint32_t w,
                      float* loss, float* grad) {
    if (!c) return PT_ERR_ARG;
    int rc = ensure_net(c);
    if (rc) return rc;
    if (!feat || !target || !loss) return c->fail(PT_ERR_ARG, "NULL tensor");
    (void)hipSetDevice(c->device);
    UNet& u = *c->unet;
    if ((rc = alloc_acts(c, u, Dims{n, h, w})) || (rc = upload_batch_features(c, u, feat, target))) return rc;
    if ((rc = hip_rc(c, fill_zero_f32(fp(u.loss), 1, c->stream), "unet loss"))) return rc;
    if ((rc = forward(c, u)) || (rc = loss_and_grad(c, u, 0))) return rc;
    if (grad && (rc = backward(c, u))) return rc;
    if ((rc = c->sync("unet loss"))) return rc;
    float sum = 0.f;
    if ((rc = hip_rc(c, hipMemcpy(&sum, u.loss.p, 4, hipMemcpyDeviceToHost), "unet loss readback"))) return rc;
    *loss = sum / (float(u.dims[0].pixels()) * 3.0f);
    if (grad)
        return hip_rc(c, hipMemcpy(grad, u.grads.p, u.n_params * 4, hipMemcpyDeviceToHost), "unet grad readback");
    return PT_OK;
}

int pt_unet_set_dataset(pt_ctx* c, const float* feat, const float* target, uint32_t n, uint32_t h, uint32_t w) {
    if (!c) return PT_ERR_ARG;
    int rc = ensure_net(c);
    if (rc) return rc;
    if (!feat || !target || n == 0) return c->fail(PT_ERR_ARG, "empty dataset");
    if (h % 16u || w % 16u || h == 0 || w == 0) return c->fail(PT_ERR_ARG, "crops must be multiples of 16");
    (void)hipSetDevice(c->device);
    UNet& u = *c->unet;
    const size_t P = size_t(n) * h * w;
    DevBuf staging;
    if ((rc = c->alloc(staging, P * kFeat * 4, "unet dataset staging"))) return rc;
    hipError_t e = hipMemcpyAsync(staging.p, feat, P * kFeat * 4, hipMemcpyHostToDevice, c->stream);
    if ((rc = c->alloc(u.ds_in, P * kInC * 2, "unet dataset")) || (rc = c->alloc(u.ds_tg, P * 3 * 4, "unet dataset"))) {
        staging.release();
        return rc;
    }
    if (e == hipSuccess) e = pack_input(static_cast<const float*>(staging.p), Dims{n, h, w}, bfp(u.ds_in), c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(u.ds_tg.p, target, P * 3 * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    staging.release();
    if (e != hipSuccess) return c->hip_fail(e, "unet dataset");
    u.ds_n = n;
    u.ds_h = h;
    u.ds_w = w;
    return PT_OK;
}

int pt_unet_dataset_reserve(pt_ctx* c, uint32_t n, uint32_t h, uint32_t w) {
    if (!c) return PT_ERR_ARG;
    int rc = ensure_net(c);
    if (rc) return rc;
    if (n == 0 || h == 0 || w == 0 || h % 16u || w % 16u)
        return c->fail(PT_ERR_ARG, "crops must be positive multiples of 16");
    (void)hipSetDevice(c->device);
    UNet& u = *c->unet;
    const size_t P = size_t(n) * h * w;
    u.ds_n = 0;
    if ((rc = c->alloc(u.ds_in, P * kInC * 2, "unet dataset")) || (rc = c->alloc(u.ds_tg, P * 3 * 4, "unet dataset")))
        return rc;
    hipError_t e = fill_zero_bf16(bfp(u.ds_in), P * kInC, c->stream);
    if (e == hipSuccess) e = fill_zero_f32(fp(u.ds_tg), P * 3, c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "unet dataset");
    u.ds_n = n;
    u.ds_h = h;
    u.ds_w = w;
    return c->sync("unet dataset");
}

int pt_unet_dataset_capture(pt_ctx* c, uint32_t index, int what, uint32_t x0, uint32_t y0) {
    if (!c) return PT_ERR_ARG;
    int rc = ensure_net(c);
    if (rc) return rc;
    UNet& u = *c->unet;
    if (u.ds_n == 0) return c->fail(PT_ERR_STATE, "no dataset (pt_unet_dataset_reserve)");
    if (index >= u.ds_n) return c->fail(PT_ERR_ARG, "crop %u outside the dataset of %u", index, u.ds_n);
    if (what != PT_UNET_INPUT && what != PT_UNET_TARGET) return c->fail(PT_ERR_ARG, "unknown capture kind %d", what);
    if (c->width == 0 || !c->accum.p) return c->fail(PT_ERR_STATE, "no frame");
    if (what == PT_UNET_INPUT && !c->gbuf.p) return c->fail(PT_ERR_STATE, "input crops need a G-buffer");
    if (uint64_t(x0) + u.ds_w > c->width || uint64_t(y0) + u.ds_h > c->height)
        return c->fail(PT_ERR_ARG, "crop window outside the %ux%u frame", c->width, c->height);
    (void)hipSetDevice(c->device);
    const Dims d{1u, u.ds_h, u.ds_w};
    const size_t crop = size_t(u.ds_h) * u.ds_w;
    const float4* acc = static_cast<const float4*>(c->accum.p);
    const hipError_t e =
        what == PT_UNET_INPUT
            ? pack_frame(acc, c->gbuf.p, c->width, c->height, x0, y0, d, bfp(u.ds_in) + index * crop * kInC,
                         c->stream)
            : crop_radiance(acc, c->width, c->height, x0, y0, d, fp(u.ds_tg) + index * crop * 3, c->stream);
    if (e != hipSuccess) return c->hip_fail(e, "unet capture");
    return c->sync("unet capture");
}

int pt_unet_train(pt_ctx* c, uint32_t steps, uint32_t batch, float lr, uint64_t seed, float* losses) {
    if (!c) return PT_ERR_ARG;
    int rc = ensure_net(c);
    if (rc) return rc;
    UNet& u = *c->unet;
    if (u.ds_n == 0) return c->fail(PT_ERR_STATE, "no dataset (pt_unet_set_dataset)");
    if (batch == 0 || steps == 0) return c->fail(PT_ERR_ARG, "steps and batch must be positive");
    if (steps > 1024) return c->fail(PT_ERR_ARG, "at most 1024 steps per call");
    (void)hipSetDevice(c->device);
    if ((rc = alloc_acts(c, u, Dims{batch, u.ds_h, u.ds_w}))) return rc;
    u.fp8 = false;  // the weights change
    const size_t crop = size_t(u.ds_h) * u.ds_w;
    hipStream_t s = c->stream;
    if ((rc = hip_rc(c, fill_zero_f32(fp(u.loss), steps, s), "unet loss"))) return rc;
    uint64_t rs = seed ^ 0xD1B54A32D192ED03ull;
    for (uint32_t step = 0; step < steps; ++step) {
        for (uint32_t bi = 0; bi < batch; ++bi) {
            const uint32_t k = uint32_t(splitmix(rs) % u.ds_n);
            hipError_t e = hipMemcpyAsync(bfp(u.in0) + bi * crop * kInC, bfp(u.ds_in) + k * crop * kInC,
                                          crop * kInC * 2, hipMemcpyDeviceToDevice, s);
            if (e == hipSuccess)
                e = hipMemcpyAsync(fp(u.target) + bi * crop * 3, fp(u.ds_tg) + k * crop * 3, crop * 3 * 4,
                                   hipMemcpyDeviceToDevice, s);
            if (e != hipSuccess) return c->hip_fail(e, "unet batch");
        }
        if ((rc = forward(c, u)) || (rc = loss_and_grad(c, u, step)) || (rc = backward(c, u))) return rc;
        ++u.adam_t;
        if ((rc = hip_rc(c, adam(fp(u.params), fp(u.grads), fp(u.m), fp(u.v), u.n_params, lr, 0.9f, 0.999f, 1e-8f,
                                 u.adam_t, s),
                         "unet adam")))
            return rc;
        if ((rc = refresh_weights(c, u))) return rc;
    }
    if ((rc = c->sync("unet train"))) return rc;
    if (losses) {
        std::vector<float> l(steps);
        if ((rc = hip_rc(c, hipMemcpy(l.data(), u.loss.p, steps * 4, hipMemcpyDeviceToHost), "unet loss readback")))
            return rc;
        const float norm = 1.0f / (float(u.dims[0].pixels()) * 3.0f);
        for (uint32_t i = 0; i < steps; ++i) losses[i] = l[i] * norm;
    }
    return PT_OK;
}

int pt_unet_denoise(pt_ctx* c, float exposure) {
    if (!c) return PT_ERR_ARG;
    int rc = ensure_net(c);
    if (rc) return rc;
    if (c->width == 0 || !c->gbuf.p) return c->fail(PT_ERR_STATE, "denoise needs a frame with a G-buffer");
    (void)hipSetDevice(c->device);
    UNet& u = *c->unet;
    const uint32_t fw = c->width, fh = c->height;
    if ((rc = c->alloc(c->dn_fin, size_t(fw) * fh * 16, "denoise out"))) return rc;
    // the output convolution writes the denoised texels (fp32 and rgba8) itself
    const ConvFuse texels{Tens{nullptr, 0, 0}, Tens{nullptr, 0, 0}, Tens{nullptr, 0, 0}, 0u, static_cast<float4*>(c->dn_fin.p),
                          static_cast<uchar4*>(c->out.p), fw, fh, exposure};
    if ((rc = frame_forward(c, u, u.fp8, &texels))) return rc;
    return c->sync("unet denoise");
}

int pt_unet_set_precision(pt_ctx* c, int precision) {
    if (!c) return PT_ERR_ARG;
    int rc = ensure_net(c);
    if (rc) return rc;
    UNet& u = *c->unet;
    if (precision == PT_UNET_BF16) {
        u.fp8 = false;
        return PT_OK;
    }
    if (precision != PT_UNET_FP8) return c->fail(PT_ERR_ARG, "unknown precision %d", precision);
    if (c->width == 0 || !c->gbuf.p) return c->fail(PT_ERR_STATE, "fp8 calibration needs a frame with a G-buffer");
    (void)hipSetDevice(c->device);
    u.fp8 = false;
    if ((rc = c->alloc(u.amax, u.n_act * 4, "unet calibration"))) return rc;
    if ((rc = hip_rc(c, fill_zero_f32(fp(u.amax), u.n_act, c->stream), "unet calibration"))) return rc;
    u.calib = static_cast<uint32_t*>(u.amax.p);
    rc = frame_forward(c, u, false);
    u.calib = nullptr;
    if (rc) return rc;
    if ((rc = c->alloc(u.hist, size_t(u.n_act) * kHistBins * 4, "unet calibration"))) return rc;
    if ((rc = hip_rc(c, fill_zero_f32(fp(u.hist), u.n_act * kHistBins, c->stream), "unet calibration"))) return rc;
    u.calib_hist = static_cast<uint32_t*>(u.hist.p);
    rc = frame_forward(c, u, false);
    u.calib_hist = nullptr;
    if (rc || (rc = c->sync("unet calibration")) || (rc = quantize_fp8(c, u))) return rc;
    u.fp8 = true;
    return PT_OK;
}

int pt_unet_fp8_state(pt_ctx* c, float* act_scales, uint64_t n_act, float* w_scales, uint64_t n_w, uint8_t* weights,
                      uint64_t n_bytes) {
    if (!c) return PT_ERR_ARG;
    int rc = ensure_net(c);
    if (rc) return rc;
    const UNet& u = *c->unet;
    if (!u.fp8) return c->fail(PT_ERR_STATE, "the network is not in fp8 mode (pt_unet_set_precision)");
    if (!act_scales || !w_scales || !weights || n_act != u.n_act || n_w != u.n_out || n_bytes != u.n_wbf)
        return c->fail(PT_ERR_ARG, "expected %llu input scales, %llu output scales and %llu weight bytes",
                       (unsigned long long)u.n_act, (unsigned long long)u.n_out, (unsigned long long)u.n_wbf);
    std::memcpy(act_scales, u.h_sa.data(), n_act * 4);
    std::memcpy(w_scales, u.h_sw.data(), n_w * 4);
    std::memcpy(weights, u.h_w8.data(), n_bytes);
    return PT_OK;
}

}  // extern "C"
