// pt_addon.cc — Node-API addon over the libpt C ABI.
//
// This is the binding the reference's JavaScript host would load in place of
// its WebGPU dispatch (tracer_native.js wraps it in the Renderer / Denoiser
// method names of tracer.js).  Every function takes the context External as
// its first argument; buffers are TypedArrays / ArrayBuffers in the tracer.js
// layouts; a nonzero pt_* status becomes a thrown JS Error carrying
// pt_last_error().  Calls are synchronous (the C ABI returns after the
// device work is done), which is what `await onSubmittedWorkDone()` gives the
// reference.  The long calls (trace, denoise, unetDenoise, batchFrame) also
// come as *Async variants returning a Promise (napi_async_work), for hosts that
// must keep their event loop running while the GPU works.
#include <node_api.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "pt_abi.h"
#include "pt_batch.h"
#include "pt_unet.h"

namespace {

#define CHECK_NAPI(call)                                                                 \
    do {                                                                                 \
        if ((call) != napi_ok) {                                                         \
            napi_throw_error(env, nullptr, "N-API call failed: " #call);                 \
            return nullptr;                                                              \
        }                                                                                \
    } while (0)

napi_value throw_pt(napi_env env, pt_ctx* ctx, int rc, const char* what) {
    std::string msg = std::string(what) + " failed (" + std::to_string(rc) + "): " + pt_last_error(ctx);
    napi_throw_error(env, nullptr, msg.c_str());
    return nullptr;
}

bool get_args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
    size_t argc = want;
    if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok || argc < want) {
        napi_throw_type_error(env, nullptr, "missing arguments");
        return false;
    }
    return true;
}

// Handles (contexts, batches) with an asynchronous call in flight.  A context
// is not re-entrant, so any other call on a busy handle throws.  Touched only
// on the JavaScript thread.
std::unordered_set<const void*> g_busy;

bool not_busy(napi_env env, const void* h) {
    if (!g_busy.count(h)) return true;
    napi_throw_error(env, nullptr, "busy: an asynchronous call on this context is still in flight");
    return false;
}

pt_ctx* get_ctx(napi_env env, napi_value v) {
    void* p = nullptr;
    if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
        napi_throw_type_error(env, nullptr, "first argument must be a pt context");
        return nullptr;
    }
    if (!not_busy(env, p)) return nullptr;
    return static_cast<pt_ctx*>(p);
}

// Raw bytes of a TypedArray, DataView-free ArrayBuffer or Buffer.
bool get_bytes(napi_env env, napi_value v, void** data, size_t* len) {
    bool is;
    if (napi_is_typedarray(env, v, &is) == napi_ok && is) {
        napi_typedarray_type t;
        size_t n, off;
        napi_value ab;
        void* base;
        if (napi_get_typedarray_info(env, v, &t, &n, &base, &ab, &off) != napi_ok) return false;
        size_t el = 1;
        switch (t) {
            case napi_int16_array: case napi_uint16_array: el = 2; break;
            case napi_int32_array: case napi_uint32_array: case napi_float32_array: el = 4; break;
            case napi_float64_array: case napi_bigint64_array: case napi_biguint64_array: el = 8; break;
            default: el = 1;
        }
        *data = base;
        *len = n * el;
        return true;
    }
    if (napi_is_arraybuffer(env, v, &is) == napi_ok && is) return napi_get_arraybuffer_info(env, v, data, len) == napi_ok;
    return false;
}

bool prop(napi_env env, napi_value obj, const char* name, napi_value* out) {
    bool has = false;
    if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) return false;
    if (napi_get_named_property(env, obj, name, out) != napi_ok) return false;
    napi_valuetype t;
    napi_typeof(env, *out, &t);
    return t != napi_undefined && t != napi_null;
}

uint32_t prop_u32(napi_env env, napi_value obj, const char* name, uint32_t dflt) {
    napi_value v;
    uint32_t r = dflt;
    if (prop(env, obj, name, &v)) napi_get_value_uint32(env, v, &r);
    return r;
}
double prop_f64(napi_env env, napi_value obj, const char* name, double dflt) {
    napi_value v;
    double r = dflt;
    if (prop(env, obj, name, &v)) napi_get_value_double(env, v, &r);
    return r;
}

bool array_field(napi_env env, napi_value obj, const char* name, uint32_t stride, pt_array* a) {
    napi_value v;
    std::memset(a, 0, sizeof(*a));
    a->stride = stride;
    if (!prop(env, obj, name, &v)) return true;  // absent: empty (the C ABI reports it)
    void* p;
    size_t n;
    if (!get_bytes(env, v, &p, &n)) return false;
    a->ptr = p;
    a->count = n / stride;
    return n % stride == 0;
}

void cleanup(napi_env, void* data, void*) { pt_destroy(static_cast<pt_ctx*>(data)); }

napi_value create(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    size_t argc = 1;
    CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
    int32_t device = 0;
    if (argc >= 1) napi_get_value_int32(env, argv[0], &device);
    pt_ctx* ctx = nullptr;
    int rc = pt_create(&ctx, device);
    if (rc) {
        napi_throw_error(env, nullptr, ("pt_create failed (" + std::to_string(rc) + "): no usable HIP device").c_str());
        return nullptr;
    }
    napi_value ext;
    CHECK_NAPI(napi_create_external(env, ctx, cleanup, nullptr, &ext));
    return ext;
}

// desc: {materials, objects, tlas, meshes, bvh, triangles, planes, lights,
//        sky (Uint16Array rgba16f), skyWidth, skyHeight, skyCDF (Float32Array|null),
//        skyExposure, skyTotalLum, flags, bounceLimit,
//        textures (Uint8Array of 512x512 rgba8 layers, optional)}
bool read_desc(napi_env env, napi_value o, pt_scene_desc* d) {
    std::memset(d, 0, sizeof(*d));
    if (!array_field(env, o, "materials", sizeof(pt_material), &d->materials) ||
        !array_field(env, o, "objects", sizeof(pt_object), &d->objects) ||
        !array_field(env, o, "tlas", sizeof(pt_bvh_node), &d->tlas) ||
        !array_field(env, o, "meshes", sizeof(pt_mesh_instance), &d->meshes) ||
        !array_field(env, o, "bvh", sizeof(pt_bvh_node), &d->bvh) ||
        !array_field(env, o, "triangles", sizeof(pt_triangle), &d->triangles) ||
        !array_field(env, o, "planes", sizeof(pt_plane), &d->planes) ||
        !array_field(env, o, "lights", sizeof(pt_light), &d->lights)) {
        napi_throw_type_error(env, nullptr, "scene arrays must be TypedArrays whose byte length is a multiple of the record size");
        return false;
    }
    napi_value v;
    void* p;
    size_t n;
    d->sky_width = prop_u32(env, o, "skyWidth", 1);
    d->sky_height = prop_u32(env, o, "skyHeight", 1);
    if (prop(env, o, "sky", &v) && get_bytes(env, v, &p, &n)) {
        if (n != size_t(d->sky_width) * d->sky_height * 8) {
            napi_throw_type_error(env, nullptr, "sky must hold skyWidth*skyHeight rgba16f texels");
            return false;
        }
        d->sky_rgba16f = static_cast<const uint16_t*>(p);
    }
    if (prop(env, o, "skyCDF", &v) && get_bytes(env, v, &p, &n)) {
        d->sky_cdf = static_cast<const float*>(p);
        d->sky_cdf_count = n / 4;
    }
    d->sky_exposure = float(prop_f64(env, o, "skyExposure", 1.0));
    d->sky_total_lum = float(prop_f64(env, o, "skyTotalLum", 0.0));
    d->flags = prop_u32(env, o, "flags", 0);
    d->bounce_limit = int32_t(prop_u32(env, o, "bounceLimit", 8));
    if (prop(env, o, "textures", &v) && get_bytes(env, v, &p, &n) && n > 0) {
        const size_t layer = size_t(512) * 512 * 4;
        if (n % layer) {
            napi_throw_type_error(env, nullptr, "textures must hold whole 512x512 rgba8 layers");
            return false;
        }
        d->textures = static_cast<const uint8_t*>(p);
        d->texture_layers = uint32_t(n / layer);
    }
    return true;
}

napi_value upload_common(napi_env env, napi_callback_info info, bool update) {
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return nullptr;
    pt_ctx* ctx = get_ctx(env, argv[0]);
    if (!ctx) return nullptr;
    pt_scene_desc d;
    if (!read_desc(env, argv[1], &d)) return nullptr;
    int rc = update ? pt_update_objects(ctx, &d) : pt_upload_scene(ctx, &d);
    if (update && (rc == PT_ERR_CAPACITY || rc == PT_ERR_FLAGS)) {
        napi_value r;
        napi_create_int32(env, rc, &r);
        return r;  // the caller falls back to uploadScene, as updateObjects falls back to setScene
    }
    if (rc) return throw_pt(env, ctx, rc, update ? "pt_update_objects" : "pt_upload_scene");
    napi_value r;
    napi_create_int32(env, 0, &r);
    return r;
}
napi_value upload_scene(napi_env env, napi_callback_info info) { return upload_common(env, info, false); }
napi_value update_objects(napi_env env, napi_callback_info info) { return upload_common(env, info, true); }

napi_value resize(napi_env env, napi_callback_info info) {
    napi_value argv[4];
    if (!get_args(env, info, 4, argv)) return nullptr;
    pt_ctx* ctx = get_ctx(env, argv[0]);
    if (!ctx) return nullptr;
    uint32_t w = 0, h = 0;
    bool g = false;
    napi_get_value_uint32(env, argv[1], &w);
    napi_get_value_uint32(env, argv[2], &h);
    napi_get_value_bool(env, argv[3], &g);
    int rc = pt_resize(ctx, w, h, g ? 1 : 0);
    if (rc) return throw_pt(env, ctx, rc, "pt_resize");
    return nullptr;
}

bool get_params(napi_env env, napi_value v, pt_scene_params* p) {
    void* data;
    size_t n;
    if (!get_bytes(env, v, &data, &n) || n != sizeof(pt_scene_params)) {
        napi_throw_type_error(env, nullptr, "params must be the 112-byte SceneParams uniform");
        return false;
    }
    std::memcpy(p, data, sizeof(*p));
    return true;
}

napi_value set_num(napi_env env, napi_value obj, const char* k, double v) {
    napi_value x;
    napi_create_double(env, v, &x);
    napi_set_named_property(env, obj, k, x);
    return obj;
}

napi_value stats_object(napi_env env, const pt_stats& st) {
    napi_value o;
    napi_create_object(env, &o);
    set_num(env, o, "extensionRays", double(st.extension_rays));
    set_num(env, o, "shadowRays", double(st.shadow_rays));
    set_num(env, o, "paths", double(st.paths));
    set_num(env, o, "kernelMs", st.kernel_ms);
    return o;
}

// The arguments of one trace call, copied off the JavaScript values.
struct TraceCall {
    pt_ctx* ctx = nullptr;
    pt_scene_params p{};
    uint32_t sw = 0, sh = 0, n = 1;
    std::vector<uint32_t> seeds;
    bool has_seeds = false;
};

// trace(ctx, params, sectionW, sectionH, nPasses, seeds|null)
bool parse_trace(napi_env env, napi_callback_info info, TraceCall* t) {
    napi_value argv[6];
    if (!get_args(env, info, 6, argv)) return false;
    if (!(t->ctx = get_ctx(env, argv[0]))) return false;
    if (!get_params(env, argv[1], &t->p)) return false;
    napi_get_value_uint32(env, argv[2], &t->sw);
    napi_get_value_uint32(env, argv[3], &t->sh);
    napi_get_value_uint32(env, argv[4], &t->n);
    void* sp;
    size_t sn;
    if (get_bytes(env, argv[5], &sp, &sn)) {
        if (sn < size_t(t->n) * 4) {
            napi_throw_type_error(env, nullptr, "seeds must hold one u32 per pass");
            return false;
        }
        const uint32_t* s = static_cast<const uint32_t*>(sp);
        t->seeds.assign(s, s + t->n);
        t->has_seeds = true;
    }
    return true;
}

// trace(ctx, params, sectionW, sectionH, nPasses, seeds|null) -> stats
napi_value trace(napi_env env, napi_callback_info info) {
    TraceCall t;
    if (!parse_trace(env, info, &t)) return nullptr;
    pt_stats st;
    int rc = pt_trace(t.ctx, &t.p, t.sw, t.sh, t.n, t.has_seeds ? t.seeds.data() : nullptr, &st);
    if (rc) return throw_pt(env, t.ctx, rc, "pt_trace");
    return stats_object(env, st);
}

napi_value gbuffer(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return nullptr;
    pt_ctx* ctx = get_ctx(env, argv[0]);
    if (!ctx) return nullptr;
    pt_scene_params p;
    if (!get_params(env, argv[1], &p)) return nullptr;
    int rc = pt_gbuffer(ctx, &p);
    if (rc) return throw_pt(env, ctx, rc, "pt_gbuffer");
    return nullptr;
}

// denoise(ctx, uniforms48, strideMode)
bool parse_denoise(napi_env env, napi_callback_info info, pt_ctx** ctx, pt_denoise_params* u, int32_t* mode) {
    napi_value argv[3];
    if (!get_args(env, info, 3, argv)) return false;
    if (!(*ctx = get_ctx(env, argv[0]))) return false;
    void* data;
    size_t n;
    if (!get_bytes(env, argv[1], &data, &n) || n != sizeof(pt_denoise_params)) {
        napi_throw_type_error(env, nullptr, "uniforms must be the 48-byte denoiser uniform");
        return false;
    }
    std::memcpy(u, data, sizeof(*u));
    *mode = 0;
    napi_get_value_int32(env, argv[2], mode);
    return true;
}

napi_value denoise(napi_env env, napi_callback_info info) {
    pt_ctx* ctx;
    pt_denoise_params u;
    int32_t mode;
    if (!parse_denoise(env, info, &ctx, &u, &mode)) return nullptr;
    int rc = pt_denoise(ctx, &u, mode);
    if (rc) return throw_pt(env, ctx, rc, "pt_denoise");
    return nullptr;
}

// read*(ctx, destTypedArray): fills a caller-owned array
template <int (*FN)(pt_ctx*, void*, uint64_t)>
napi_value read_into(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return nullptr;
    pt_ctx* ctx = get_ctx(env, argv[0]);
    if (!ctx) return nullptr;
    void* data;
    size_t n;
    if (!get_bytes(env, argv[1], &data, &n)) {
        napi_throw_type_error(env, nullptr, "destination must be a TypedArray");
        return nullptr;
    }
    int rc = FN(ctx, data, n);
    if (rc) return throw_pt(env, ctx, rc, "readback");
    return argv[1];
}
int rd_accum(pt_ctx* c, void* d, uint64_t n) { return pt_read_accum(c, static_cast<float*>(d), n); }
int rd_denoised(pt_ctx* c, void* d, uint64_t n) { return pt_read_denoised(c, static_cast<float*>(d), n); }
int rd_output(pt_ctx* c, void* d, uint64_t n) { return pt_read_output(c, static_cast<uint8_t*>(d), n); }
int rd_gbuf(pt_ctx* c, void* d, uint64_t n) { return pt_read_gbuffer(c, static_cast<pt_gbuffer_texel*>(d), n); }
int wr_accum(pt_ctx* c, void* d, uint64_t n) { return pt_write_accum(c, static_cast<const float*>(d), n); }

// readSkyCdf(ctx, Float32Array of (h + 1) + h (w + 1) floats) -> total weight:
// the CDF the device built when the scene came without one
napi_value read_sky_cdf(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return nullptr;
    pt_ctx* ctx = get_ctx(env, argv[0]);
    if (!ctx) return nullptr;
    void* data;
    size_t n;
    if (!get_bytes(env, argv[1], &data, &n) || n % 4) {
        napi_throw_type_error(env, nullptr, "destination must be a Float32Array");
        return nullptr;
    }
    float total = 0.f;
    int rc = pt_read_sky_cdf(ctx, static_cast<float*>(data), n / 4, &total);
    if (rc) return throw_pt(env, ctx, rc, "pt_read_sky_cdf");
    napi_value r;
    napi_create_double(env, double(total), &r);
    return r;
}

// setOption(ctx, key, value): pt_set_option (the PT_OPT_* keys)
napi_value set_option(napi_env env, napi_callback_info info) {
    napi_value argv[3];
    if (!get_args(env, info, 3, argv)) return nullptr;
    pt_ctx* ctx = get_ctx(env, argv[0]);
    if (!ctx) return nullptr;
    int32_t key = 0;
    int64_t value = 0;
    napi_get_value_int32(env, argv[1], &key);
    napi_get_value_int64(env, argv[2], &value);
    int rc = pt_set_option(ctx, key, value);
    if (rc) return throw_pt(env, ctx, rc, "pt_set_option");
    return nullptr;
}

// The learned denoiser (pt_unet.h): unetInit(ctx, seed), unetSetParams(ctx,
// Float32Array), unetDenoise(ctx, exposure), unetSetPrecision(ctx, 0 bf16 |
// 1 fp8, calibrated on the current frame); the result is read with
// readOutput / readDenoised like the a-trous filter's.
napi_value unet_init(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return nullptr;
    pt_ctx* ctx = get_ctx(env, argv[0]);
    if (!ctx) return nullptr;
    int64_t seed = 0;
    napi_get_value_int64(env, argv[1], &seed);
    int rc = pt_unet_init(ctx, uint64_t(seed));
    if (rc) return throw_pt(env, ctx, rc, "pt_unet_init");
    return nullptr;
}

napi_value unet_set_params(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return nullptr;
    pt_ctx* ctx = get_ctx(env, argv[0]);
    if (!ctx) return nullptr;
    void* data;
    size_t n;
    if (!get_bytes(env, argv[1], &data, &n) || n % 4) {
        napi_throw_type_error(env, nullptr, "parameters must be a Float32Array");
        return nullptr;
    }
    int rc = pt_unet_set_params(ctx, static_cast<const float*>(data), n / 4);
    if (rc) return throw_pt(env, ctx, rc, "pt_unet_set_params");
    return nullptr;
}

// unetDenoise(ctx, exposure)
bool parse_unet(napi_env env, napi_callback_info info, pt_ctx** ctx, double* exposure) {
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return false;
    if (!(*ctx = get_ctx(env, argv[0]))) return false;
    *exposure = 1.0;
    napi_get_value_double(env, argv[1], exposure);
    return true;
}

napi_value unet_denoise(napi_env env, napi_callback_info info) {
    pt_ctx* ctx;
    double exposure;
    if (!parse_unet(env, info, &ctx, &exposure)) return nullptr;
    int rc = pt_unet_denoise(ctx, float(exposure));
    if (rc) return throw_pt(env, ctx, rc, "pt_unet_denoise");
    return nullptr;
}

napi_value unet_set_precision(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return nullptr;
    pt_ctx* ctx = get_ctx(env, argv[0]);
    if (!ctx) return nullptr;
    int32_t precision = 0;
    napi_get_value_int32(env, argv[1], &precision);
    int rc = pt_unet_set_precision(ctx, precision);
    if (rc) return throw_pt(env, ctx, rc, "pt_unet_set_precision");
    return nullptr;
}

// Batches (pt_batch.h): batchCreate(device) -> External, batchAdd(batch, ctx),
// batchFrame(batch, [params ArrayBuffer...], [passes...], [seeds Uint32Array|null...])
// -> [{extensionRays, shadowRays, paths, kernelMs}...]
void batch_cleanup(napi_env, void* data, void*) { pt_batch_destroy(static_cast<pt_batch*>(data)); }

// The contexts of each batch (their jobs), so an asynchronous batch frame can
// mark them busy too.
std::unordered_map<const void*, std::vector<const void*>> g_batch_jobs;

pt_batch* get_batch(napi_env env, napi_value v) {
    void* p = nullptr;
    if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
        napi_throw_type_error(env, nullptr, "first argument must be a pt batch");
        return nullptr;
    }
    if (!not_busy(env, p)) return nullptr;
    return static_cast<pt_batch*>(p);
}

napi_value throw_batch(napi_env env, pt_batch* b, int rc, const char* what) {
    std::string msg = std::string(what) + " failed (" + std::to_string(rc) + "): " + pt_batch_last_error(b);
    napi_throw_error(env, nullptr, msg.c_str());
    return nullptr;
}

napi_value batch_create(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    size_t argc = 1;
    CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
    int32_t device = 0;
    if (argc >= 1) napi_get_value_int32(env, argv[0], &device);
    pt_batch* b = nullptr;
    int rc = pt_batch_create(&b, device);
    if (rc) {
        napi_throw_error(env, nullptr, ("pt_batch_create failed (" + std::to_string(rc) + ")").c_str());
        return nullptr;
    }
    napi_value ext;
    CHECK_NAPI(napi_create_external(env, b, batch_cleanup, nullptr, &ext));
    return ext;
}

napi_value batch_add(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return nullptr;
    pt_batch* b = get_batch(env, argv[0]);
    if (!b) return nullptr;
    pt_ctx* ctx = get_ctx(env, argv[1]);
    if (!ctx) return nullptr;
    uint32_t idx = 0;
    int rc = pt_batch_add(b, ctx, &idx);
    if (rc) return throw_batch(env, b, rc, "pt_batch_add");
    g_batch_jobs[b].push_back(ctx);
    napi_value r;
    napi_create_uint32(env, idx, &r);
    return r;
}

// The arguments of one batch frame, copied off the JavaScript values.
struct BatchCall {
    pt_batch* b = nullptr;
    std::vector<pt_scene_params> params;
    std::vector<uint32_t> passes;
    std::vector<std::vector<uint32_t>> seeds;
    std::vector<char> has_seeds;
};

// batchFrame(batch, [params ArrayBuffer...], [passes...], [seeds Uint32Array|null...])
bool parse_batch(napi_env env, napi_callback_info info, BatchCall* c) {
    napi_value argv[4];
    if (!get_args(env, info, 4, argv)) return false;
    if (!(c->b = get_batch(env, argv[0]))) return false;
    uint32_t n = 0;
    pt_batch_size(c->b, &n);
    for (const void* job : g_batch_jobs[c->b])
        if (!not_busy(env, job)) return false;
    c->params.resize(n);
    c->passes.assign(n, 0);
    c->seeds.resize(n);
    c->has_seeds.assign(n, 0);
    for (uint32_t i = 0; i < n; ++i) {
        napi_value pv, nv, sv;
        void* data;
        size_t len;
        if (napi_get_element(env, argv[1], i, &pv) != napi_ok || !get_bytes(env, pv, &data, &len) ||
            len != sizeof(pt_scene_params)) {
            napi_throw_type_error(env, nullptr, "params must hold one 112-byte SceneParams per job");
            return false;
        }
        std::memcpy(&c->params[i], data, sizeof(pt_scene_params));
        if (napi_get_element(env, argv[2], i, &nv) != napi_ok ||
            napi_get_value_uint32(env, nv, &c->passes[i]) != napi_ok) {
            napi_throw_type_error(env, nullptr, "passes must hold one count per job");
            return false;
        }
        if (napi_get_element(env, argv[3], i, &sv) == napi_ok && get_bytes(env, sv, &data, &len)) {
            if (len < size_t(c->passes[i]) * 4) {
                napi_throw_type_error(env, nullptr, "seeds must hold one u32 per pass");
                return false;
            }
            const uint32_t* s = static_cast<const uint32_t*>(data);
            c->seeds[i].assign(s, s + c->passes[i]);
            c->has_seeds[i] = 1;
        }
    }
    return true;
}

int run_batch(BatchCall& c, std::vector<pt_stats>& st) {
    std::vector<const uint32_t*> sp(c.passes.size(), nullptr);
    for (size_t i = 0; i < sp.size(); ++i)
        if (c.has_seeds[i]) sp[i] = c.seeds[i].data();
    st.assign(c.passes.size(), pt_stats{});
    return pt_batch_frame(c.b, c.params.data(), c.passes.data(), sp.data(), st.data());
}

napi_value stats_array(napi_env env, const std::vector<pt_stats>& st) {
    napi_value arr;
    napi_create_array_with_length(env, st.size(), &arr);
    for (uint32_t i = 0; i < st.size(); ++i) napi_set_element(env, arr, i, stats_object(env, st[i]));
    return arr;
}

napi_value batch_frame(napi_env env, napi_callback_info info) {
    BatchCall c;
    if (!parse_batch(env, info, &c)) return nullptr;
    std::vector<pt_stats> st;
    int rc = run_batch(c, st);
    if (rc) return throw_batch(env, c.b, rc, "pt_batch_frame");
    return stats_array(env, st);
}

// ---- asynchronous variants --------------------------------------------------
// traceAsync / denoiseAsync / unetDenoiseAsync / batchFrameAsync take the same
// arguments as the synchronous calls and return a Promise, as the reference's
// host awaits the GPU (await device.queue.onSubmittedWorkDone(),
// tracer.js:2280-2281, 2294).  The arguments are copied on the JavaScript
// thread, the C call runs on a libuv worker thread, and the Promise settles on
// the JavaScript thread with the synchronous call's result, or rejects with
// its Error.  The handles involved are marked busy (and kept alive by
// references) until then.
enum class Op { Trace, Denoise, Unet, Batch };

struct AsyncCall {
    Op op;
    napi_async_work work = nullptr;
    napi_deferred deferred = nullptr;
    std::vector<napi_ref> refs;           // the handle Externals, alive while in flight
    std::vector<const void*> claimed;     // busy handles
    TraceCall t;
    pt_ctx* ctx = nullptr;
    pt_denoise_params u{};
    int32_t mode = 0;
    double exposure = 1.0;
    BatchCall bc;
    int rc = 0;
    std::string err;
    pt_stats st{};
    std::vector<pt_stats> bst;
};

void async_execute(napi_env, void* data) {
    AsyncCall* w = static_cast<AsyncCall*>(data);
    const char* what = "";
    switch (w->op) {
        case Op::Trace:
            what = "pt_trace";
            w->rc = pt_trace(w->t.ctx, &w->t.p, w->t.sw, w->t.sh, w->t.n, w->t.has_seeds ? w->t.seeds.data() : nullptr,
                             &w->st);
            break;
        case Op::Denoise:
            what = "pt_denoise";
            w->rc = pt_denoise(w->ctx, &w->u, w->mode);
            break;
        case Op::Unet:
            what = "pt_unet_denoise";
            w->rc = pt_unet_denoise(w->ctx, float(w->exposure));
            break;
        case Op::Batch:
            what = "pt_batch_frame";
            w->rc = run_batch(w->bc, w->bst);
            break;
    }
    if (w->rc) {
        const char* msg = w->op == Op::Batch ? pt_batch_last_error(w->bc.b)
                                             : pt_last_error(w->op == Op::Trace ? w->t.ctx : w->ctx);
        w->err = std::string(what) + " failed (" + std::to_string(w->rc) + "): " + msg;
    }
}

void async_complete(napi_env env, napi_status status, void* data) {
    AsyncCall* w = static_cast<AsyncCall*>(data);
    for (const void* h : w->claimed) g_busy.erase(h);
    if (status != napi_ok || w->rc) {
        napi_value m, e;
        napi_create_string_utf8(env, status != napi_ok ? "asynchronous call cancelled" : w->err.c_str(),
                                NAPI_AUTO_LENGTH, &m);
        napi_create_error(env, nullptr, m, &e);
        napi_reject_deferred(env, w->deferred, e);
    } else {
        napi_value r;
        if (w->op == Op::Trace) r = stats_object(env, w->st);
        else if (w->op == Op::Batch) r = stats_array(env, w->bst);
        else napi_get_undefined(env, &r);
        napi_resolve_deferred(env, w->deferred, r);
    }
    for (napi_ref ref : w->refs) napi_delete_reference(env, ref);
    napi_delete_async_work(env, w->work);
    delete w;
}

napi_value queue_async(napi_env env, napi_callback_info info, AsyncCall* w, const char* name) {
    // keep the handle Externals (the first argument) alive while in flight
    napi_value argv[1];
    size_t argc = 1;
    napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr);
    napi_ref ref;
    if (napi_create_reference(env, argv[0], 1, &ref) == napi_ok) w->refs.push_back(ref);
    napi_value promise, res_name;
    if (napi_create_promise(env, &w->deferred, &promise) != napi_ok ||
        napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &res_name) != napi_ok ||
        napi_create_async_work(env, nullptr, res_name, async_execute, async_complete, w, &w->work) != napi_ok ||
        napi_queue_async_work(env, w->work) != napi_ok) {
        for (napi_ref r : w->refs) napi_delete_reference(env, r);
        delete w;
        napi_throw_error(env, nullptr, "could not queue the asynchronous call");
        return nullptr;
    }
    for (const void* h : w->claimed) g_busy.insert(h);
    return promise;
}

napi_value trace_async(napi_env env, napi_callback_info info) {
    auto* w = new AsyncCall{Op::Trace};
    if (!parse_trace(env, info, &w->t)) { delete w; return nullptr; }
    w->claimed = {w->t.ctx};
    return queue_async(env, info, w, "pt_trace");
}

napi_value denoise_async(napi_env env, napi_callback_info info) {
    auto* w = new AsyncCall{Op::Denoise};
    if (!parse_denoise(env, info, &w->ctx, &w->u, &w->mode)) { delete w; return nullptr; }
    w->claimed = {w->ctx};
    return queue_async(env, info, w, "pt_denoise");
}

napi_value unet_denoise_async(napi_env env, napi_callback_info info) {
    auto* w = new AsyncCall{Op::Unet};
    if (!parse_unet(env, info, &w->ctx, &w->exposure)) { delete w; return nullptr; }
    w->claimed = {w->ctx};
    return queue_async(env, info, w, "pt_unet_denoise");
}

napi_value batch_frame_async(napi_env env, napi_callback_info info) {
    auto* w = new AsyncCall{Op::Batch};
    if (!parse_batch(env, info, &w->bc)) { delete w; return nullptr; }
    w->claimed = g_batch_jobs[w->bc.b];
    w->claimed.push_back(w->bc.b);
    return queue_async(env, info, w, "pt_batch_frame");
}

napi_value last_error(napi_env env, napi_callback_info info) {
    napi_value argv[1];
    if (!get_args(env, info, 1, argv)) return nullptr;
    pt_ctx* ctx = get_ctx(env, argv[0]);
    if (!ctx) return nullptr;
    napi_value s;
    napi_create_string_utf8(env, pt_last_error(ctx), NAPI_AUTO_LENGTH, &s);
    return s;
}

napi_value init(napi_env env, napi_value exports) {
    napi_property_descriptor props[] = {
        {"create", nullptr, create, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"uploadScene", nullptr, upload_scene, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"updateObjects", nullptr, update_objects, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"resize", nullptr, resize, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"trace", nullptr, trace, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"gbuffer", nullptr, gbuffer, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"denoise", nullptr, denoise, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"readAccum", nullptr, read_into<rd_accum>, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"writeAccum", nullptr, read_into<wr_accum>, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"readOutput", nullptr, read_into<rd_output>, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"readGBuffer", nullptr, read_into<rd_gbuf>, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"readDenoised", nullptr, read_into<rd_denoised>, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"lastError", nullptr, last_error, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"setOption", nullptr, set_option, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"readSkyCdf", nullptr, read_sky_cdf, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"unetInit", nullptr, unet_init, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"unetSetParams", nullptr, unet_set_params, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"unetDenoise", nullptr, unet_denoise, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"unetSetPrecision", nullptr, unet_set_precision, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"batchCreate", nullptr, batch_create, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"batchAdd", nullptr, batch_add, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"batchFrame", nullptr, batch_frame, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"traceAsync", nullptr, trace_async, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"denoiseAsync", nullptr, denoise_async, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"unetDenoiseAsync", nullptr, unet_denoise_async, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
        {"batchFrameAsync", nullptr, batch_frame_async, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
    };
    napi_define_properties(env, exports, sizeof(props) / sizeof(props[0]), props);
    return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
