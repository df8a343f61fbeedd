// pt_addon.cc — Node-API addon over the libpt C ABI.
//
// This is the binding the reference's JavaScript host would load in place of
// its WebGPU dispatch (tracer_native.js wraps it in the Renderer / Denoiser
// method names of tracer.js).  Every function takes the context External as
// its first argument; buffers are TypedArrays / ArrayBuffers in the tracer.js
// layouts; a nonzero pt_* status becomes a thrown JS Error carrying
// pt_last_error().  Calls are synchronous (the C ABI returns after the
// device work is done), which is what `await onSubmittedWorkDone()` gives the
// reference.
#include <node_api.h>

#include <cstdint>
#include <cstring>
#include <string>
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

pt_ctx* get_ctx(napi_env env, napi_value v) {
    void* p = nullptr;
    if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
        napi_throw_type_error(env, nullptr, "first argument must be a pt context");
        return nullptr;
    }
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

// trace(ctx, params, sectionW, sectionH, nPasses, seeds|null) -> stats
napi_value trace(napi_env env, napi_callback_info info) {
    napi_value argv[6];
    if (!get_args(env, info, 6, argv)) return nullptr;
    pt_ctx* ctx = get_ctx(env, argv[0]);
    if (!ctx) return nullptr;
    pt_scene_params p;
    if (!get_params(env, argv[1], &p)) return nullptr;
    uint32_t sw = 0, sh = 0, n = 1;
    napi_get_value_uint32(env, argv[2], &sw);
    napi_get_value_uint32(env, argv[3], &sh);
    napi_get_value_uint32(env, argv[4], &n);
    const uint32_t* seeds = nullptr;
    void* sp;
    size_t sn;
    if (get_bytes(env, argv[5], &sp, &sn)) {
        if (sn < size_t(n) * 4) {
            napi_throw_type_error(env, nullptr, "seeds must hold one u32 per pass");
            return nullptr;
        }
        seeds = static_cast<const uint32_t*>(sp);
    }
    pt_stats st;
    int rc = pt_trace(ctx, &p, sw, sh, n, seeds, &st);
    if (rc) return throw_pt(env, ctx, rc, "pt_trace");
    napi_value o;
    napi_create_object(env, &o);
    set_num(env, o, "extensionRays", double(st.extension_rays));
    set_num(env, o, "shadowRays", double(st.shadow_rays));
    set_num(env, o, "paths", double(st.paths));
    set_num(env, o, "kernelMs", st.kernel_ms);
    return o;
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

napi_value denoise(napi_env env, napi_callback_info info) {
    napi_value argv[3];
    if (!get_args(env, info, 3, argv)) return nullptr;
    pt_ctx* ctx = get_ctx(env, argv[0]);
    if (!ctx) return nullptr;
    void* data;
    size_t n;
    if (!get_bytes(env, argv[1], &data, &n) || n != sizeof(pt_denoise_params)) {
        napi_throw_type_error(env, nullptr, "uniforms must be the 48-byte denoiser uniform");
        return nullptr;
    }
    pt_denoise_params u;
    std::memcpy(&u, data, sizeof(u));
    int32_t mode = 0;
    napi_get_value_int32(env, argv[2], &mode);
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

napi_value unet_denoise(napi_env env, napi_callback_info info) {
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return nullptr;
    pt_ctx* ctx = get_ctx(env, argv[0]);
    if (!ctx) return nullptr;
    double exposure = 1.0;
    napi_get_value_double(env, argv[1], &exposure);
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

pt_batch* get_batch(napi_env env, napi_value v) {
    void* p = nullptr;
    if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
        napi_throw_type_error(env, nullptr, "first argument must be a pt batch");
        return nullptr;
    }
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
    napi_value r;
    napi_create_uint32(env, idx, &r);
    return r;
}

napi_value batch_frame(napi_env env, napi_callback_info info) {
    napi_value argv[4];
    if (!get_args(env, info, 4, argv)) return nullptr;
    pt_batch* b = get_batch(env, argv[0]);
    if (!b) return nullptr;
    uint32_t n = 0;
    pt_batch_size(b, &n);
    std::vector<pt_scene_params> params(n);
    std::vector<uint32_t> passes(n, 0);
    std::vector<const uint32_t*> seeds(n, nullptr);
    for (uint32_t i = 0; i < n; ++i) {
        napi_value pv, nv, sv;
        void* data;
        size_t len;
        if (napi_get_element(env, argv[1], i, &pv) != napi_ok || !get_bytes(env, pv, &data, &len) ||
            len != sizeof(pt_scene_params)) {
            napi_throw_type_error(env, nullptr, "params must hold one 112-byte SceneParams per job");
            return nullptr;
        }
        std::memcpy(&params[i], data, sizeof(pt_scene_params));
        if (napi_get_element(env, argv[2], i, &nv) != napi_ok || napi_get_value_uint32(env, nv, &passes[i]) != napi_ok) {
            napi_throw_type_error(env, nullptr, "passes must hold one count per job");
            return nullptr;
        }
        if (napi_get_element(env, argv[3], i, &sv) == napi_ok && get_bytes(env, sv, &data, &len)) {
            if (len < size_t(passes[i]) * 4) {
                napi_throw_type_error(env, nullptr, "seeds must hold one u32 per pass");
                return nullptr;
            }
            seeds[i] = static_cast<const uint32_t*>(data);
        }
    }
    std::vector<pt_stats> st(n);
    int rc = pt_batch_frame(b, params.data(), passes.data(), seeds.data(), st.data());
    if (rc) return throw_batch(env, b, rc, "pt_batch_frame");
    napi_value arr;
    napi_create_array_with_length(env, n, &arr);
    for (uint32_t i = 0; i < n; ++i) {
        napi_value o;
        napi_create_object(env, &o);
        set_num(env, o, "extensionRays", double(st[i].extension_rays));
        set_num(env, o, "shadowRays", double(st[i].shadow_rays));
        set_num(env, o, "paths", double(st[i].paths));
        set_num(env, o, "kernelMs", st[i].kernel_ms);
        napi_set_element(env, arr, i, o);
    }
    return arr;
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
    };
    napi_define_properties(env, exports, sizeof(props) / sizeof(props[0]), props);
    return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
