// shade.h — environment map, light sampling, MIS and the Disney-style BSDF
// (shader.wgsl:170-254, 1442-2271), for the gfx950 kernels.
#pragma once
#include "traverse.h"

namespace pt {

// ------------------------------------------------------------------ sky ----
__device__ __forceinline__ v3 sky_texel(const DevScene& S, int x, int y) {
    // one 8-byte load: 4 halfs
    const uint2 raw = *reinterpret_cast<const uint2*>(S.sky + (size_t(y) * S.sky_w + x) * 4);
    const _Float16 r = __builtin_bit_cast(_Float16, uint16_t(raw.x & 0xffffu));
    const _Float16 g = __builtin_bit_cast(_Float16, uint16_t(raw.x >> 16));
    const _Float16 b = __builtin_bit_cast(_Float16, uint16_t(raw.y & 0xffffu));
    return mk(float(r), float(g), float(b));
}
// textureSampleLevel on the sky: bilinear, fp32 weights, clamp-to-edge.
static __device__ v3 sky_bilinear(const DevScene& S, float u, float v) {
    const float W = float(S.sky_w), H = float(S.sky_h);
    float x = u * W - 0.5f, y = v * H - 0.5f;
    x = mn(mx(x, -1.0f), W);
    y = mn(mx(y, -1.0f), H);
    const float fx = floorf(x), fy = floorf(y);
    const float ax = x - fx, ay = y - fy;
    const int mxi = int(S.sky_w) - 1, myi = int(S.sky_h) - 1;
    int x0 = int(fx), y0 = int(fy);
    const int x1 = min(max(x0 + 1, 0), mxi), y1 = min(max(y0 + 1, 0), myi);
    x0 = min(max(x0, 0), mxi);
    y0 = min(max(y0, 0), myi);
    const v3 top = sky_texel(S, x0, y0) * (1.0f - ax) + sky_texel(S, x1, y0) * ax;
    const v3 bot = sky_texel(S, x0, y1) * (1.0f - ax) + sky_texel(S, x1, y1) * ax;
    return top * (1.0f - ay) + bot * ay;
}
// sample_sky, shader.wgsl:176-188
static __device__ v3 sample_sky(const DevScene& S, v3 dir) {
    const float phi = atan2f(dir.z, dir.x);
    const float th = asinf(clampf(dir.y, -1.0f, 1.0f));
    return sky_bilinear(S, 0.5f + phi / kTWO_PI, 0.5f - th / kPI);
}
// binary_search_cdf, shader.wgsl:192-206
__device__ __forceinline__ uint32_t cdf_search(const float* cdf, uint32_t off, uint32_t size, float targ) {
    uint32_t lo = 0, hi = size;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2u;
        if (cdf[off + mid] < targ) lo = mid + 1u; else hi = mid;
    }
    return max(1u, lo) - 1u;
}

struct Params {
    float sky_total_lum;
    uint32_t sky_width, sky_height, sky_cdf_index;
};

// get_sky_pdf, shader.wgsl:208-229
static __device__ float sky_pdf(const DevScene& S, const Params& P, v3 dir) {
    const float phi = atan2f(dir.z, dir.x);
    const float th = acosf(clampf(dir.y, -1.0f, 1.0f));
    const float u = (phi + kPI) / (2.0f * kPI);
    const float v = th / kPI;
    const v3 c = sky_bilinear(S, u, v);
    const float lum = dot(c, mk(0.2126f, 0.7152f, 0.0722f));
    const float st = sinf(th);
    if (P.sky_total_lum <= 0.0f || st <= 0.0f) return 1.0f / (4.0f * kPI);
    return (lum * float(P.sky_width) * float(P.sky_height)) / (P.sky_total_lum * 2.0f * kPI * kPI);
}

struct LightSample { v3 dir, color; float dist, pdf; };

// sample_env_cdf, shader.wgsl:231-254
static __device__ LightSample sample_env(const DevScene& S, const Params& P, float ux, float uy) {
    LightSample ls;
    const uint32_t vi = cdf_search(S.cdf, 0u, P.sky_height, uy);
    const uint32_t row = P.sky_cdf_index + vi * (P.sky_width + 1u);
    const uint32_t ui = cdf_search(S.cdf, row, P.sky_width, ux);
    const float uc = (float(ui) + 0.5f) / float(P.sky_width);
    const float vc = (float(vi) + 0.5f) / float(P.sky_height);
    const float th = kPI * vc;
    const float phi = 2.0f * kPI * uc - kPI;
    const float st = sinf(th);
    ls.dir = mk(st * cosf(phi), cosf(th), st * sinf(phi));
    ls.color = sky_bilinear(S, uc, vc);
    ls.pdf = sky_pdf(S, P, ls.dir);
    ls.dist = kINF_DIST;
    return ls;
}

// --------------------------------------------------------------- lights ----
// light_sphere_pdf, shader.wgsl:1473-1499
static __device__ float sphere_light_pdf(const pt_light& L, const pt_object& sph, v3 hp, v3 dir) {
    const v3 d = ld3(sph.world_position) - hp;
    const float d2 = dot(d, d);
    const float r2 = L.scale * L.scale;
    if (d2 <= r2) return 1.0f / (4.0f * kPI);
    const float tc = dot(d, dir);
    if (tc <= 0.0f) return 0.0f;
    const float dp = d2 - tc * tc;
    if (dp > r2) return 0.0f;
    const float ctm = sqrtf(mx(0.0f, 1.0f - r2 / d2));
    const float sa = kTWO_PI * (1.0f - ctm);
    return 1.0f / mx(sa, 0.0001f);
}

// build_tangent_space, shader.wgsl:1463-1468
__device__ __forceinline__ void tangent_space(v3 N, v3& T, v3& B) {
    const v3 a = fabsf(N.y) > 0.99999f ? mk(0.f, 0.f, 1.f) : mk(0.f, 1.f, 0.f);
    T = normalize(cross(N, a));
    B = cross(T, N);
}

// sample_light_sphere, shader.wgsl:1501-1549
static __device__ LightSample sphere_light_sample(const DevScene& S, const pt_light& L, const pt_object& sph, v3 hp,
                                           float x0, float x1) {
    const v3 d = ld3(sph.world_position) - hp;
    const float d2 = dot(d, d);
    const float r2 = L.scale * L.scale;
    LightSample ls;
    ls.color = ld3(S.mats[sph.material_idx].emittance);
    if (d2 <= r2) {
        const float z = 1.0f - 2.0f * x0;
        const float st = sqrtf(mx(0.0f, 1.0f - z * z));
        const float phi = kTWO_PI * x1;
        ls.dir = mk(st * cosf(phi), st * sinf(phi), z);
        ls.dist = 0.0f;
        ls.pdf = 1.0f / (4.0f * kPI);
        return ls;
    }
    const float ctm = sqrtf(mx(0.0f, 1.0f - r2 / d2));
    const float ct = 1.0f - x0 * (1.0f - ctm);
    const float st = sqrtf(mx(0.0f, 1.0f - ct * ct));
    const float phi = kTWO_PI * x1;
    const v3 l = mk(st * cosf(phi), st * sinf(phi), ct);
    const v3 w = normalize(d);
    v3 T, B;
    tangent_space(w, T, B);
    ls.dir = normalize(T * l.x + B * l.y + w * l.z);
    const float tc = dot(d, ls.dir);
    const float dp = d2 - tc * tc;
    ls.dist = tc - sqrtf(mx(0.0f, r2 - dp));
    const float sa = kTWO_PI * (1.0f - ctm);
    ls.pdf = 1.0f / mx(sa, 0.0001f);
    return ls;
}

// Projected visible-face areas of a box light (shader.wgsl:1557-1562, 1584-1589).
__device__ __forceinline__ void box_areas(const pt_light& L, v3 hp, float s[3], float a[3]) {
    const v3 v1 = ld3(L.matrix), v2 = ld3(L.matrix + 4), v3_ = ld3(L.matrix + 8), cen = ld3(L.matrix + 12);
    const v3 d = cen - hp;
    s[0] = dot(d, v1) < 0.0f ? 1.0f : -1.0f;
    s[1] = dot(d, v2) < 0.0f ? 1.0f : -1.0f;
    s[2] = dot(d, v3_) < 0.0f ? 1.0f : -1.0f;
    const v3 nd = normalize(-d);
    a[0] = len(cross(v2 * 2.0f, v3_ * 2.0f)) * mx(dot(nd, normalize(v1 * s[0])), 0.01f);
    a[1] = len(cross(v1 * 2.0f, v3_ * 2.0f)) * mx(dot(nd, normalize(v2 * s[1])), 0.01f);
    a[2] = len(cross(v1 * 2.0f, v2 * 2.0f)) * mx(dot(nd, normalize(v3_ * s[2])), 0.01f);
}

// light_box_pdf, shader.wgsl:1551-1573
static __device__ float box_light_pdf(const pt_light& L, v3 hp, v3 rd, float hit_t, v3 hit_n) {
    float s[3], a[3];
    box_areas(L, hp, s, a);
    const float total = a[0] + a[1] + a[2];
    const float cl = mx(dot(-rd, hit_n), 1e-6f);
    return (hit_t * hit_t) / (total * cl);
}

// sample_light_box, shader.wgsl:1575-1624 (draws the face pick itself)
static __device__ LightSample box_light_sample(const DevScene& S, const pt_light& L, const pt_object& ob, v3 hp, float x0,
                                        float x1, Pcg& rng) {
    const v3 v1 = ld3(L.matrix), v2 = ld3(L.matrix + 4), v3_ = ld3(L.matrix + 8), cen = ld3(L.matrix + 12);
    float s[3], a[3];
    box_areas(L, hp, s, a);
    const float total = a[0] + a[1] + a[2];
    v3 p = cen, n;
    const float u = x0 * 2.0f - 1.0f, v = x1 * 2.0f - 1.0f;
    const float pick = rng.next() * total;
    if (pick < a[0]) {
        p = p + (v1 * s[0] + v2 * u + v3_ * v);
        n = normalize(v1 * s[0]);
    } else if (pick < a[0] + a[1]) {
        p = p + (v2 * s[1] + v1 * u + v3_ * v);
        n = normalize(v2 * s[1]);
    } else {
        p = p + (v3_ * s[2] + v1 * u + v2 * v);
        n = normalize(v3_ * s[2]);
    }
    LightSample ls;
    const v3 delta = p - hp;
    const float dsq = dot(delta, delta);
    ls.dist = sqrtf(dsq);
    ls.dir = delta / ls.dist;
    const float cl = mx(dot(-ls.dir, n), 1e-6f);
    ls.pdf = dsq / (total * cl);
    ls.color = ld3(S.mats[ob.material_idx].emittance);
    return ls;
}

// mis_weight, shader.wgsl:1648-1659
__device__ __forceinline__ float mis_weight(float pa, float pb) {
    if (pa <= 0.0f) return 0.0f;
    const float m = mx(pa, pb);
    const float a = pa / m, b = pb / m;
    const float a2 = a * a, b2 = b * b;
    const float sum = a2 + b2;
    if (sum <= 0.0f) return 0.0f;
    return a2 / sum;
}

// ---------------------------------------------------------------- media ----
struct Medium { float ior; v3 sigma, emission; };

// MediumStack (shader.wgsl:1661-1697): up to 4 media of (ior, sigma,
// emission) per lane, kept in LDS and interleaved across the block's lanes
// (slot s, field f of lane l at ((s*7+f)*blockDim + l)), so the rarely
// touched stack costs no registers and no scratch traffic.
constexpr int kMediaFloats = 4 * 7;
struct Media {
    float* base;
    uint32_t stride;
    int n;
    __device__ __forceinline__ Medium at(int s) const {
        const float* p = base + uint32_t(s) * 7u * stride;
        return Medium{p[0], mk(p[stride], p[2 * stride], p[3 * stride]),
                      mk(p[4 * stride], p[5 * stride], p[6 * stride])};
    }
    __device__ __forceinline__ Medium top() const {
        if (n > 0) return at(n - 1);
        return Medium{1.0f, splat(0.f), splat(0.f)};
    }
    __device__ __forceinline__ float top_ior() const { return n > 0 ? base[uint32_t(n - 1) * 7u * stride] : 1.0f; }
    __device__ __forceinline__ float outer_ior() const { return n > 1 ? base[uint32_t(n - 2) * 7u * stride] : 1.0f; }
    __device__ __forceinline__ void push(const Medium& x) {
        if (n < 4) {
            float* p = base + uint32_t(n) * 7u * stride;
            p[0] = x.ior;
            p[stride] = x.sigma.x; p[2 * stride] = x.sigma.y; p[3 * stride] = x.sigma.z;
            p[4 * stride] = x.emission.x; p[5 * stride] = x.emission.y; p[6 * stride] = x.emission.z;
            n += 1;
        }
    }
    __device__ __forceinline__ void pop() { if (n > 0) n -= 1; }
};

// ------------------------------------------------------------------ BSDF ---
struct Surface { v3 normal, snormal, albedo; float roughness, metallic; v3 emittance; float alpha; };

// get_surface_context's maps (shader.wgsl:1393-1440) at the texture
// coordinates uv (hit_uv * uv_scale, moved by parallax mapping where the
// material has a height map): albedo (and its alpha) and emittance times
// texel^2.2, roughness times green, metallic times blue, and the normal map
// in the hit's tangent frame (T, B, N) with its x and y scaled by
// height_params.x.  sc.normal must hold N on entry.
__device__ __forceinline__ void apply_textures(const DevScene& S, const pt_material& m, v2 uv, v3 T, v3 B,
                                               Surface& sc) {
    if (m.normal_idx >= 0) {
        const float4 t = sample_texture(S, m.normal_idx, uv);
        const float k = m.height_params[0];
        const v3 nm = normalize(mk((t.x * 2.0f - 1.0f) * k, (t.y * 2.0f - 1.0f) * k, t.z * 2.0f - 1.0f));
        sc.normal = normalize(T * nm.x + B * nm.y + sc.normal * nm.z);
    }
    if (m.albedo_idx >= 0) {
        const float4 t = sample_texture(S, m.albedo_idx, uv);
        sc.alpha = t.w;
        sc.albedo = sc.albedo * mk(powf(t.x, 2.2f), powf(t.y, 2.2f), powf(t.z, 2.2f));
    }
    if (m.roughness_idx >= 0) sc.roughness *= sample_texture(S, m.roughness_idx, uv).y;
    if (m.metallic_idx >= 0) sc.metallic *= sample_texture(S, m.metallic_idx, uv).z;
    if (m.emissive_idx >= 0) {
        const float4 t = sample_texture(S, m.emissive_idx, uv);
        sc.emittance = sc.emittance * mk(powf(t.x, 2.2f), powf(t.y, 2.2f), powf(t.z, 2.2f));
    }
}

// The material's own surface before any map.
__device__ __forceinline__ Surface base_surface(const pt_material& m, v3 n) {
    Surface sc;
    sc.normal = n;
    sc.snormal = n;
    sc.albedo = ld3(m.color);
    sc.alpha = 1.0f;
    sc.metallic = m.metallic;
    sc.roughness = m.roughness;
    sc.emittance = ld3(m.emittance);
    return sc;
}

struct PomRes {
    v2 uv;
    float h;
    bool hit;
};

// Depth (0 at the surface, 1 at the bottom) the height map gives at uv:
// height_params.y < 0 marks a height map (white high), else a depth map.
__device__ __forceinline__ float pom_depth(const DevScene& S, const pt_material& m, v2 uv) {
    const float s = sample_texture(S, m.height_idx, uv).x;
    return m.height_params[1] < 0.0f ? 1.0f - s : s;
}

// calculate_pom (shader.wgsl:1238-1278) from depth 0: march the view ray in
// height_params.z layers (at most 64) until it falls below the map, then
// interpolate between the last two layers.  A march that leaves [0, 1]
// returns uv (0, 0) and depth 0, as the reference's zero-initialised result.
__device__ PomRes pom_march(const DevScene& S, const pt_material& m, v2 uv0, v3 vts) {
    PomRes r{v2{0.0f, 0.0f}, 0.0f, false};
    const float scale = fabsf(m.height_params[1]);
    const float nl = m.height_params[2];
    const float ld = 1.0f / nl;
    float cur = 0.0f;
    const v2 P{-(vts.x / vts.z) * scale, -(vts.y / vts.z) * scale};
    const v2 dt{P.x / nl, P.y / nl};
    v2 tc{uv0.x - P.x * m.height_params[3], uv0.y - P.y * m.height_params[3]};
    float dv = pom_depth(S, m, tc);
    for (int i = 0; i < 64; ++i) {
        if (float(i) >= nl || cur >= dv) break;
        if (cur > 1.0f || cur < 0.0f) return r;
        tc = v2{tc.x + dt.x, tc.y + dt.y};
        dv = pom_depth(S, m, tc);
        cur += ld;
    }
    const v2 pc{tc.x - dt.x, tc.y - dt.y};
    const float pdv = pom_depth(S, m, pc);
    const float after = dv - cur;
    const float before = pdv - cur + ld;
    const float w = after / (after - before);
    r.uv = v2{tc.x * (1.0f - w) + pc.x * w, tc.y * (1.0f - w) + pc.y * w};
    r.h = cur * (1.0f - w) + (cur - ld) * w;
    r.hit = true;
    return r;
}

// calculate_shadow_pom (shader.wgsl:1280-1329): march from (uv, depth)
// toward the surface along lts (tangent space); hit when the map rises
// above the ray or when lts points into the surface.
__device__ PomRes pom_shadow(const DevScene& S, const pt_material& m, v2 uv, float h, v3 lts) {
    PomRes r{uv, h, false};
    if (lts.z <= 0.0f) {
        r.hit = true;
        return r;
    }
    const float scale = fabsf(m.height_params[1]);
    const float nl = m.height_params[2];
    const float ld = 1.0f / nl;
    const v2 p{(lts.x / lts.z) * scale * ld, (lts.y / lts.z) * scale * ld};
    v2 su = uv;
    float sd = h;
    for (int i = 0; i < 32; ++i) {
        if (float(i) / 2.0f >= nl || sd <= 0.0f) break;
        su = v2{su.x + p.x, su.y + p.y};
        sd -= ld;
        const float md = pom_depth(S, m, su);
        if (md < sd) {
            r.hit = true;
            r.uv = su;
            r.h = md;
            return r;
        }
    }
    return r;
}

// Tangent-space coordinates of a world direction, (T, B, N)^T * v.
__device__ __forceinline__ v3 to_tangent(v3 T, v3 B, v3 N, v3 v) { return mk(dot(T, v), dot(B, v), dot(N, v)); }

__device__ __forceinline__ float d_ggx(float nh, float a2) {
    const float den = (nh * nh * (a2 - 1.0f) + 1.0f);
    return a2 / (kPI * den * den + 1e-10f);
}
static __device__ float fresnel_diel(float ci, float etai, float etat) {
    const float cosi = clampf(fabsf(ci), 0.0f, 1.0f);
    const float sint = (etai / etat) * sqrtf(mx(0.0f, 1.0f - cosi * cosi));
    if (sint >= 1.0f) return 1.0f;
    const float cost = sqrtf(mx(0.0f, 1.0f - sint * sint));
    const float Rs = ((etat * cosi) - (etai * cost)) / ((etat * cosi) + (etai * cost));
    const float Rp = ((etai * cosi) - (etat * cost)) / ((etai * cosi) + (etat * cost));
    return (Rs * Rs + Rp * Rp) / 2.0f;
}
__device__ __forceinline__ float g2_div_nv(float nv, float nl, float a2) {
    const float sv = sqrtf(mx(0.0f, a2 + nv * nv * (1.0f - a2)));
    const float sl = sqrtf(mx(0.0f, a2 + nl * nl * (1.0f - a2)));
    return (2.0f * nl) / (nl * sv + nv * sl + 1e-10f);
}
__device__ __forceinline__ float g2_unc_div_nv(float nv, float nl, float a2) {
    const float lv = nv + sqrtf(mx(0.0f, a2 + nv * nv * (1.0f - a2)));
    const float ll = nl + sqrtf(mx(0.0f, a2 + nl * nl * (1.0f - a2)));
    return (4.0f * nl) / (lv * ll + 1e-7f);
}
__device__ __forceinline__ float g1_div_nv(float nv, float a2) {
    return 2.0f / (nv + sqrtf(mx(0.0f, a2 + (1.0f - a2) * nv * nv)) + 1e-7f);
}
__device__ __forceinline__ float pow5(float x) { return powf(x, 5.0f); }

// ImportanceSampleCosine, shader.wgsl:1751-1763
static __device__ v3 cosine_sample(float x0, float x1, v3 N) {
    const float phi = kTWO_PI * x0;
    const float ct = sqrtf(clampf(x1, 0.0f, 1.0f));
    const float st = sqrtf(mx(0.0f, 1.0f - x1));
    const v3 L = mk(cosf(phi) * st, sinf(phi) * st, ct);
    const v3 up = fabsf(N.y) > 0.999f ? mk(0.f, 0.f, 1.f) : mk(0.f, 1.f, 0.f);
    const v3 T = normalize(cross(up, N));
    const v3 B = cross(N, T);
    return normalize(T * L.x + B * L.y + N * L.z);
}
// ImportanceSampleVNDF_GGX, shader.wgsl:1806-1829
static __device__ v3 vndf_sample(float x0, float x1, v3 V, v3 N, float ag) {
    const v3 up = fabsf(N.y) > 0.999f ? mk(0.f, 0.f, 1.f) : mk(0.f, 1.f, 0.f);
    const v3 T = normalize(cross(up, N));
    const v3 B = cross(N, T);
    const v3 vl = mk(dot(V, T), dot(V, B), dot(V, N));
    const v3 Vh = normalize(mk(ag * vl.x, ag * vl.y, mx(vl.z, 0.0f)));
    const float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    const v3 T1 = lensq > 0.0f ? mk(-Vh.y, Vh.x, 0.0f) / sqrtf(mx(lensq, 1e-7f)) : mk(1.0f, 0.0f, 0.0f);
    const v3 T2 = cross(Vh, T1);
    const float r = sqrtf(x0);
    const float phi = 2.0f * kPI * x1;
    const float t1 = r * cosf(phi);
    const float t2 = r * sinf(phi);
    const float s = 0.5f * (1.0f + Vh.z);
    const float t2m = lerp(sqrtf(mx(0.0f, 1.0f - t1 * t1)), t2, s);
    const v3 Nh = t1 * T1 + t2m * T2 + sqrtf(mx(0.0f, 1.0f - t1 * t1 - t2m * t2m)) * Vh;
    const v3 Ne = normalize(mk(ag * Nh.x, ag * Nh.y, mx(0.0f, Nh.z)));
    return normalize(T * Ne.x + B * Ne.y + N * Ne.z);
}

__device__ __forceinline__ void iors(bool entering, const pt_material& m, const Media& md, float& etai, float& etat) {
    etai = md.top_ior();
    etat = entering ? m.ior : md.outer_ior();
}

// eval_surface, shader.wgsl:1831-1947
static __device__ v3 eval_bsdf(v3 V, v3 L, const pt_material& m, const Surface& sc, const Media& md) {
    const bool entering = dot(sc.normal, V) > 0.0f;
    const v3 no = entering ? sc.normal : -sc.normal;
    const float nv = dot(no, V);
    const float nl = dot(no, L);
    if (nv <= 0.0f) return splat(0.0f);
    const float rv = clampf(sc.roughness, 0.01f, 1.0f);
    const float alpha = rv * rv;
    const float a2 = mx(alpha * alpha, 1e-6f);
    float etai, etat;
    iors(entering, m, md, etai, etat);
    if (nl > 0.0f) {
        const v3 Hv = V + L;
        const v3 H = dot(Hv, Hv) > 1e-6f ? normalize(Hv) : no;
        const float nh = mx(dot(no, H), 0.0f);
        const float vh = mx(dot(V, H), 0.0f);
        v3 dc = sc.albedo;
        if (m.sheen > 0.0f) dc = dc + lerp(splat(1.0f), sc.albedo, m.sheen_tint) * pow5(mx(0.0f, 1.0f - nv)) * m.sheen;
        if (m.subsurface > 0.0f) dc = lerp(dc, ld3(m.subsurface_tint), m.subsurface);
        const v3 diffuse = (dc / kPI) * (1.0f - sc.metallic) * (1.0f - m.transmission);
        const float D = d_ggx(nh, a2);
        const float G = g2_div_nv(nv, nl, a2);
        const float Fd = fresnel_diel(vh, etai, etat);
        const float lum = dot(sc.albedo, mk(0.2126f, 0.7152f, 0.0722f));
        const v3 tint = lum <= 0.0f ? splat(1.0f) : sc.albedo / mx(lum, 0.0001f);
        const v3 Fdt = lerp(splat(Fd), splat(Fd) * tint, m.specular_tint);
        const v3 F90 = lerp(splat(1.0f), tint, m.specular_tint);
        const v3 Fm = sc.albedo + (F90 - sc.albedo) * pow5(mx(0.0f, 1.0f - vh));
        const v3 Fa = lerp(Fdt, Fm, sc.metallic);
        const float Fmr = fresnel_diel(nv, etai, etat);
        const float F0d = powf((etai - etat) / (etai + etat), 2.0f);
        const float Fme = lerp(Fmr, F0d, rv);
        const v3 Fmm = sc.albedo + (splat(1.0f) - sc.albedo) * pow5(mx(0.0f, 1.0f - nv));
        const v3 Fma = lerp(splat(Fme), Fmm, sc.metallic);
        const v3 diffuse_cos = diffuse * nl * (splat(1.0f) - Fma);
        const float ms = 1.0f + alpha * 0.4f;
        const v3 spec_cos = (D * Fa * G) / 4.0f * ms;
        v3 cc_cos = splat(0.0f);
        float cc_att = 1.0f;
        if (m.clearcoat > 0.0f && entering) {
            const float cr = clampf(1.0f - m.clearcoat_gloss, 0.01f, 1.0f);
            const float ca = cr * cr;
            const float ca2 = mx(ca * ca, 1e-6f);
            const float cD = d_ggx(nh, ca2);
            const float cG = g2_div_nv(nv, nl, ca2);
            const float cF = fresnel_diel(vh, 1.0f, m.clearcoat_ior);
            cc_cos = (m.clearcoat * cD * cG * splat(cF)) / 4.0f;
            const float cFm = fresnel_diel(nv, 1.0f, m.clearcoat_ior);
            cc_att = 1.0f - (m.clearcoat * cFm);
        }
        return ((diffuse_cos + spec_cos) * cc_att + cc_cos) * sc.alpha;
    } else if (nl < 0.0f && m.transmission > 0.0f) {
        v3 H = normalize(-(etai * V + etat * L));
        if (dot(H, no) < 0.0f) H = -H;
        const float nh = mx(dot(no, H), 0.0f);
        const float vh = dot(V, H);
        const float lh = dot(L, H);
        if (vh * lh >= 0.0f) return splat(0.0f);
        const float D = d_ggx(nh, a2);
        const float G = g2_unc_div_nv(nv, fabsf(nl), a2);
        const float Fd = fresnel_diel(vh, etai, etat);
        const float den = etai * vh + etat * lh;
        const float den2 = mx(den * den, 1e-7f);
        const float ms = 1.0f + alpha * 0.25f;
        const float btdf = (fabsf(vh) * fabsf(lh) * etat * etat * (1.0f - Fd) * D * G) / den2 * ms;
        float cc_att = 1.0f;
        if (m.clearcoat > 0.0f && entering) {
            const float cFm = fresnel_diel(nv, 1.0f, m.clearcoat_ior);
            cc_att = 1.0f - (m.clearcoat * cFm);
        }
        const v3 tc = splat(m.transmission * (1.0f - sc.metallic));
        return tc * btdf * cc_att * sc.alpha;
    }
    return splat(0.0f);
}

struct Lobes { float cc, spec, trans, diff; };

// Lobe weights of pdf_surface (shader.wgsl:1975-1987) and sample_surface
// (shader.wgsl:2078-2090), evaluated at cosine cv.
static __device__ Lobes lobes(float cv, bool entering, float etai, float etat, float rv, const pt_material& m, const Surface& sc) {
    const float Fcc = fresnel_diel(cv, 1.0f, m.clearcoat_ior);
    const float pcc = entering ? m.clearcoat * Fcc : 0.0f;
    const float Fr = fresnel_diel(cv, etai, etat);
    const float F0 = powf((etai - etat) / (etai + etat), 2.0f);
    const float Fe = lerp(Fr, F0, rv);
    const v3 Fm = sc.albedo + (splat(1.0f) - sc.albedo) * pow5(mx(0.0f, 1.0f - cv));
    const v3 Fa = lerp(splat(Fe), Fm, sc.metallic);
    const float favg = clampf((Fa.x + Fa.y + Fa.z) / 3.0f, 0.0f, 1.0f);
    Lobes l;
    l.cc = pcc;
    l.spec = (1.0f - pcc) * favg;
    l.trans = (1.0f - pcc) * (1.0f - favg) * m.transmission * (1.0f - sc.metallic);
    l.diff = (1.0f - pcc) * (1.0f - favg) * (1.0f - m.transmission) * (1.0f - sc.metallic);
    return l;
}

// pdf_surface, shader.wgsl:1952-2047
static __device__ float pdf_bsdf(v3 V, v3 L, const pt_material& m, const Surface& sc, const Media& md) {
    const bool entering = dot(sc.normal, V) > 0.0f;
    const v3 no = entering ? sc.normal : -sc.normal;
    const float nv = dot(no, V);
    const float nl = dot(no, L);
    if (nv <= 0.0f) return 1e-7f;
    const float rv = clampf(sc.roughness, 0.01f, 1.0f);
    const float alpha = rv * rv;
    const float a2 = mx(alpha * alpha, 1e-6f);
    float etai, etat;
    iors(entering, m, md, etai, etat);
    const Lobes P = lobes(nv, entering, etai, etat, rv, m, sc);
    float pdf = 0.0f;
    if (nl > 0.0f) {
        const v3 Hv = V + L;
        const v3 H = dot(Hv, Hv) > 1e-6f ? normalize(Hv) : no;
        const float nh = mx(dot(no, H), 0.0f);
        const float vh = mx(dot(V, H), 0.0f);
        if (P.diff > 0.0f) pdf += P.diff * (nl / kPI);
        if (P.spec > 0.0f) {
            const float D = d_ggx(nh, a2);
            const float G1 = g1_div_nv(nv, a2);
            float sw = P.spec;
            if (P.trans > 0.0f) {
                const float ci = clampf(vh, 0.0f, 1.0f);
                const float si = (etai / etat) * sqrtf(mx(0.0f, 1.0f - ci * ci));
                if (si >= 1.0f) sw += P.trans;
            }
            const float pH = G1 * D * mx(0.0f, vh);
            pdf += sw * pH / mx(4.0f * vh, 1e-7f);
        }
        if (P.cc > 0.0f) {
            const float cr = clampf(1.0f - m.clearcoat_gloss, 0.01f, 1.0f);
            const float ca = cr * cr;
            const float ca2 = mx(ca * ca, 1e-6f);
            const float cD = d_ggx(nh, ca2);
            const float cG1 = g1_div_nv(nv, ca2);
            const float pHc = cG1 * cD * mx(0.0f, vh);
            pdf += P.cc * pHc / mx(4.0f * vh, 1e-7f);
        }
    } else if (nl < 0.0f && P.trans > 0.0f) {
        v3 H = normalize(-(etai * V + etat * L));
        if (dot(H, no) < 0.0f) H = -H;
        const float nh = mx(dot(no, H), 0.0f);
        const float vh = dot(V, H);
        const float lh = dot(L, H);
        if (vh * lh < 0.0f) {
            const float D = d_ggx(nh, a2);
            const float G1 = g1_div_nv(nv, a2);
            const float pH = G1 * D * mx(0.0f, vh);
            const float den = etai * vh + etat * lh;
            const float j = (etat * etat * fabsf(lh)) / mx(den * den, 1e-7f);
            pdf += P.trans * pH * j;
        }
    }
    return mx(pdf, 1e-7f) * sc.alpha;
}

__device__ __forceinline__ v3 medium_sigma(const Surface& sc, const pt_material& m) {
    return -mk(logf(mx(sc.albedo.x, 0.0001f)), logf(mx(sc.albedo.y, 0.0001f)), logf(mx(sc.albedo.z, 0.0001f))) *
           m.concentration;
}

// sample_surface, shader.wgsl:2052-2175.  Returns false when the path ends.
static __device__ bool sample_bsdf(v3& ro, v3& rd, v3& tp, const pt_material& m, const Surface& sc, v3 hp, Media& md,
                            float& last_pdf, Pcg& rng) {
    const v3 V = -rd;
    const bool entering = dot(sc.normal, V) > 0.0f;
    const v3 no = entering ? sc.normal : -sc.normal;
    const v3 sno = entering ? sc.snormal : -sc.snormal;
    const float snv = mx(dot(sno, V), 1e-6f);
    const float rv = clampf(sc.roughness, 0.01f, 1.0f);
    const float alpha = rv * rv;
    const float cr = clampf(1.0f - m.clearcoat_gloss, 0.01f, 1.0f);
    const float ca = cr * cr;
    float etai, etat;
    iors(entering, m, md, etai, etat);
    const Lobes P = lobes(snv, entering, etai, etat, rv, m, sc);
    const float total = P.cc + P.spec + P.trans + P.diff;
    if (rng.next() > sc.alpha) {
        ro = hp - no * 0.005f;
        return true;
    }
    const float pick = rng.next();
    const float x0 = rng.next();
    const float x1 = rng.next();
    v3 L = splat(0.0f);
    bool trans = false;
    if (pick < P.cc) {
        L = reflect(-V, vndf_sample(x0, x1, V, sno, ca));
    } else if (pick < P.cc + P.spec) {
        v3 Ns = no;
        if (m.anisotropic > 0.0f) {
            const v3 up = fabsf(no.y) > 0.99999f ? mk(0.f, 0.f, 1.f) : mk(0.f, 1.f, 0.f);
            v3 T = normalize(cross(up, no));
            const v3 B = cross(no, T);
            const float ang = m.aniso_rotation * kTWO_PI;
            T = T * cosf(ang) + B * sinf(ang);
            Ns = normalize(lerp(no, cross(cross(V, T), T), m.anisotropic));
        }
        L = reflect(-V, vndf_sample(x0, x1, V, Ns, alpha));
    } else if (pick < P.cc + P.spec + P.trans) {
        if (fabsf(etai - etat) < 1e-4f) {
            rd = -V;
            ro = hp - no * 0.001f;
            tp = tp * splat(m.transmission * (1.0f - sc.metallic));
            last_pdf = P.trans;
            if (entering) md.push(Medium{m.ior, medium_sigma(sc, m), ld3(m.emittance)});
            else md.pop();
            return true;
        }
        const v3 H = vndf_sample(x0, x1, V, no, alpha);
        L = refract(-V, H, etai / etat);
        if (len(L) < 0.1f) L = reflect(-V, H);
        else trans = true;
    } else if (pick < total) {
        L = cosine_sample(x0, x1, no);
    } else {
        return false;
    }
    const float nl = dot(no, L);
    if (trans) {
        if (nl >= 0.0f) return false;
        ro = hp - no * 0.001f;
    } else {
        if (nl <= 0.0f) return false;
        ro = hp + no * 0.001f;
    }
    rd = L;
    const v3 f = eval_bsdf(V, L, m, sc, md);
    const float pdf = pdf_bsdf(V, L, m, sc, md);
    if (pdf > 0.0f) tp = tp * (f / pdf);
    else return false;
    last_pdf = pdf;
    if (trans) {
        if (entering) md.push(Medium{m.ior, medium_sigma(sc, m), ld3(m.emittance)});
        else md.pop();
    }
    return true;
}

// clamp_firefly, shader.wgsl:2179-2189
__device__ __forceinline__ v3 firefly(v3 r, int bounce) {
    if (bounce == 0) return r;
    const float l = dot(r, mk(0.2126f, 0.7152f, 0.0722f));
    if (l > 20.0f) return r * (20.0f / l);
    return r;
}

}  // namespace pt
