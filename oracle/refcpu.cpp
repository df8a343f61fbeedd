// refcpu.cpp — the oracle (test infrastructure only; see refcpu.h).
//
// A line-by-line FP32 restatement of shader.wgsl / denoiser.wgsl in plain
// C++, written for readability and checkability rather than speed: private
// 64-entry stacks, 32 B BVH nodes read straight from the tracer.js arrays,
// no precomputation.  Built with -ffp-contract=off so each WGSL expression
// rounds exactly once per operation, left to right.
#include "refcpu.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <new>
#include <thread>
#include <type_traits>
#include <vector>

namespace {

// ----------------------------------------------------------- vectors -------
struct f2 { float x, y; };
struct f3 { float x, y, z; };

inline f3 V(float x, float y, float z) { return f3{x, y, z}; }
inline f3 V(float s) { return f3{s, s, s}; }
inline f3 operator+(f3 a, f3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
inline f3 operator-(f3 a, f3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
inline f3 operator*(f3 a, f3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
inline f3 operator/(f3 a, f3 b) { return V(a.x / b.x, a.y / b.y, a.z / b.z); }
inline f3 operator*(f3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
inline f3 operator*(float s, f3 a) { return V(s * a.x, s * a.y, s * a.z); }
inline f3 operator/(f3 a, float s) { return V(a.x / s, a.y / s, a.z / s); }
inline f3 operator-(f3 a) { return V(-a.x, -a.y, -a.z); }
inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline f3 cross(f3 a, f3 b) { return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline float length(f3 a) { return std::sqrt(dot(a, a)); }
inline f3 normalize(f3 a) { float l = length(a); return a / l; }
inline float fmn(float a, float b) { return std::fmin(a, b); }
inline float fmx(float a, float b) { return std::fmax(a, b); }
inline f3 fmn(f3 a, f3 b) { return V(fmn(a.x, b.x), fmn(a.y, b.y), fmn(a.z, b.z)); }
inline f3 fmx(f3 a, f3 b) { return V(fmx(a.x, b.x), fmx(a.y, b.y), fmx(a.z, b.z)); }
inline float clampf(float x, float lo, float hi) { return fmn(fmx(x, lo), hi); }
inline float max_component(f3 v) { return fmx(fmx(v.x, v.y), v.z); }
inline float min_component(f3 v) { return fmn(fmn(v.x, v.y), v.z); }
inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
inline f3 mix3(f3 a, f3 b, float t) { return a * (1.0f - t) + b * t; }
inline f3 mix3(f3 a, f3 b, f3 t) { return a * (V(1.0f) - t) + b * t; }
inline float signf(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }
inline f3 reflect(f3 e1, f3 e2) { return e1 - (2.0f * dot(e2, e1)) * e2; }
inline f3 refract(f3 e1, f3 e2, float e3) {
    float d = dot(e2, e1);
    float k = 1.0f - e3 * e3 * (1.0f - d * d);
    if (k < 0.0f) return V(0.f);
    return e3 * e1 - (e3 * d + std::sqrt(k)) * e2;
}
inline f3 ld3(const float* p) { return V(p[0], p[1], p[2]); }

// mat4 (column-major, 16 floats) times vec4(v, w).
inline f3 mul_point(const float* m, f3 v) {
    return V(m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * 1.0f,
             m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * 1.0f,
             m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * 1.0f);
}
inline f3 mul_dir(const float* m, f3 v) {
    return V(m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * 0.0f,
             m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * 0.0f,
             m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * 0.0f);
}
// transpose(mat4) * vec4(v, 0): row r of the result is dot(column r, v).
inline f3 mul_transpose_dir(const float* m, f3 v) {
    return V(m[0] * v.x + m[1] * v.y + m[2] * v.z + m[3] * 0.0f,
             m[4] * v.x + m[5] * v.y + m[6] * v.z + m[7] * 0.0f,
             m[8] * v.x + m[9] * v.y + m[10] * v.z + m[11] * 0.0f);
}
// transpose(mat3x3(col0.xyz, col1.xyz, col2.xyz)) * v
inline f3 mul_normal(const float* m, f3 v) {
    return V(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
             m[8] * v.x + m[9] * v.y + m[10] * v.z);
}
// mat3x3(col0.xyz, col1.xyz, col2.xyz) * v (torus normal, no transpose)
inline f3 mul_mat3(const float* m, f3 v) {
    return V(m[0] * v.x + m[4] * v.y + m[8] * v.z, m[1] * v.x + m[5] * v.y + m[9] * v.z,
             m[2] * v.x + m[6] * v.y + m[10] * v.z);
}

constexpr float PI = 3.14159265359f;
constexpr float TWO_PI = 6.28318530718f;
constexpr float INFINITY_ = 65504.0f;
constexpr int STACK_SIZE = 64;

float half_to_float(uint16_t h) {
    int sign = (h >> 15) & 1, ex = (h >> 10) & 0x1f, mant = h & 0x3ff;
    float v;
    if (ex == 0) v = std::ldexp(float(mant), -24);
    else if (ex == 31) v = mant ? NAN : INFINITY;
    else v = std::ldexp(float(mant + 1024), ex - 25);
    return sign ? -v : v;
}

uint8_t to_unorm8(float x) {
    float v = fmn(fmx(x, 0.0f), 1.0f);
    return uint8_t(v * 255.0f + 0.5f);
}

struct Counters {
    uint64_t ext = 0, shadow = 0, paths = 0, pops = 0, inner = 0, objs = 0, tris = 0, shade = 0;
    void add(const Counters& o) {
        ext += o.ext; shadow += o.shadow; paths += o.paths; pops += o.pops; inner += o.inner;
        objs += o.objs; tris += o.tris; shade += o.shade;
    }
};

}  // namespace

struct ref_scene {
    std::vector<pt_material> mats;
    std::vector<pt_object> objs;
    std::vector<pt_bvh_node> tlas;
    std::vector<pt_mesh_instance> meshes;
    std::vector<pt_bvh_node> bvh;
    std::vector<pt_triangle> tris;
    std::vector<pt_plane> planes;
    std::vector<pt_light> lights;
    uint32_t sky_w = 1, sky_h = 1;
    std::vector<float> sky;  // RGBA f32 widened from f16
    std::vector<float> cdf;
    uint32_t flags = 0;
    int bounce_limit = 8;
    std::vector<uint8_t> tex;  // texture array: layers of 512 x 512 rgba8
    uint32_t tex_layers = 0;
    bool has(uint32_t f) const { return (flags & f) != 0; }
};

namespace {

struct Ray { f3 origin, direction; };

struct SurfaceHit {
    float t;
    int m_idx;
    f3 hit_n;
    f2 hit_uv;
    int o_idx;
    int count;
    int prim_id;   // not in the WGSL record: which triangle won
    int plane_idx; // not in the WGSL record: which plane won
    f3 tangent, bitangent;  // set by sphere, cube and plane hits only
};

SurfaceHit empty_hit() {
    SurfaceHit h;
    h.t = 1e10f; h.m_idx = -1; h.hit_n = V(0.f); h.hit_uv = f2{0.f, 0.f}; h.o_idx = -1; h.count = 0;
    h.prim_id = -1; h.plane_idx = -1;
    h.tangent = V(0.f); h.bitangent = V(0.f);
    return h;
}

// --------------------------------------------------------------- sky ------
// textureSampleLevel(skyTex, skySampler(linear, clamp-to-edge), uv, 0).rgb
f3 sky_texel(const ref_scene& s, int x, int y) {
    const float* p = &s.sky[(size_t(y) * s.sky_w + x) * 4];
    return V(p[0], p[1], p[2]);
}
f3 sample_sky_tex(const ref_scene& s, float u, float v) {
    const float W = float(s.sky_w), H = float(s.sky_h);
    float x = u * W - 0.5f, y = v * H - 0.5f;
    x = fmn(fmx(x, -1.0f), W);
    y = fmn(fmx(y, -1.0f), H);
    const float fx = std::floor(x), fy = std::floor(y);
    const float ax = x - fx, ay = y - fy;
    int x0 = int(fx), y0 = int(fy), x1 = x0 + 1, y1 = y0 + 1;
    const int mx = int(s.sky_w) - 1, my = int(s.sky_h) - 1;
    x0 = std::min(std::max(x0, 0), mx); x1 = std::min(std::max(x1, 0), mx);
    y0 = std::min(std::max(y0, 0), my); y1 = std::min(std::max(y1, 0), my);
    const f3 top = sky_texel(s, x0, y0) * (1.0f - ax) + sky_texel(s, x1, y0) * ax;
    const f3 bot = sky_texel(s, x0, y1) * (1.0f - ax) + sky_texel(s, x1, y1) * ax;
    return top * (1.0f - ay) + bot * ay;
}

// shader.wgsl:176-188
f3 sample_sky(const ref_scene& s, f3 dir) {
    float phi = std::atan2(dir.z, dir.x);
    float theta = std::asin(clampf(dir.y, -1.0f, 1.0f));
    float u = 0.5f + phi / TWO_PI;
    float v = 0.5f - theta / PI;
    return sample_sky_tex(s, u, v);
}

// shader.wgsl:192-206
uint32_t binary_search_cdf(const ref_scene& s, uint32_t offset, uint32_t size, float targ) {
    uint32_t low = 0, high = size;
    while (low < high) {
        uint32_t mid = low + (high - low) / 2u;
        float val = s.cdf[offset + mid];
        if (val < targ) low = mid + 1u; else high = mid;
    }
    return std::max(1u, low) - 1u;
}

// shader.wgsl:208-229
float get_sky_pdf(const ref_scene& s, const pt_scene_params& p, f3 dir) {
    float phi = std::atan2(dir.z, dir.x);
    float theta = std::acos(clampf(dir.y, -1.0f, 1.0f));
    float u = (phi + PI) / (2.0f * PI);
    float v = theta / PI;
    f3 c = sample_sky_tex(s, u, v);
    float lum = dot(c, V(0.2126f, 0.7152f, 0.0722f));
    float sin_theta = std::sin(theta);
    if (p.sky_total_lum <= 0.0f || sin_theta <= 0.0f) return 1.0f / (4.0f * PI);
    return (lum * float(p.sky_width) * float(p.sky_height)) / (p.sky_total_lum * 2.0f * PI * PI);
}

struct LightSample { f3 dir; f3 color; float dist; float pdf; };

// shader.wgsl:231-254
LightSample sample_env_cdf(const ref_scene& s, const pt_scene_params& p, f2 u) {
    LightSample ls;
    uint32_t v_idx = binary_search_cdf(s, 0u, p.sky_height, u.y);
    uint32_t row_offset = p.sky_cdf_index + (v_idx * (p.sky_width + 1u));
    uint32_t u_idx = binary_search_cdf(s, row_offset, p.sky_width, u.x);
    float u_coord = (float(u_idx) + 0.5f) / float(p.sky_width);
    float v_coord = (float(v_idx) + 0.5f) / float(p.sky_height);
    float theta = PI * v_coord;
    float phi = 2.0f * PI * u_coord - PI;
    float sin_theta = std::sin(theta);
    ls.dir = V(sin_theta * std::cos(phi), std::cos(theta), sin_theta * std::sin(phi));
    ls.color = sample_sky_tex(s, u_coord, v_coord);
    ls.pdf = get_sky_pdf(s, p, ls.dir);
    ls.dist = INFINITY_;
    return ls;
}

// -------------------------------------------------------- intersection ------
// shader.wgsl:256-265
float intersect_aabb(f3 origin, f3 inv_dir, const float* mn, const float* mx) {
    f3 t0 = (ld3(mn) - origin) * inv_dir;
    f3 t1 = (ld3(mx) - origin) * inv_dir;
    f3 tmin = fmn(t0, t1), tmax = fmx(t0, t1);
    float t_near = max_component(tmin), t_far = min_component(tmax);
    if (t_near > t_far || t_far < 0.0f) return -1.0f;
    return t_near < 0.0f ? 0.0f : t_near;
}

// shader.wgsl:267-288
bool intersect_triangle(const Ray& ray, const pt_triangle& tri, float* hit_t, f3* bary) {
    f3 v0 = ld3(tri.v0);
    f3 edge1 = ld3(tri.v1) - v0;
    f3 edge2 = ld3(tri.v2) - v0;
    f3 h = cross(ray.direction, edge2);
    float a = dot(edge1, h);
    if (std::fabs(a) < 0.000001f) return false;
    float f = 1.0f / a;
    f3 s = ray.origin - v0;
    float u = f * dot(s, h);
    if (u < 0.0f || u > 1.0f) return false;
    f3 q = cross(s, edge1);
    float v = f * dot(ray.direction, q);
    if (v < 0.0f || u + v > 1.0f) return false;
    float t = f * dot(edge2, q);
    if (t > 0.000001f) {
        *hit_t = t;
        *bary = V(1.0f - u - v, u, v);
        return true;
    }
    return false;
}

// shader.wgsl:290-300
float hit_unit_sphere(const Ray& r) {
    float a = dot(r.direction, r.direction);
    float half_b = dot(r.origin, r.direction);
    float c = dot(r.origin, r.origin) - 1.0f;
    float disc = half_b * half_b - a * c;
    if (disc < 0.0f) return -1.0f;
    float sqrtd = std::sqrt(disc);
    float root = (-half_b - sqrtd) / a;
    if (root < 0.001f) root = (-half_b + sqrtd) / a;
    return root >= 0.001f ? root : -1.0f;
}

// shader.wgsl:302-312
float hit_unit_cube(const Ray& r) {
    f3 inv_dir = V(1.0f) / r.direction;
    f3 t0 = (V(-1.0f) - r.origin) * inv_dir;
    f3 t1 = (V(1.0f) - r.origin) * inv_dir;
    f3 tmin = fmn(t0, t1), tmax = fmx(t0, t1);
    float tnear = max_component(tmin), tfar = min_component(tmax);
    if (tnear < tfar && tfar > 0.0f) return tnear > 0.001f ? tnear : tfar;
    return -1.0f;
}

// shader.wgsl:314-321
float hit_plane(f3 normal, float d, const Ray& r) {
    float denom = dot(normal, r.direction);
    if (std::fabs(denom) > 1e-6f) {
        float t = (d - dot(normal, r.origin)) / denom;
        if (t > 0.001f) return t;
    }
    return -1.0f;
}

// shader.wgsl:323-379
bool hit_cylinder(const Ray& r, float r0, float r1, float* t_out, f3* n_out, f2* uv_out) {
    const float h = 1.0f;
    float dr = r1 - r0;
    float k = dr / h;
    float origin_eff_radius = r0 + k * r.origin.y;
    float A = r.direction.x * r.direction.x + r.direction.z * r.direction.z - k * k * r.direction.y * r.direction.y;
    float B = 2.0f * (r.origin.x * r.direction.x + r.origin.z * r.direction.z - origin_eff_radius * k * r.direction.y);
    float C = r.origin.x * r.origin.x + r.origin.z * r.origin.z - origin_eff_radius * origin_eff_radius;
    float t_min = 1e10f;
    bool hit = false;
    float disc = B * B - 4.0f * A * C;
    if (disc > 0.0f) {
        float sqrt_d = std::sqrt(disc);
        float ts[2] = {(-B - sqrt_d) / (2.0f * A), (-B + sqrt_d) / (2.0f * A)};
        for (int i = 0; i < 2; ++i) {
            float t = ts[i];
            float y = r.origin.y + t * r.direction.y;
            if (t > 0.001f && t < t_min && y >= 0.0f && y <= h) {
                t_min = t;
                f3 p = r.origin + t * r.direction;
                float slant = -k * (r0 + k * y);
                *n_out = normalize(V(p.x, slant, p.z));
                *uv_out = f2{std::atan2(p.z, p.x) / TWO_PI + 0.5f, y};
                hit = true;
            }
        }
    }
    const float caps[2] = {0.0f, 1.0f};
    const float radii[2] = {r0, r1};
    for (int i = 0; i < 2; ++i) {
        float py = caps[i];
        float t = (py - r.origin.y) / r.direction.y;
        if (t > 0.001f && t < t_min) {
            f3 p = r.origin + t * r.direction;
            if (p.x * p.x + p.z * p.z <= radii[i] * radii[i]) {
                t_min = t;
                *n_out = V(0.0f, float(i) * 2.0f - 1.0f, 0.0f);
                float m2 = fmx(r0, r1) * 2.0f;
                *uv_out = f2{p.x / m2 + 0.5f, p.z / m2 + 0.5f};
                hit = true;
            }
        }
    }
    if (hit) *t_out = t_min;
    return hit;
}

// shader.wgsl:381-471
float hit_torus(const Ray& r, float Ra, float ra) {
    float po = 1.0f;
    float Ra2 = Ra * Ra, ra2 = ra * ra;
    float m = dot(r.origin, r.origin);
    float n = dot(r.origin, r.direction);
    float bounds_radius = Ra + ra;
    float h_bound = n * n - m + bounds_radius * bounds_radius;
    if (h_bound < 0.0f) return -1.0f;
    float k = (m - ra2 - Ra2) / 2.0f;
    float k3 = n;
    float k2 = n * n + Ra2 * r.direction.y * r.direction.y + k;
    float k1 = k * n + Ra2 * r.origin.y * r.direction.y;
    float k0 = k * k + Ra2 * r.origin.y * r.origin.y - Ra2 * ra2;
    float K3 = k3, K2 = k2, K1 = k1, K0 = k0;
    if (std::fabs(k3 * (k3 * k3 - k2) + k1) < 0.01f) {
        po = -1.0f;
        float inv_k0 = 1.0f / k0;
        K1 = k3 * inv_k0;
        K2 = k2 * inv_k0;
        K3 = k1 * inv_k0;
        K0 = inv_k0;
    }
    float c2 = (2.0f * K2 - 3.0f * K3 * K3) / 3.0f;
    float c1 = 2.0f * (K3 * (K3 * K3 - K2) + K1);
    float c0 = (K3 * (K3 * (-3.0f * K3 * K3 + 4.0f * K2) - 8.0f * K1) + 4.0f * K0) / 3.0f;
    float Q = c2 * c2 + c0;
    float R = 3.0f * c0 * c2 - c2 * c2 * c2 - c1 * c1;
    float h = R * R - Q * Q * Q;
    float z;
    if (h < 0.0f) {
        float sQ = std::sqrt(Q);
        z = 2.0f * sQ * std::cos(std::acos(clampf(R / (sQ * Q), -1.0f, 1.0f)) / 3.0f);
    } else {
        float sQ = std::pow(std::sqrt(h) + std::fabs(R), 1.0f / 3.0f);
        z = signf(R) * std::fabs(sQ + Q / sQ);
    }
    z = c2 - z;
    float d1 = z - 3.0f * c2;
    float d2 = z * z - 3.0f * c0;
    if (std::fabs(d1) < 1.0e-4f) {
        if (d2 < 0.0f) return -1.0f;
        d2 = std::sqrt(d2);
    } else {
        if (d1 < 0.0f) return -1.0f;
        d1 = std::sqrt(d1 / 2.0f);
        d2 = c1 / d1;
    }
    float result = 1e20f;
    float h1 = d1 * d1 - z + d2;
    if (h1 > 0.0f) {
        h1 = std::sqrt(h1);
        float t1 = -d1 - h1 - K3; if (po < 0.0f) t1 = 2.0f / t1;
        float t2 = -d1 + h1 - K3; if (po < 0.0f) t2 = 2.0f / t2;
        if (t1 > 0.001f) result = t1;
        if (t2 > 0.001f) result = fmn(result, t2);
    }
    float h2 = d1 * d1 - z - d2;
    if (h2 > 0.0f) {
        h2 = std::sqrt(h2);
        float t3 = d1 - h2 - K3; if (po < 0.0f) t3 = 2.0f / t3;
        float t4 = d1 + h2 - K3; if (po < 0.0f) t4 = 2.0f / t4;
        if (t3 > 0.001f) result = fmn(result, t3);
        if (t4 > 0.001f) result = fmn(result, t4);
    }
    return result > 1e10f ? -1.0f : result;
}

// ------------------------------------------------------- closest hit -------
// shader.wgsl:473-542
void trace_mesh(const ref_scene& S, const Ray& ray_world, const float* inv, int material, uint32_t node_offset,
                uint32_t tri_offset, SurfaceHit* hit, int idx, Counters* c) {
    Ray rl;
    rl.origin = mul_point(inv, ray_world.origin);
    rl.direction = mul_dir(inv, ray_world.direction);
    f3 inv_dir = V(1.0f) / rl.direction;
    uint32_t stack[STACK_SIZE];
    int sp = 0;
    stack[sp++] = node_offset;
    float base_t = intersect_aabb(rl.origin, inv_dir, S.bvh[node_offset].aabb_min, S.bvh[node_offset].aabb_max);
    if (base_t < 0.0f || base_t >= hit->t) return;
    while (sp > 0) {
        hit->count++;
        c->pops++;
        uint32_t node_idx = stack[--sp];
        const pt_bvh_node& node = S.bvh[node_idx];
        if (node.count > 0u) {
            uint32_t start = tri_offset + node.next, end = start + node.count;
            for (uint32_t i = start; i < end; ++i) {
                const pt_triangle& tri = S.tris[i];
                float t_tri = 0.0f;
                f3 bary = V(0.f);
                c->tris++;
                if (intersect_triangle(rl, tri, &t_tri, &bary)) {
                    if (t_tri >= 0.0f && t_tri < hit->t) {
                        hit->t = t_tri;
                        hit->m_idx = material;
                        f3 ln = normalize(ld3(tri.n0) * bary.x + ld3(tri.n1) * bary.y + ld3(tri.n2) * bary.z);
                        hit->hit_n = normalize(mul_transpose_dir(inv, ln));
                        hit->hit_uv = f2{tri.uv0[0] * bary.x + tri.uv1[0] * bary.y + tri.uv2[0] * bary.z,
                                         tri.uv0[1] * bary.x + tri.uv1[1] * bary.y + tri.uv2[1] * bary.z};
                        hit->o_idx = idx;
                        hit->prim_id = int(i);
                        hit->plane_idx = -1;
                    }
                }
            }
        } else {
            c->inner++;
            uint32_t left_idx = node_idx + 1u, right_idx = node_offset + node.next;
            float left_t = intersect_aabb(rl.origin, inv_dir, S.bvh[left_idx].aabb_min, S.bvh[left_idx].aabb_max);
            float right_t = intersect_aabb(rl.origin, inv_dir, S.bvh[right_idx].aabb_min, S.bvh[right_idx].aabb_max);
            if (left_t < right_t) {
                if (right_t >= 0.0f && right_t < hit->t && sp < STACK_SIZE) stack[sp++] = right_idx;
                if (left_t >= 0.0f && left_t < hit->t && sp < STACK_SIZE) stack[sp++] = left_idx;
            } else {
                if (left_t >= 0.0f && left_t < hit->t && sp < STACK_SIZE) stack[sp++] = left_idx;
                if (right_t >= 0.0f && right_t < hit->t && sp < STACK_SIZE) stack[sp++] = right_idx;
            }
        }
    }
}

void set_object_hit(SurfaceHit* hit, float t, int material, int idx) {
    hit->t = t;
    hit->m_idx = material;
    hit->o_idx = idx;
    hit->prim_id = -1;
    hit->plane_idx = -1;
}

// shader.wgsl:544-575
void trace_sphere(const Ray& rw, const pt_object& o, SurfaceHit* hit, int idx) {
    Ray lr{mul_point(o.inv_matrix, rw.origin), mul_dir(o.inv_matrix, rw.direction)};
    float t = hit_unit_sphere(lr);
    if (t > 0.001f && t < hit->t) {
        set_object_hit(hit, t, o.material_idx, idx);
        f3 ln = lr.origin + lr.direction * t;
        float phi = std::atan2(ln.z, ln.x);
        float theta = std::asin(clampf(ln.y, -1.0f, 1.0f));
        hit->hit_uv = f2{0.5f + phi / TWO_PI, 0.5f + theta / PI};
        f3 lt = std::fabs(ln.y) > 0.9999f ? V(1.0f, 0.0f, 0.0f) : V(-ln.z, 0.0f, ln.x);
        lt = normalize(lt);
        hit->hit_n = normalize(mul_normal(o.inv_matrix, ln));
        hit->tangent = normalize(mul_normal(o.inv_matrix, lt));
        hit->bitangent = normalize(mul_normal(o.inv_matrix, cross(ln, lt)));
    }
}

// shader.wgsl:577-620
void trace_cube(const Ray& rw, const pt_object& o, SurfaceHit* hit, int idx) {
    Ray lr{mul_point(o.inv_matrix, rw.origin), mul_dir(o.inv_matrix, rw.direction)};
    float t = hit_unit_cube(lr);
    if (t > 0.001f && t < hit->t) {
        set_object_hit(hit, t, o.material_idx, idx);
        f3 lh = lr.origin + lr.direction * t;
        f3 d = V(std::fabs(lh.x), std::fabs(lh.y), std::fabs(lh.z));
        float max_d = max_component(d);
        f3 ln, lt, lb = V(0.0f, 1.0f, 0.0f);
        if (max_d == d.x) {
            float s = signf(lh.x);
            ln = V(s, 0.0f, 0.0f);
            lt = V(0.0f, 0.0f, -s);
            hit->hit_uv = f2{(-s * lh.z) * 0.5f + 0.5f, lh.y * 0.5f + 0.5f};
        } else if (max_d == d.y) {
            float s = signf(lh.y);
            ln = V(0.0f, s, 0.0f);
            lt = V(1.0f, 0.0f, 0.0f);
            lb = V(0.0f, 0.0f, s);
            hit->hit_uv = f2{lh.x * 0.5f + 0.5f, (s * lh.z) * 0.5f + 0.5f};
        } else {
            float s = signf(lh.z);
            ln = V(0.0f, 0.0f, s);
            lt = V(s, 0.0f, 0.0f);
            hit->hit_uv = f2{(s * lh.x) * 0.5f + 0.5f, lh.y * 0.5f + 0.5f};
        }
        hit->hit_n = normalize(mul_normal(o.inv_matrix, ln));
        hit->tangent = normalize(mul_normal(o.inv_matrix, lt));
        hit->bitangent = normalize(mul_normal(o.inv_matrix, lb));
    }
}

// shader.wgsl:622-636
void trace_plane(const Ray& rw, const pt_plane& p, SurfaceHit* hit, int plane_idx) {
    f3 n = ld3(p.normal);
    float t = hit_plane(n, p.d, rw);
    if (t > 0.001f && t < hit->t) {
        hit->t = t;
        hit->m_idx = p.material_idx;
        hit->hit_n = n;
        f3 hp = rw.origin + rw.direction * t;
        f3 tangent = V(1.0f, 0.0f, 0.0f);
        if (std::fabs(n.x) > 0.9f) tangent = V(0.0f, 0.0f, 1.0f);
        f3 tg = normalize(cross(n, tangent));
        f3 bt = normalize(cross(n, tg));
        hit->hit_uv = f2{dot(hp, tg), dot(hp, bt)};
        hit->tangent = tg;
        hit->bitangent = bt;
        hit->prim_id = -1;
        hit->plane_idx = plane_idx;
    }
}

// shader.wgsl:638-656
void trace_cylinder(const Ray& rw, const pt_object& o, SurfaceHit* hit, int idx) {
    Ray lr{mul_point(o.inv_matrix, rw.origin), mul_dir(o.inv_matrix, rw.direction)};
    float t; f3 n; f2 uv;
    if (hit_cylinder(lr, 1.0f, o.param0, &t, &n, &uv)) {
        if (t < hit->t) {
            set_object_hit(hit, t, o.material_idx, idx);
            hit->hit_uv = uv;
            hit->hit_n = normalize(mul_normal(o.inv_matrix, n));
        }
    }
}

// shader.wgsl:658-692
void trace_torus(const Ray& rw, const pt_object& o, SurfaceHit* hit, int idx) {
    Ray lr{mul_point(o.inv_matrix, rw.origin), mul_dir(o.inv_matrix, rw.direction)};
    float ray_scale = length(lr.direction);
    lr.direction = lr.direction / ray_scale;
    float t = hit_torus(lr, 1.0f, o.param0 * 1.0f);
    if (t > 0.0f) {
        float world_t = t / ray_scale;
        if (world_t < hit->t) {
            set_object_hit(hit, world_t, o.material_idx, idx);
            f3 p = lr.origin + lr.direction * t;
            f3 ring_p = normalize(V(p.x, 0.0f, p.z));
            f3 ln = normalize(p - ring_p);
            hit->hit_n = normalize(mul_mat3(o.inv_matrix, ln));
            float u = (std::atan2(p.z, p.x) / TWO_PI) + 0.5f;
            float v = (std::atan2(p.y, std::sqrt(p.x * p.x + p.z * p.z) - 1.0f) / TWO_PI) + 0.5f;
            hit->hit_uv = f2{u, v};
        }
    }
}

// shader.wgsl:694-752
void trace_tlas(const ref_scene& S, const Ray& ray, SurfaceHit* hit, Counters* c) {
    f3 inv_dir = V(1.0f) / ray.direction;
    uint32_t stack[STACK_SIZE];
    int sp = 0;
    stack[sp++] = 0;
    float base_t = intersect_aabb(ray.origin, inv_dir, S.tlas[0].aabb_min, S.tlas[0].aabb_max);
    if (base_t < 0.0f || base_t >= hit->t) return;
    while (sp > 0) {
        hit->count++;
        c->pops++;
        uint32_t node_idx = stack[--sp];
        const pt_bvh_node& node = S.tlas[node_idx];
        if (node.count > 0u) {
            uint32_t start = node.next, end = start + node.count;
            for (uint32_t i = start; i < end; ++i) {
                const pt_object& obj = S.objs[i];
                int idx = int(i);
                c->objs++;
                if (S.has(PT_HAS_MESHES) && obj.object_type == 0) {
                    uint32_t no, to;
                    std::memcpy(&no, &obj.param0, 4);
                    std::memcpy(&to, &obj.param1, 4);
                    trace_mesh(S, ray, obj.inv_matrix, obj.material_idx, no, to, hit, idx, c);
                } else if (S.has(PT_HAS_SPHERES) && obj.object_type == 1) {
                    trace_sphere(ray, obj, hit, idx);
                } else if (S.has(PT_HAS_CUBES) && obj.object_type == 2) {
                    trace_cube(ray, obj, hit, idx);
                } else if (S.has(PT_HAS_CYLINDERS) && obj.object_type == 3) {
                    trace_cylinder(ray, obj, hit, idx);
                } else if (S.has(PT_HAS_TORI) && obj.object_type == 4) {
                    trace_torus(ray, obj, hit, idx);
                }
            }
        } else {
            c->inner++;
            uint32_t left_idx = node_idx + 1u, right_idx = node.next;
            float left_t = intersect_aabb(ray.origin, inv_dir, S.tlas[left_idx].aabb_min, S.tlas[left_idx].aabb_max);
            float right_t = intersect_aabb(ray.origin, inv_dir, S.tlas[right_idx].aabb_min, S.tlas[right_idx].aabb_max);
            if (left_t < right_t) {
                if (right_t >= 0.0f && right_t < hit->t && sp < STACK_SIZE) stack[sp++] = right_idx;
                if (left_t >= 0.0f && left_t < hit->t && sp < STACK_SIZE) stack[sp++] = left_idx;
            } else {
                if (left_t >= 0.0f && left_t < hit->t && sp < STACK_SIZE) stack[sp++] = left_idx;
                if (right_t >= 0.0f && right_t < hit->t && sp < STACK_SIZE) stack[sp++] = right_idx;
            }
        }
    }
}

// shader.wgsl:754-774
SurfaceHit trace_scene(const ref_scene& S, const Ray& ray, Counters* c) {
    SurfaceHit hit = empty_hit();
    c->ext++;
    if (S.has(PT_HAS_PLANES))
        for (size_t i = 0; i < S.planes.size(); ++i) trace_plane(ray, S.planes[i], &hit, int(i));
    if (S.has(PT_HAS_SPHERES | PT_HAS_CUBES | PT_HAS_CYLINDERS | PT_HAS_TORI | PT_HAS_MESHES)) trace_tlas(S, ray, &hit, c);
    if (S.has(PT_HAS_LIST_MESHES))
        for (const auto& m : S.meshes)
            trace_mesh(S, ray, m.inv_matrix, m.material_idx, m.node_offset, m.tri_offset, &hit, -1, c);
    return hit;
}

// ------------------------------------------------------------ shadows -------
// With no albedo texture an occluder blocks outright (shader.wgsl:777-811),
// and textured materials are outside this build, so each shadow routine
// reduces to "is there an intersection inside (0.001 | 1e-6, target_dist)".

// shader.wgsl:813-872
// sample_texture (shader.wgsl:170-173): textureSampleLevel on the material
// texture array, level 0 of a 512 x 512 rgba8unorm layer, linear filtering,
// repeat addressing (tracer.js:1965-1968); the filter is restated with fp32
// weights (texel centres at uv * 512 - 0.5).  idx < 0 gives 1.
constexpr int kTexSize = 512;
struct f4 { float x, y, z, w; };
f4 sample_texture(const ref_scene& S, int idx, f2 uv) {
    if (idx < 0 || uint32_t(idx) >= S.tex_layers) return f4{1.f, 1.f, 1.f, 1.f};
    const float x = uv.x * float(kTexSize) - 0.5f, y = uv.y * float(kTexSize) - 0.5f;
    const float x0 = std::floor(x), y0 = std::floor(y);
    const float fx = x - x0, fy = y - y0;
    const long ix = long(x0), iy = long(y0);
    const uint8_t* L = S.tex.data() + size_t(idx) * kTexSize * kTexSize * 4;
    auto texel = [&](long tx, long ty, int c) {
        const size_t xx = size_t(tx & (kTexSize - 1)), yy = size_t(ty & (kTexSize - 1));
        return float(L[(yy * kTexSize + xx) * 4 + size_t(c)]) / 255.0f;
    };
    float o[4];
    for (int c = 0; c < 4; ++c) {
        const float t00 = texel(ix, iy, c), t10 = texel(ix + 1, iy, c);
        const float t01 = texel(ix, iy + 1, c), t11 = texel(ix + 1, iy + 1, c);
        const float a = t00 + (t10 - t00) * fx, b = t01 + (t11 - t01) * fx;
        o[c] = a + (b - a) * fy;
    }
    return f4{o[0], o[1], o[2], o[3]};
}

// evaluate_shadow_attenuation (shader.wgsl:777-811): a material without an
// albedo texture stops the light (true); with one, the shadow is scaled by
// (1 - alpha) of the texel at uv * uv_scale and stopped above alpha 0.9999.
bool evaluate_shadow_attenuation(const ref_scene& S, const pt_material& mat, f2 uv, f3* shadow) {
    if (mat.albedo_idx < 0) return true;
    f4 t = sample_texture(S, mat.albedo_idx, f2{uv.x * mat.uv_scale[0], uv.y * mat.uv_scale[1]});
    *shadow = *shadow * (1.0f - t.w);
    return t.w > 0.9999f;
}

// shader.wgsl:893-901 (sphere), 904-922 (get_cube_uv), 958-962
// (get_frustum_uv), 1090-1094 (get_torus_uv)
f2 sphere_shadow_uv(f3 p) {
    return f2{0.5f + std::atan2(p.z, p.x) / TWO_PI, 0.5f + std::asin(clampf(p.y, -1.0f, 1.0f)) / PI};
}
f2 get_cube_uv(f3 p) {
    f3 d = V(std::fabs(p.x), std::fabs(p.y), std::fabs(p.z));
    float max_d = max_component(d);
    if (max_d == d.x) { float s = signf(p.x); return f2{(-s * p.z) * 0.5f + 0.5f, p.y * 0.5f + 0.5f}; }
    if (max_d == d.y) { float s = signf(p.y); return f2{p.x * 0.5f + 0.5f, (s * p.z) * 0.5f + 0.5f}; }
    float s = signf(p.z);
    return f2{(s * p.x) * 0.5f + 0.5f, p.y * 0.5f + 0.5f};
}
f2 get_frustum_uv(f3 p, float r0, float r1) {
    if (p.y > 0.999f) return f2{p.x / (r1 * 2.0f) + 0.5f, p.z / (r1 * 2.0f) + 0.5f};
    if (p.y < 0.001f) return f2{p.x / (r0 * 2.0f) + 0.5f, p.z / (r0 * 2.0f) + 0.5f};
    return f2{std::atan2(p.z, p.x) / TWO_PI + 0.5f, p.y};
}
f2 get_torus_uv(f3 p) {
    return f2{(std::atan2(p.z, p.x) / TWO_PI) + 0.5f,
              (std::atan2(p.y, std::sqrt(p.x * p.x + p.z * p.z) - 1.0f) / TWO_PI) + 0.5f};
}

bool trace_mesh_shadow(const ref_scene& S, const Ray& rw, const float* inv, uint32_t node_offset, uint32_t tri_offset,
                       int material, float target_dist, f3* shadow, Counters* c) {
    Ray rl{mul_point(inv, rw.origin), mul_dir(inv, rw.direction)};
    f3 inv_dir = V(1.0f) / rl.direction;
    uint32_t stack[STACK_SIZE];
    int sp = 0;
    stack[sp++] = node_offset;
    float base_t = intersect_aabb(rl.origin, inv_dir, S.bvh[node_offset].aabb_min, S.bvh[node_offset].aabb_max);
    if (base_t < 0.0f || base_t >= target_dist) return false;
    while (sp > 0) {
        c->pops++;
        uint32_t node_idx = stack[--sp];
        const pt_bvh_node& node = S.bvh[node_idx];
        if (node.count > 0u) {
            uint32_t start = tri_offset + node.next, end = start + node.count;
            for (uint32_t i = start; i < end; ++i) {
                float t_tri = 0.0f;
                f3 bary;
                c->tris++;
                const pt_triangle& tri = S.tris[i];
                if (intersect_triangle(rl, tri, &t_tri, &bary))
                    if (t_tri >= 0.0f && t_tri < target_dist) {
                        f2 uv{tri.uv0[0] * bary.x + tri.uv1[0] * bary.y + tri.uv2[0] * bary.z,
                              tri.uv0[1] * bary.x + tri.uv1[1] * bary.y + tri.uv2[1] * bary.z};
                        if (evaluate_shadow_attenuation(S, S.mats[material], uv, shadow)) return true;
                    }
            }
        } else {
            c->inner++;
            uint32_t left_idx = node_idx + 1u, right_idx = node_offset + node.next;
            float left_t = intersect_aabb(rl.origin, inv_dir, S.bvh[left_idx].aabb_min, S.bvh[left_idx].aabb_max);
            float right_t = intersect_aabb(rl.origin, inv_dir, S.bvh[right_idx].aabb_min, S.bvh[right_idx].aabb_max);
            if (left_t < right_t) {
                if (right_t >= 0.0f && right_t < target_dist && sp < STACK_SIZE) stack[sp++] = right_idx;
                if (left_t >= 0.0f && left_t < target_dist && sp < STACK_SIZE) stack[sp++] = left_idx;
            } else {
                if (left_t >= 0.0f && left_t < target_dist && sp < STACK_SIZE) stack[sp++] = left_idx;
                if (right_t >= 0.0f && right_t < target_dist && sp < STACK_SIZE) stack[sp++] = right_idx;
            }
        }
    }
    return false;
}

// shader.wgsl:874-902
bool trace_sphere_shadow(const ref_scene& S, const Ray& rw, const pt_object& o, float target_dist, f3* shadow) {
    Ray lr{mul_point(o.inv_matrix, rw.origin), mul_dir(o.inv_matrix, rw.direction)};
    float a = dot(lr.direction, lr.direction);
    float half_b = dot(lr.origin, lr.direction);
    float cc = dot(lr.origin, lr.origin) - 1.0f;
    float disc = half_b * half_b - a * cc;
    if (disc < 0.0f) return false;
    float sqrtd = std::sqrt(disc);
    float t1 = (-half_b - sqrtd) / a;
    float t2 = (-half_b + sqrtd) / a;
    float h1 = fmx(0.001f, t1);
    float h2 = fmn(target_dist, t2);
    if (!(h1 < h2)) return false;
    const pt_material& mat = S.mats[o.material_idx];
    if (evaluate_shadow_attenuation(S, mat, sphere_shadow_uv(lr.origin + lr.direction * h1), shadow)) return true;
    return evaluate_shadow_attenuation(S, mat, sphere_shadow_uv(lr.origin + lr.direction * h2), shadow);
}

// shader.wgsl:924-953
bool trace_cube_shadow(const ref_scene& S, const Ray& rw, const pt_object& o, float target_dist, f3* shadow) {
    Ray lr{mul_point(o.inv_matrix, rw.origin), mul_dir(o.inv_matrix, rw.direction)};
    f3 inv_dir = V(1.0f) / lr.direction;
    f3 tn = (V(-1.0f) - lr.origin) * inv_dir;
    f3 tf = (V(1.0f) - lr.origin) * inv_dir;
    f3 t_min = fmn(tn, tf), t_max = fmx(tn, tf);
    float t1 = max_component(t_min), t2 = min_component(t_max);
    float h1 = fmx(0.001f, t1), h2 = fmn(target_dist, t2);
    if (!(h1 < h2)) return false;
    const pt_material& mat = S.mats[o.material_idx];
    if (evaluate_shadow_attenuation(S, mat, get_cube_uv(lr.origin + lr.direction * h1), shadow)) return true;
    return evaluate_shadow_attenuation(S, mat, get_cube_uv(lr.origin + lr.direction * h2), shadow);
}

// shader.wgsl:961-1016
bool trace_cylinder_shadow(const ref_scene& S, const Ray& rw, const pt_object& o, float target_dist, f3* shadow) {
    Ray r{mul_point(o.inv_matrix, rw.origin), mul_dir(o.inv_matrix, rw.direction)};
    float k = (o.param0 - 1.0f) / 1.0f;
    float origin_eff_r = 1.0f + k * r.origin.y;
    float A = r.direction.x * r.direction.x + r.direction.z * r.direction.z - k * k * r.direction.y * r.direction.y;
    float B = 2.0f * (r.origin.x * r.direction.x + r.origin.z * r.direction.z - origin_eff_r * k * r.direction.y);
    float C = r.origin.x * r.origin.x + r.origin.z * r.origin.z - origin_eff_r * origin_eff_r;
    float h1 = 1e10f, h2 = -1e10f;
    uint32_t hit_count = 0;
    float disc = B * B - 4.0f * A * C;
    if (disc >= 0.0f) {
        float sqrt_d = std::sqrt(disc);
        float ts[2] = {(-B - sqrt_d) / (2.0f * A), (-B + sqrt_d) / (2.0f * A)};
        for (int i = 0; i < 2; ++i) {
            float t = ts[i];
            float y = r.origin.y + t * r.direction.y;
            if (y >= 0.0f && y <= 1.0f) { h1 = fmn(h1, t); h2 = fmx(h2, t); hit_count++; }
        }
    }
    const float caps[2] = {0.0f, 1.0f};
    const float radii[2] = {1.0f, o.param0};
    for (int i = 0; i < 2; ++i) {
        float t = (caps[i] - r.origin.y) / r.direction.y;
        f3 p = r.origin + t * r.direction;
        if (p.x * p.x + p.z * p.z <= radii[i] * radii[i]) { h1 = fmn(h1, t); h2 = fmx(h2, t); hit_count++; }
    }
    if (hit_count >= 2u) {
        float fh1 = fmx(0.001f, h1), fh2 = fmn(target_dist, h2);
        if (fh1 < fh2) {
            const pt_material& mat = S.mats[o.material_idx];
            if (evaluate_shadow_attenuation(S, mat, get_frustum_uv(r.origin + r.direction * fh1, 1.0f, o.param0), shadow))
                return true;
            return evaluate_shadow_attenuation(S, mat, get_frustum_uv(r.origin + r.direction * fh2, 1.0f, o.param0),
                                               shadow);
        }
    }
    return false;
}

// shader.wgsl:1019-1036
float solve_cubic_real(float a, float b, float c) {
    float q = (a * a - 3.0f * b) / 9.0f;
    float r = (2.0f * a * a * a - 9.0f * a * b + 27.0f * c) / 54.0f;
    float r2 = r * r, q3 = q * q * q;
    if (r2 < q3) {
        float theta = std::acos(clampf(r / std::sqrt(q3), -1.0f, 1.0f));
        return -2.0f * std::sqrt(q) * std::cos(theta / 3.0f) - a / 3.0f;
    }
    float aa = -signf(r) * std::pow(std::fabs(r) + std::sqrt(r2 - q3), 1.0f / 3.0f);
    float bb = 0.0f;
    if (aa != 0.0f) bb = q / aa;
    return (aa + bb) - a / 3.0f;
}

// shader.wgsl:1039-1087
void solve_quartic_real(float a, float b, float c, float d, float roots[4]) {
    roots[0] = roots[1] = roots[2] = roots[3] = -1.0f;
    uint32_t count = 0;
    float a2 = a * a;
    float y = solve_cubic_real(-b, a * c - 4.0f * d, 4.0f * b * d - a2 * d - c * c);
    float R2 = 0.25f * a2 - b + y;
    if (R2 < 0.0f) return;
    float R = std::sqrt(R2);
    float D2, E2;
    if (R < 1e-6f) {
        D2 = 0.75f * a2 - 2.0f * b + 2.0f * std::sqrt(y * y - 4.0f * d);
        E2 = 0.75f * a2 - 2.0f * b - 2.0f * std::sqrt(y * y - 4.0f * d);
    } else {
        D2 = 0.75f * a2 - R2 - 2.0f * b + (4.0f * a * b - 8.0f * c - a2 * a) / (4.0f * R);
        E2 = 0.75f * a2 - R2 - 2.0f * b - (4.0f * a * b - 8.0f * c - a2 * a) / (4.0f * R);
    }
    if (D2 >= 0.0f) {
        float D = std::sqrt(D2);
        roots[0] = -0.25f * a + 0.5f * R + 0.5f * D;
        roots[1] = -0.25f * a + 0.5f * R - 0.5f * D;
        count += 2u;
    }
    if (E2 >= 0.0f) {
        float E = std::sqrt(E2);
        roots[count] = -0.25f * a - 0.5f * R + 0.5f * E;
        roots[count + 1u] = -0.25f * a - 0.5f * R - 0.5f * E;
        count += 2u;
    }
    for (uint32_t i = 0; i < 3u; ++i)
        for (uint32_t j = i + 1u; j < 4u; ++j)
            if (roots[i] > roots[j]) std::swap(roots[i], roots[j]);
}

// shader.wgsl:1096-1132
bool trace_torus_shadow(const ref_scene& S, const Ray& rw, const pt_object& o, float target_dist, f3* shadow) {
    Ray l{mul_point(o.inv_matrix, rw.origin), mul_dir(o.inv_matrix, rw.direction)};
    float ray_scale = length(l.direction);
    l.direction = l.direction / ray_scale;
    float R = 1.0f, r = o.param0;
    float K = dot(l.origin, l.origin) + R * R - r * r;
    float G = dot(l.origin, l.direction);
    float B = 4.0f * G;
    float C = 2.0f * K + 4.0f * G * G - 4.0f * R * R * (1.0f - l.direction.y * l.direction.y);
    float D = 4.0f * K * G - 8.0f * R * R * l.origin.y * l.direction.y;
    float E = K * K - 4.0f * R * R * (l.origin.x * l.origin.x + l.origin.z * l.origin.z);
    float roots[4];
    solve_quartic_real(B, C, D, E, roots);
    for (int i = 0; i < 4; i += 2) {
        float t_near = roots[i], t_far = roots[i + 1];
        if (t_near < 0.0f || t_near > target_dist * ray_scale) continue;
        float h1 = fmx(0.001f, t_near) / ray_scale;
        float h2 = fmn(target_dist * ray_scale, t_far) / ray_scale;
        if (h1 < h2) {
            const pt_material& mat = S.mats[o.material_idx];
            if (evaluate_shadow_attenuation(S, mat, get_torus_uv(l.origin + l.direction * (h1 * ray_scale)), shadow))
                return true;
            if (evaluate_shadow_attenuation(S, mat, get_torus_uv(l.origin + l.direction * (h2 * ray_scale)), shadow))
                return true;
        }
    }
    return false;
}

// shader.wgsl:1134-1184
bool trace_tlas_shadow(const ref_scene& S, const Ray& ray, int target_idx, float target_dist, f3* shadow, Counters* c) {
    f3 inv_dir = V(1.0f) / ray.direction;
    uint32_t stack[STACK_SIZE];
    int sp = 0;
    stack[sp++] = 0u;
    while (sp > 0) {
        c->pops++;
        uint32_t node_idx = stack[--sp];
        const pt_bvh_node& node = S.tlas[node_idx];
        if (node.count > 0u) {
            uint32_t start = node.next, end = start + node.count;
            for (uint32_t i = start; i < end; ++i) {
                if (int(i) == target_idx) continue;
                const pt_object& obj = S.objs[i];
                c->objs++;
                if (S.has(PT_HAS_MESHES) && obj.object_type == 0) {
                    uint32_t no, to;
                    std::memcpy(&no, &obj.param0, 4);
                    std::memcpy(&to, &obj.param1, 4);
                    if (trace_mesh_shadow(S, ray, obj.inv_matrix, no, to, obj.material_idx, target_dist, shadow, c))
                        return true;
                } else if (S.has(PT_HAS_SPHERES) && obj.object_type == 1) {
                    if (trace_sphere_shadow(S, ray, obj, target_dist, shadow)) return true;
                } else if (S.has(PT_HAS_CUBES) && obj.object_type == 2) {
                    if (trace_cube_shadow(S, ray, obj, target_dist, shadow)) return true;
                } else if (S.has(PT_HAS_CYLINDERS) && obj.object_type == 3) {
                    if (trace_cylinder_shadow(S, ray, obj, target_dist, shadow)) return true;
                } else if (S.has(PT_HAS_TORI) && obj.object_type == 4) {
                    if (trace_torus_shadow(S, ray, obj, target_dist, shadow)) return true;
                }
            }
        } else {
            c->inner++;
            uint32_t left_idx = node_idx + 1u, right_idx = node.next;
            float left_t = intersect_aabb(ray.origin, inv_dir, S.tlas[left_idx].aabb_min, S.tlas[left_idx].aabb_max);
            float right_t = intersect_aabb(ray.origin, inv_dir, S.tlas[right_idx].aabb_min, S.tlas[right_idx].aabb_max);
            if (left_t < right_t) {
                if (right_t >= 0.0f && right_t < target_dist && sp < STACK_SIZE) stack[sp++] = right_idx;
                if (left_t >= 0.0f && left_t < target_dist && sp < STACK_SIZE) stack[sp++] = left_idx;
            } else {
                if (left_t >= 0.0f && left_t < target_dist && sp < STACK_SIZE) stack[sp++] = left_idx;
                if (right_t >= 0.0f && right_t < target_dist && sp < STACK_SIZE) stack[sp++] = right_idx;
            }
        }
    }
    return false;
}

// shader.wgsl:1186-1204
bool trace_plane_shadow(const ref_scene& S, const Ray& ray, const pt_plane& p, float target_dist, f3* shadow) {
    f3 n = ld3(p.normal);
    float denom = dot(n, ray.direction);
    if (std::fabs(denom) > 1e-6f) {
        float t = (p.d - dot(n, ray.origin)) / denom;
        if (t > 0.001f && t < target_dist) {
            f3 hit_pos = ray.origin + ray.direction * t;
            f3 tangent = std::fabs(n.x) > 0.9999f ? V(0.0f, 0.0f, 1.0f) : V(1.0f, 0.0f, 0.0f);
            tangent = normalize(cross(n, tangent));
            f3 bitangent = normalize(cross(n, tangent));
            f2 uv{dot(hit_pos, tangent), dot(hit_pos, bitangent)};
            return evaluate_shadow_attenuation(S, S.mats[p.material_idx], uv, shadow);
        }
    }
    return false;
}

// shader.wgsl:1206-1229
f3 trace_scene_shadow(const ref_scene& S, const Ray& ray, int target_idx, float target_dist, Counters* c) {
    c->shadow++;
    f3 shadow = V(1.0f);
    if (S.has(PT_HAS_PLANES))
        for (const auto& p : S.planes)
            if (trace_plane_shadow(S, ray, p, target_dist, &shadow)) return V(0.0f);
    if (S.has(PT_HAS_SPHERES | PT_HAS_CUBES | PT_HAS_CYLINDERS | PT_HAS_TORI | PT_HAS_MESHES))
        if (trace_tlas_shadow(S, ray, target_idx, target_dist, &shadow, c)) return V(0.0f);
    if (S.has(PT_HAS_LIST_MESHES))
        for (const auto& m : S.meshes)
            if (trace_mesh_shadow(S, ray, m.inv_matrix, m.node_offset, m.tri_offset, m.material_idx, target_dist,
                                  &shadow, c))
                return V(0.0f);
    return shadow;
}

// ------------------------------------------------------------ shading -------
struct Rng {
    uint32_t state;
    // shader.wgsl:147-154: the output permutation of the current state, then
    // one LCG step.
    uint32_t next_word() {
        uint32_t s = state;
        state = s * 747796405u + 2891336453u;
        uint32_t word = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
        word = (word >> 22u) ^ word;
        return word;
    }
    float next() { return float(next_word()) * 2.3283064365386963e-10f; }
};

// shader.wgsl:2280-2287 (rng == nullptr: gbuf_main's pixel centre,
// shader.wgsl:2517-2522): seed the pixel's stream, draw the warm-up value and
// the two jitter values, and build the camera ray through the jittered point.
Ray camera_ray(const pt_scene_params& P, uint32_t x, uint32_t y, Rng* rng, f2* jitter) {
    float jx = 0.5f, jy = 0.5f;
    if (rng) {
        const uint32_t idx = y * P.width + x;
        rng->state = (idx * 1973u + P.sample_number * 9277u + P.seed * 26699u) | 1u;
        rng->next();
        jx = rng->next();
        jy = rng->next();
    }
    if (jitter) *jitter = f2{jx, jy};
    f2 screen_uv{(float(x) + jx) / float(P.width), (float(y) + jy) / float(P.height)};
    f3 r00 = ld3(P.ray00), r10 = ld3(P.ray10), r01 = ld3(P.ray01), r11 = ld3(P.ray11);
    f3 ray_dir = normalize(mix3(mix3(r00, r10, screen_uv.x), mix3(r01, r11, screen_uv.x), 1.0f - screen_uv.y));
    return Ray{ld3(P.eye), ray_dir};
}

struct SurfaceContext { f3 normal, surface_normal, albedo; float roughness, metallic; f3 emittance; float alpha; };


// shader.wgsl:1393-1440 at texture coordinates uv (hit_uv * uv_scale, moved
// by parallax mapping): the normal map in the frame (tangent, bitangent,
// hit_n) with x and y scaled by height_params.x, albedo (its alpha too),
// roughness, metallic and emissive maps.
SurfaceContext get_surface_context(const ref_scene& S, const SurfaceHit& hit, const pt_material& mat, f2 uv) {
    SurfaceContext ctx;
    ctx.normal = hit.hit_n;
    ctx.surface_normal = hit.hit_n;
    ctx.albedo = ld3(mat.color);
    ctx.alpha = 1.0f;
    ctx.metallic = mat.metallic;
    ctx.roughness = mat.roughness;
    ctx.emittance = ld3(mat.emittance);
    if (S.tex_layers == 0) return ctx;
    if (mat.normal_idx >= 0) {
        const f4 t = sample_texture(S, mat.normal_idx, uv);
        const float k = mat.height_params[0];
        f3 nm = normalize(V((t.x * 2.0f - 1.0f) * k, (t.y * 2.0f - 1.0f) * k, t.z * 2.0f - 1.0f));
        ctx.normal = normalize(hit.tangent * nm.x + hit.bitangent * nm.y + hit.hit_n * nm.z);
    }
    if (mat.albedo_idx >= 0) {
        const f4 t = sample_texture(S, mat.albedo_idx, uv);
        ctx.alpha = t.w;
        ctx.albedo = V(ctx.albedo.x * std::pow(t.x, 2.2f), ctx.albedo.y * std::pow(t.y, 2.2f),
                       ctx.albedo.z * std::pow(t.z, 2.2f));
    }
    if (mat.roughness_idx >= 0) ctx.roughness *= sample_texture(S, mat.roughness_idx, uv).y;
    if (mat.metallic_idx >= 0) ctx.metallic *= sample_texture(S, mat.metallic_idx, uv).z;
    if (mat.emissive_idx >= 0) {
        const f4 t = sample_texture(S, mat.emissive_idx, uv);
        ctx.emittance = V(ctx.emittance.x * std::pow(t.x, 2.2f), ctx.emittance.y * std::pow(t.y, 2.2f),
                          ctx.emittance.z * std::pow(t.z, 2.2f));
    }
    return ctx;
}

f2 material_uv(const SurfaceHit& hit, const pt_material& mat) {
    return f2{hit.hit_uv.x * mat.uv_scale[0], hit.hit_uv.y * mat.uv_scale[1]};
}

struct PomResult { f2 uv; float height; bool hit; };

// Depth under the surface the height map gives at uv (height_params.y < 0:
// a height map, white high; else a depth map, white deep).
float pom_depth(const ref_scene& S, const pt_material& mat, f2 uv) {
    float s = sample_texture(S, mat.height_idx, uv).x;
    return mat.height_params[1] < 0.0f ? 1.0f - s : s;
}

// shader.wgsl:1238-1278 with start depth 0.  The early return leaves the
// zero-initialised result: uv (0, 0), height 0.
PomResult calculate_pom(const ref_scene& S, f2 uv0, f3 view_ts, const pt_material& mat) {
    PomResult res{f2{0.0f, 0.0f}, 0.0f, false};
    float scale = std::fabs(mat.height_params[1]);
    float num_layers = mat.height_params[2];
    float layer_depth = 1.0f / num_layers;
    float cur_depth = 0.0f;
    f2 P{-(view_ts.x / view_ts.z) * scale, -(view_ts.y / view_ts.z) * scale};
    f2 delta{P.x / num_layers, P.y / num_layers};
    f2 cur{uv0.x - P.x * mat.height_params[3], uv0.y - P.y * mat.height_params[3]};
    float map_depth = pom_depth(S, mat, cur);
    for (int i = 0; i < 64; ++i) {
        if (float(i) >= num_layers || cur_depth >= map_depth) break;
        if (cur_depth > 1.0f || cur_depth < 0.0f) return res;
        cur = f2{cur.x + delta.x, cur.y + delta.y};
        map_depth = pom_depth(S, mat, cur);
        cur_depth += layer_depth;
    }
    f2 prev{cur.x - delta.x, cur.y - delta.y};
    float prev_depth = pom_depth(S, mat, prev);
    float after = map_depth - cur_depth;
    float before = prev_depth - cur_depth + layer_depth;
    float w = after / (after - before);
    res.uv = f2{cur.x * (1.0f - w) + prev.x * w, cur.y * (1.0f - w) + prev.y * w};
    res.height = cur_depth * (1.0f - w) + (cur_depth - layer_depth) * w;
    res.hit = true;
    return res;
}

// shader.wgsl:1280-1329
PomResult calculate_shadow_pom(const ref_scene& S, f2 uv, float height, f3 light_ts, const pt_material& mat) {
    PomResult res{uv, height, false};
    if (light_ts.z <= 0.0f) {
        res.hit = true;
        return res;
    }
    float scale = std::fabs(mat.height_params[1]);
    float num_layers = mat.height_params[2];
    float layer_depth = 1.0f / num_layers;
    f2 p{(light_ts.x / light_ts.z) * scale * layer_depth, (light_ts.y / light_ts.z) * scale * layer_depth};
    f2 su = uv;
    float sd = height;
    for (int i = 0; i < 32; ++i) {
        if (float(i) / 2.0f >= num_layers || sd <= 0.0f) break;
        su = f2{su.x + p.x, su.y + p.y};
        sd -= layer_depth;
        float md = pom_depth(S, mat, su);
        if (md < sd) {
            res.hit = true;
            res.uv = su;
            res.height = md;
            return res;
        }
    }
    return res;
}

// transpose(tbn) * v
f3 tangent_space(const SurfaceHit& hit, f3 v) { return V(dot(hit.tangent, v), dot(hit.bitangent, v), dot(hit.hit_n, v)); }

// shader.wgsl:1463-1468
void build_tangent_space(f3 N, f3* T, f3* B) {
    f3 a = std::fabs(N.y) > 0.99999f ? V(0.f, 0.f, 1.f) : V(0.f, 1.f, 0.f);
    *T = normalize(cross(N, a));
    *B = cross(*T, N);
}

// shader.wgsl:1473-1499
float light_sphere_pdf(const pt_light& light, const pt_object& sphere, f3 hit_pos, f3 dir) {
    f3 d = ld3(sphere.world_position) - hit_pos;
    float d2 = dot(d, d);
    float r2 = light.scale * light.scale;
    if (d2 <= r2) return 1.0f / (4.0f * PI);
    float tc = dot(d, dir);
    if (tc <= 0.0f) return 0.0f;
    float d_perp_sq = d2 - tc * tc;
    if (d_perp_sq > r2) return 0.0f;
    float sin_theta_max_sq = r2 / d2;
    float cos_theta_max = std::sqrt(fmx(0.0f, 1.0f - sin_theta_max_sq));
    float solid_angle = TWO_PI * (1.0f - cos_theta_max);
    return 1.0f / fmx(solid_angle, 0.0001f);
}

// shader.wgsl:1501-1549
LightSample sample_light_sphere(const ref_scene& S, const pt_light& light, const pt_object& sphere, f3 hit_pos, f2 xi) {
    f3 d = ld3(sphere.world_position) - hit_pos;
    float d2 = dot(d, d);
    float r2 = light.scale * light.scale;
    LightSample ls;
    ls.color = ld3(S.mats[sphere.material_idx].emittance);
    if (d2 <= r2) {
        float z = 1.0f - 2.0f * xi.x;
        float sin_theta = std::sqrt(fmx(0.0f, 1.0f - z * z));
        float phi = TWO_PI * xi.y;
        ls.dir = V(sin_theta * std::cos(phi), sin_theta * std::sin(phi), z);
        ls.dist = 0.0f;
        ls.pdf = 1.0f / (4.0f * PI);
        return ls;
    }
    float sin_theta_max_sq = r2 / d2;
    float cos_theta_max = std::sqrt(fmx(0.0f, 1.0f - sin_theta_max_sq));
    float cos_theta = 1.0f - xi.x * (1.0f - cos_theta_max);
    float sin_theta = std::sqrt(fmx(0.0f, 1.0f - cos_theta * cos_theta));
    float phi = TWO_PI * xi.y;
    f3 local_dir = V(sin_theta * std::cos(phi), sin_theta * std::sin(phi), cos_theta);
    f3 w = normalize(d);
    f3 T, B;
    build_tangent_space(w, &T, &B);
    ls.dir = normalize(T * local_dir.x + B * local_dir.y + w * local_dir.z);
    float tc = dot(d, ls.dir);
    float d_perp_sq = d2 - tc * tc;
    float t_c = std::sqrt(fmx(0.0f, r2 - d_perp_sq));
    ls.dist = tc - t_c;
    float solid_angle = TWO_PI * (1.0f - cos_theta_max);
    ls.pdf = 1.0f / fmx(solid_angle, 0.0001f);
    return ls;
}

// Visible-face areas shared by light_box_pdf and sample_light_box.
void box_face_areas(const pt_light& light, f3 hit_pos, f3 s[3], float a[3]) {
    f3 v1 = ld3(light.matrix), v2 = ld3(light.matrix + 4), v3 = ld3(light.matrix + 8), cen = ld3(light.matrix + 12);
    f3 d = cen - hit_pos;
    float s1 = dot(d, v1) < 0.0f ? 1.0f : -1.0f;
    float s2 = dot(d, v2) < 0.0f ? 1.0f : -1.0f;
    float s3 = dot(d, v3) < 0.0f ? 1.0f : -1.0f;
    f3 nd = normalize(-d);
    a[0] = length(cross(v2 * 2.0f, v3 * 2.0f)) * fmx(dot(nd, normalize(v1 * s1)), 0.01f);
    a[1] = length(cross(v1 * 2.0f, v3 * 2.0f)) * fmx(dot(nd, normalize(v2 * s2)), 0.01f);
    a[2] = length(cross(v1 * 2.0f, v2 * 2.0f)) * fmx(dot(nd, normalize(v3 * s3)), 0.01f);
    s[0] = V(s1); s[1] = V(s2); s[2] = V(s3);
}

// shader.wgsl:1551-1573
float light_box_pdf(const pt_light& light, f3 hit_pos, f3 ray_dir, const SurfaceHit& hit) {
    f3 s[3]; float a[3];
    box_face_areas(light, hit_pos, s, a);
    float total_area = a[0] + a[1] + a[2];
    float cos_l = fmx(dot(-ray_dir, hit.hit_n), 1e-6f);
    float dist_sq = hit.t * hit.t;
    return dist_sq / (total_area * cos_l);
}

// shader.wgsl:1575-1624
LightSample sample_light_box(const ref_scene& S, const pt_light& light, const pt_object& obj, f3 hit_pos, f2 xi, Rng& rng) {
    f3 v1 = ld3(light.matrix), v2 = ld3(light.matrix + 4), v3 = ld3(light.matrix + 8), cen = ld3(light.matrix + 12);
    f3 s[3]; float a[3];
    box_face_areas(light, hit_pos, s, a);
    float total_area = a[0] + a[1] + a[2];
    f3 p = cen, n;
    float u = xi.x * 2.0f - 1.0f, v = xi.y * 2.0f - 1.0f;
    float pick = rng.next() * total_area;
    // `p += a + b + c` adds the whole right-hand side to p
    if (pick < a[0]) {
        p = p + (v1 * s[0].x + v2 * u + v3 * v);
        n = normalize(v1 * s[0].x);
    } else if (pick < a[0] + a[1]) {
        p = p + (v2 * s[1].x + v1 * u + v3 * v);
        n = normalize(v2 * s[1].x);
    } else {
        p = p + (v3 * s[2].x + v1 * u + v2 * v);
        n = normalize(v3 * s[2].x);
    }
    LightSample ls;
    f3 delta = p - hit_pos;
    float dist_sq = dot(delta, delta);
    ls.dist = std::sqrt(dist_sq);
    ls.dir = delta / ls.dist;
    float cos_l = fmx(dot(-ls.dir, n), 1e-6f);
    ls.pdf = dist_sq / (total_area * cos_l);
    ls.color = ld3(S.mats[obj.material_idx].emittance);
    return ls;
}

// shader.wgsl:1626-1646
float get_light_pdf(const ref_scene& S, const pt_light& light, const pt_object& obj, f3 hit_pos, f3 dir, const SurfaceHit& hit) {
    if (S.has(PT_HAS_SPHERES) && obj.object_type == 1) return light_sphere_pdf(light, obj, hit_pos, dir);
    if (S.has(PT_HAS_CUBES) && obj.object_type == 2) return light_box_pdf(light, hit_pos, dir, hit);
    return 0.0f;
}

LightSample sample_light(const ref_scene& S, const pt_light& light, const pt_object& obj, f3 hit_pos, Rng& rng) {
    LightSample samp;
    samp.dir = V(0.f); samp.color = V(0.f); samp.dist = 0.f; samp.pdf = 0.f;  // WGSL zero-initialises
    if (S.has(PT_HAS_SPHERES) && obj.object_type == 1) {
        float x = rng.next(); float y = rng.next();
        samp = sample_light_sphere(S, light, obj, hit_pos, f2{x, y});
    } else if (S.has(PT_HAS_CUBES) && obj.object_type == 2) {
        float x = rng.next(); float y = rng.next();
        samp = sample_light_box(S, light, obj, hit_pos, f2{x, y}, rng);
    }
    return samp;
}

// shader.wgsl:1648-1659
float mis_weight(float pdf_a, float pdf_b) {
    if (pdf_a <= 0.0f) return 0.0f;
    float max_pdf = fmx(pdf_a, pdf_b);
    float a = pdf_a / max_pdf, b = pdf_b / max_pdf;
    float a2 = a * a, b2 = b * b;
    float sum = a2 + b2;
    if (sum <= 0.0f) return 0.0f;
    return a2 / sum;
}

struct Medium { float ior; f3 sigma; f3 emission; };
struct MediumStack { Medium media[4]; int count = 0; };
Medium air() { return Medium{1.0f, V(0.f), V(0.f)}; }
Medium peek_medium(const MediumStack& s) { return s.count > 0 ? s.media[s.count - 1] : air(); }
Medium peek_outer_medium(const MediumStack& s) { return s.count > 1 ? s.media[s.count - 2] : air(); }
void push_medium(MediumStack& s, const Medium& m) { if (s.count < 4) { s.media[s.count] = m; s.count += 1; } }
void pop_medium(MediumStack& s) { if (s.count > 0) s.count -= 1; }

// shader.wgsl:1700-1704
float D_GGX(float NdotH, float alpha2) {
    float denom = (NdotH * NdotH * (alpha2 - 1.0f) + 1.0f);
    return alpha2 / (PI * denom * denom + 1e-10f);
}
// shader.wgsl:1721-1733
float fresnel_dielectric(float cosThetaI, float etai, float etat) {
    float cosi = clampf(std::fabs(cosThetaI), 0.0f, 1.0f);
    float sint = (etai / etat) * std::sqrt(fmx(0.0f, 1.0f - cosi * cosi));
    if (sint >= 1.0f) return 1.0f;
    float cost = std::sqrt(fmx(0.0f, 1.0f - sint * sint));
    float Rs = ((etat * cosi) - (etai * cost)) / ((etat * cosi) + (etai * cost));
    float Rp = ((etai * cosi) - (etat * cost)) / ((etai * cosi) + (etat * cost));
    return (Rs * Rs + Rp * Rp) / 2.0f;
}
// shader.wgsl:1751-1763
f3 importance_sample_cosine(f2 xi, f3 N) {
    float phi = TWO_PI * xi.x;
    float cosTheta = std::sqrt(clampf(xi.y, 0.0f, 1.0f));
    float sinTheta = std::sqrt(fmx(0.0f, 1.0f - xi.y));
    f3 L = V(std::cos(phi) * sinTheta, std::sin(phi) * sinTheta, cosTheta);
    f3 up = std::fabs(N.y) > 0.999f ? V(0.f, 0.f, 1.f) : V(0.f, 1.f, 0.f);
    f3 tangent = normalize(cross(up, N));
    f3 bitangent = cross(N, tangent);
    return normalize(tangent * L.x + bitangent * L.y + N * L.z);
}
// shader.wgsl:1787-1791
float G_Smith_GGX_div_NV(float NdotV, float NdotL, float alpha2) {
    float sqrt_v = std::sqrt(fmx(0.0f, alpha2 + NdotV * NdotV * (1.0f - alpha2)));
    float sqrt_l = std::sqrt(fmx(0.0f, alpha2 + NdotL * NdotL * (1.0f - alpha2)));
    return (2.0f * NdotL) / (NdotL * sqrt_v + NdotV * sqrt_l + 1e-10f);
}
// shader.wgsl:1794-1798
float G_Smith_GGX_Uncorrelated_div_NV(float NdotV, float NdotL, float alpha2) {
    float lambda_v = NdotV + std::sqrt(fmx(0.0f, alpha2 + NdotV * NdotV * (1.0f - alpha2)));
    float lambda_l = NdotL + std::sqrt(fmx(0.0f, alpha2 + NdotL * NdotL * (1.0f - alpha2)));
    return (4.0f * NdotL) / (lambda_v * lambda_l + 1e-7f);
}
// shader.wgsl:1801-1803
float G1_GGX_div_NV(float NdotV, float alpha2) {
    return 2.0f / (NdotV + std::sqrt(fmx(0.0f, alpha2 + (1.0f - alpha2) * NdotV * NdotV)) + 1e-7f);
}
// shader.wgsl:1806-1829
f3 importance_sample_vndf_ggx(f2 xi, f3 Vv, f3 N, float alpha_ggx) {
    f3 up = std::fabs(N.y) > 0.999f ? V(0.f, 0.f, 1.f) : V(0.f, 1.f, 0.f);
    f3 T = normalize(cross(up, N));
    f3 B = cross(N, T);
    f3 V_local = V(dot(Vv, T), dot(Vv, B), dot(Vv, N));
    f3 Vh = normalize(V(alpha_ggx * V_local.x, alpha_ggx * V_local.y, fmx(V_local.z, 0.0f)));
    float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    f3 T1 = lensq > 0.0f ? V(-Vh.y, Vh.x, 0.0f) / std::sqrt(fmx(lensq, 1e-7f)) : V(1.0f, 0.0f, 0.0f);
    f3 T2 = cross(Vh, T1);
    float r = std::sqrt(xi.x);
    float phi = 2.0f * PI * xi.y;
    float t1 = r * std::cos(phi);
    float t2 = r * std::sin(phi);
    float s = 0.5f * (1.0f + Vh.z);
    float t2_mod = mixf(std::sqrt(fmx(0.0f, 1.0f - t1 * t1)), t2, s);
    f3 Nh_local = t1 * T1 + t2_mod * T2 + std::sqrt(fmx(0.0f, 1.0f - t1 * t1 - t2_mod * t2_mod)) * Vh;
    f3 Ne_local = normalize(V(alpha_ggx * Nh_local.x, alpha_ggx * Nh_local.y, fmx(0.0f, Nh_local.z)));
    return normalize(T * Ne_local.x + B * Ne_local.y + N * Ne_local.z);
}

void media_iors(bool entering, const pt_material& mat, const MediumStack& stack, float* etai, float* etat) {
    if (entering) { *etai = peek_medium(stack).ior; *etat = mat.ior; }
    else { *etai = peek_medium(stack).ior; *etat = peek_outer_medium(stack).ior; }
}

inline float pow5(float x) { return std::pow(x, 5.0f); }

// shader.wgsl:1831-1947
f3 eval_surface(f3 Vv, f3 L, const pt_material& mat, const SurfaceContext& ctx, const MediumStack& stack) {
    bool entering = dot(ctx.normal, Vv) > 0.0f;
    f3 n_orient = entering ? ctx.normal : -ctx.normal;
    float dotNV = dot(n_orient, Vv);
    float dotNL = dot(n_orient, L);
    if (dotNV <= 0.0f) return V(0.0f);
    float roughness_val = clampf(ctx.roughness, 0.01f, 1.0f);
    float alpha = roughness_val * roughness_val;
    float alpha2 = fmx(alpha * alpha, 1e-6f);
    float etai, etat;
    media_iors(entering, mat, stack, &etai, &etat);
    if (dotNL > 0.0f) {
        f3 H_vec = Vv + L;
        f3 H = dot(H_vec, H_vec) > 1e-6f ? normalize(H_vec) : n_orient;
        float dotNH = fmx(dot(n_orient, H), 0.0f);
        float dotVH = fmx(dot(Vv, H), 0.0f);
        f3 diff_col = ctx.albedo;
        if (mat.sheen > 0.0f)
            diff_col = diff_col + mix3(V(1.0f), ctx.albedo, mat.sheen_tint) * pow5(fmx(0.0f, 1.0f - dotNV)) * mat.sheen;
        if (mat.subsurface > 0.0f) diff_col = mix3(diff_col, ld3(mat.subsurface_tint), mat.subsurface);
        f3 diffuse = (diff_col / PI) * (1.0f - ctx.metallic) * (1.0f - mat.transmission);
        float D = D_GGX(dotNH, alpha2);
        float G2_div_NV = G_Smith_GGX_div_NV(dotNV, dotNL, alpha2);
        float F_dielectric = fresnel_dielectric(dotVH, etai, etat);
        float lum = dot(ctx.albedo, V(0.2126f, 0.7152f, 0.0722f));
        f3 tint = lum <= 0.0f ? V(1.0f) : ctx.albedo / fmx(lum, 0.0001f);
        f3 F_dielectric_tinted = mix3(V(F_dielectric), V(F_dielectric) * tint, mat.specular_tint);
        f3 F0_metal = ctx.albedo;
        f3 F90_metal = mix3(V(1.0f), tint, mat.specular_tint);
        f3 F_metal = F0_metal + (F90_metal - F0_metal) * pow5(fmx(0.0f, 1.0f - dotVH));
        f3 F_actual = mix3(F_dielectric_tinted, F_metal, ctx.metallic);
        float F_macro_raw = fresnel_dielectric(dotNV, etai, etat);
        float F0_macro_d = std::pow((etai - etat) / (etai + etat), 2.0f);
        float F_macro_est = mixf(F_macro_raw, F0_macro_d, roughness_val);
        f3 F_macro_metal = ctx.albedo + (V(1.0f) - ctx.albedo) * pow5(fmx(0.0f, 1.0f - dotNV));
        f3 F_macro_actual = mix3(V(F_macro_est), F_macro_metal, ctx.metallic);
        f3 diffuse_cos = diffuse * dotNL * (V(1.0f) - F_macro_actual);
        float ms_factor = 1.0f + alpha * 0.4f;
        f3 specular_cos = (D * F_actual * G2_div_NV) / 4.0f * ms_factor;
        f3 clearcoat_cos = V(0.0f);
        float cc_attenuation = 1.0f;
        if (mat.clearcoat > 0.0f && entering) {
            float cc_roughness = clampf(1.0f - mat.clearcoat_gloss, 0.01f, 1.0f);
            float cc_alpha = cc_roughness * cc_roughness;
            float cc_alpha2 = fmx(cc_alpha * cc_alpha, 1e-6f);
            float cc_D = D_GGX(dotNH, cc_alpha2);
            float cc_G2_div_NV = G_Smith_GGX_div_NV(dotNV, dotNL, cc_alpha2);
            float cc_F = fresnel_dielectric(dotVH, 1.0f, mat.clearcoat_ior);
            clearcoat_cos = (mat.clearcoat * cc_D * cc_G2_div_NV * V(cc_F)) / 4.0f;
            float cc_F_macro = fresnel_dielectric(dotNV, 1.0f, mat.clearcoat_ior);
            cc_attenuation = 1.0f - (mat.clearcoat * cc_F_macro);
        }
        return ((diffuse_cos + specular_cos) * cc_attenuation + clearcoat_cos) * ctx.alpha;
    } else if (dotNL < 0.0f && mat.transmission > 0.0f) {
        f3 H_vec = -(etai * Vv + etat * L);
        f3 H = normalize(H_vec);
        if (dot(H, n_orient) < 0.0f) H = -H;
        float dotNH = fmx(dot(n_orient, H), 0.0f);
        float dotVH = dot(Vv, H);
        float dotLH = dot(L, H);
        if (dotVH * dotLH >= 0.0f) return V(0.0f);
        float D = D_GGX(dotNH, alpha2);
        float G2_unc_div_NV = G_Smith_GGX_Uncorrelated_div_NV(dotNV, std::fabs(dotNL), alpha2);
        float F_dielectric = fresnel_dielectric(dotVH, etai, etat);
        float denom = etai * dotVH + etat * dotLH;
        float denom2 = fmx(denom * denom, 1e-7f);
        float ms_factor = 1.0f + alpha * 0.25f;
        float btdf_cos = (std::fabs(dotVH) * std::fabs(dotLH) * etat * etat * (1.0f - F_dielectric) * D * G2_unc_div_NV) /
                         denom2 * ms_factor;
        float cc_attenuation = 1.0f;
        if (mat.clearcoat > 0.0f && entering) {
            float cc_F_macro = fresnel_dielectric(dotNV, 1.0f, mat.clearcoat_ior);
            cc_attenuation = 1.0f - (mat.clearcoat * cc_F_macro);
        }
        f3 trans_color = V(mat.transmission * (1.0f - ctx.metallic));
        return trans_color * btdf_cos * cc_attenuation * ctx.alpha;
    }
    return V(0.0f);
}

struct LobeProbs { float p_cc, p_spec, p_trans, p_diff; };

// The lobe-selection weights shared (with different cosines) by pdf_surface
// (shader.wgsl:1975-1987) and sample_surface (shader.wgsl:2078-2090).
LobeProbs lobe_probs(float cos_v, bool entering, float etai, float etat, float roughness_val, const pt_material& mat,
                     const SurfaceContext& ctx) {
    float F_cc = fresnel_dielectric(cos_v, 1.0f, mat.clearcoat_ior);
    float p_cc = entering ? mat.clearcoat * F_cc : 0.0f;
    float F_est_raw = fresnel_dielectric(cos_v, etai, etat);
    float F0_d = std::pow((etai - etat) / (etai + etat), 2.0f);
    float F_est = mixf(F_est_raw, F0_d, roughness_val);
    f3 F_metal = ctx.albedo + (V(1.0f) - ctx.albedo) * pow5(fmx(0.0f, 1.0f - cos_v));
    f3 F_actual = mix3(V(F_est), F_metal, ctx.metallic);
    float f_avg = clampf((F_actual.x + F_actual.y + F_actual.z) / 3.0f, 0.0f, 1.0f);
    LobeProbs p;
    p.p_cc = p_cc;
    p.p_spec = (1.0f - p_cc) * f_avg;
    p.p_trans = (1.0f - p_cc) * (1.0f - f_avg) * mat.transmission * (1.0f - ctx.metallic);
    p.p_diff = (1.0f - p_cc) * (1.0f - f_avg) * (1.0f - mat.transmission) * (1.0f - ctx.metallic);
    return p;
}

// shader.wgsl:1952-2047
float pdf_surface(f3 Vv, f3 L, const pt_material& mat, const SurfaceContext& ctx, const MediumStack& stack) {
    bool entering = dot(ctx.normal, Vv) > 0.0f;
    f3 n_orient = entering ? ctx.normal : -ctx.normal;
    float dotNV = dot(n_orient, Vv);
    float dotNL = dot(n_orient, L);
    if (dotNV <= 0.0f) return 1e-7f;
    float roughness_val = clampf(ctx.roughness, 0.01f, 1.0f);
    float alpha = roughness_val * roughness_val;
    float alpha2 = fmx(alpha * alpha, 1e-6f);
    float etai, etat;
    media_iors(entering, mat, stack, &etai, &etat);
    LobeProbs P = lobe_probs(dotNV, entering, etai, etat, roughness_val, mat, ctx);
    float pdf = 0.0f;
    if (dotNL > 0.0f) {
        f3 H_vec = Vv + L;
        f3 H = dot(H_vec, H_vec) > 1e-6f ? normalize(H_vec) : n_orient;
        float dotNH = fmx(dot(n_orient, H), 0.0f);
        float dotVH = fmx(dot(Vv, H), 0.0f);
        if (P.p_diff > 0.0f) pdf += P.p_diff * (dotNL / PI);
        if (P.p_spec > 0.0f) {
            float D = D_GGX(dotNH, alpha2);
            float G1_div_NV = G1_GGX_div_NV(dotNV, alpha2);
            float spec_weight = P.p_spec;
            if (P.p_trans > 0.0f) {
                float cosi = clampf(dotVH, 0.0f, 1.0f);
                float sint = (etai / etat) * std::sqrt(fmx(0.0f, 1.0f - cosi * cosi));
                if (sint >= 1.0f) spec_weight += P.p_trans;
            }
            float pdf_H = G1_div_NV * D * fmx(0.0f, dotVH);
            pdf += spec_weight * pdf_H / fmx(4.0f * dotVH, 1e-7f);
        }
        if (P.p_cc > 0.0f) {
            float cc_roughness = clampf(1.0f - mat.clearcoat_gloss, 0.01f, 1.0f);
            float cc_alpha = cc_roughness * cc_roughness;
            float cc_alpha2 = fmx(cc_alpha * cc_alpha, 1e-6f);
            float cc_D = D_GGX(dotNH, cc_alpha2);
            float cc_G1_div_NV = G1_GGX_div_NV(dotNV, cc_alpha2);
            float pdf_H_cc = cc_G1_div_NV * cc_D * fmx(0.0f, dotVH);
            pdf += P.p_cc * pdf_H_cc / fmx(4.0f * dotVH, 1e-7f);
        }
    } else if (dotNL < 0.0f && P.p_trans > 0.0f) {
        f3 H_vec = -(etai * Vv + etat * L);
        f3 H = normalize(H_vec);
        if (dot(H, n_orient) < 0.0f) H = -H;
        float dotNH = fmx(dot(n_orient, H), 0.0f);
        float dotVH = dot(Vv, H);
        float dotLH = dot(L, H);
        if (dotVH * dotLH < 0.0f) {
            float D = D_GGX(dotNH, alpha2);
            float G1_div_NV = G1_GGX_div_NV(dotNV, alpha2);
            float pdf_H = G1_div_NV * D * fmx(0.0f, dotVH);
            float denom = etai * dotVH + etat * dotLH;
            float dwh_dwi = (etat * etat * std::fabs(dotLH)) / fmx(denom * denom, 1e-7f);
            pdf += P.p_trans * pdf_H * dwh_dwi;
        }
    }
    return fmx(pdf, 1e-7f) * ctx.alpha;
}

// shader.wgsl:2052-2175 (the radiance argument is unused there)
bool sample_surface(Ray* ray, f3* throughput, const pt_material& mat, const SurfaceContext& ctx, f3 hit_pos,
                    MediumStack* stack, float* last_surface_pdf, Rng& rng) {
    f3 Vv = -ray->direction;
    bool entering = dot(ctx.normal, Vv) > 0.0f;
    f3 n_orient = entering ? ctx.normal : -ctx.normal;
    f3 sn_orient = entering ? ctx.surface_normal : -ctx.surface_normal;
    float dotSNV = fmx(dot(sn_orient, Vv), 1e-6f);
    float roughness_val = clampf(ctx.roughness, 0.01f, 1.0f);
    float alpha = roughness_val * roughness_val;
    float cc_roughness = clampf(1.0f - mat.clearcoat_gloss, 0.01f, 1.0f);
    float cc_alpha = cc_roughness * cc_roughness;
    float etai, etat;
    media_iors(entering, mat, *stack, &etai, &etat);
    LobeProbs P = lobe_probs(dotSNV, entering, etai, etat, roughness_val, mat, ctx);
    float total_p = P.p_cc + P.p_spec + P.p_trans + P.p_diff;
    if (rng.next() > ctx.alpha) {
        ray->origin = hit_pos - n_orient * 0.005f;
        return true;
    }
    float lobe = rng.next();
    float xi_x = rng.next();
    float xi_y = rng.next();
    f2 xi{xi_x, xi_y};
    f3 L = V(0.0f);
    bool is_transmission = false;
    if (lobe < P.p_cc) {
        f3 H = importance_sample_vndf_ggx(xi, Vv, sn_orient, cc_alpha);
        L = reflect(-Vv, H);
    } else if (lobe < P.p_cc + P.p_spec) {
        f3 N_spec = n_orient;
        if (mat.anisotropic > 0.0f) {
            f3 up = std::fabs(n_orient.y) > 0.99999f ? V(0.f, 0.f, 1.f) : V(0.f, 1.f, 0.f);
            f3 T = normalize(cross(up, n_orient));
            f3 B = cross(n_orient, T);
            float angle = mat.aniso_rotation * TWO_PI;
            T = T * std::cos(angle) + B * std::sin(angle);
            N_spec = normalize(mix3(n_orient, cross(cross(Vv, T), T), mat.anisotropic));
        }
        f3 H = importance_sample_vndf_ggx(xi, Vv, N_spec, alpha);
        L = reflect(-Vv, H);
    } else if (lobe < P.p_cc + P.p_spec + P.p_trans) {
        if (std::fabs(etai - etat) < 1e-4f) {
            ray->direction = -Vv;
            ray->origin = hit_pos - n_orient * 0.001f;
            *throughput = *throughput * V(mat.transmission * (1.0f - ctx.metallic));
            *last_surface_pdf = P.p_trans;
            if (entering) {
                f3 sigma = -V(std::log(fmx(ctx.albedo.x, 0.0001f)), std::log(fmx(ctx.albedo.y, 0.0001f)),
                              std::log(fmx(ctx.albedo.z, 0.0001f))) * mat.concentration;
                push_medium(*stack, Medium{mat.ior, sigma, ld3(mat.emittance)});
            } else {
                pop_medium(*stack);
            }
            return true;
        }
        f3 H = importance_sample_vndf_ggx(xi, Vv, n_orient, alpha);
        float eta_ratio = etai / etat;
        L = refract(-Vv, H, eta_ratio);
        if (length(L) < 0.1f) L = reflect(-Vv, H);
        else is_transmission = true;
    } else if (lobe < total_p) {
        L = importance_sample_cosine(xi, n_orient);
    } else {
        return false;
    }
    float dotNL = dot(n_orient, L);
    if (is_transmission) {
        if (dotNL >= 0.0f) return false;
        ray->origin = hit_pos - n_orient * 0.001f;
    } else {
        if (dotNL <= 0.0f) return false;
        ray->origin = hit_pos + n_orient * 0.001f;
    }
    ray->direction = L;
    f3 bsdf_val = eval_surface(Vv, L, mat, ctx, *stack);
    float pdf_val = pdf_surface(Vv, L, mat, ctx, *stack);
    if (pdf_val > 0.0f) *throughput = *throughput * (bsdf_val / pdf_val);
    else return false;
    *last_surface_pdf = pdf_val;
    if (is_transmission) {
        if (entering) {
            f3 sigma = -V(std::log(fmx(ctx.albedo.x, 0.0001f)), std::log(fmx(ctx.albedo.y, 0.0001f)),
                          std::log(fmx(ctx.albedo.z, 0.0001f))) * mat.concentration;
            push_medium(*stack, Medium{mat.ior, sigma, ld3(mat.emittance)});
        } else {
            pop_medium(*stack);
        }
    }
    return true;
}

// shader.wgsl:2179-2189
f3 clamp_firefly(f3 rad, int bounce) {
    if (bounce == 0) return rad;
    float lum = dot(rad, V(0.2126f, 0.7152f, 0.0722f));
    if (lum > 20.0f) return rad * (20.0f / lum);
    return rad;
}

// shader.wgsl:2226-2271
f3 evaluate_nee(const ref_scene& S, const LightSample& ls, f3 hit_pos, f3 n_orient, f3 Vv, const pt_material& mat,
                const SurfaceContext& ctx, const SurfaceHit& hit, float cur_height, f2 final_uv,
                const MediumStack& stack, int ignore_obj_idx, Counters* c) {
    if (ls.pdf <= 0.0f) return V(0.0f);
    float dotNL = dot(n_orient, ls.dir);
    if (dotNL <= 0.0f && mat.transmission <= 0.0f) return V(0.0f);
    float bsdf_pdf = pdf_surface(Vv, ls.dir, mat, ctx, stack);
    float weight = mis_weight(ls.pdf, bsdf_pdf);
    f3 bsdf_val = eval_surface(Vv, ls.dir, mat, ctx, stack);
    f3 sample_color = ls.color;
    f3 c_weight = weight * bsdf_val / fmx(ls.pdf, 1e-6f);
    if (max_component(sample_color * c_weight) < 1e-5f) return V(0.0f);
    if (S.has(PT_HAS_HEIGHTMAPS) && mat.height_idx >= 0) {  // shader.wgsl:2254-2260
        f3 light_ts = normalize(tangent_space(hit, ls.dir));
        if (calculate_shadow_pom(S, final_uv, cur_height, light_ts, mat).hit) return V(0.0f);
    }
    f3 ray_offset = dotNL < 0.0f ? -n_orient * 0.001f : n_orient * 0.001f;
    Ray shadow_ray{hit_pos + ray_offset, ls.dir};
    f3 shadow_hit = trace_scene_shadow(S, shadow_ray, ignore_obj_idx, ls.dist, c);
    sample_color = sample_color * shadow_hit;
    if (max_component(shadow_hit) < 0.001f) return V(0.0f);
    return sample_color * c_weight;
}

struct HitRecorder { pt_hit* out; uint32_t n; float* rays; };

void store_hit(pt_hit* out, const SurfaceHit& h) {
    std::memset(out, 0, sizeof(*out));
    out->t = h.t;
    out->o_idx = h.o_idx;
    out->m_idx = h.m_idx;
    out->prim_id = h.prim_id;
    out->plane_idx = h.plane_idx;
    out->count = uint32_t(h.count);
    out->n[0] = h.hit_n.x; out->n[1] = h.hit_n.y; out->n[2] = h.hit_n.z;
    out->uv[0] = h.hit_uv.x; out->uv[1] = h.hit_uv.y;
}

// shader.wgsl:2273-2480: radiance of one path through pixel (x, y).
f3 trace_path(const ref_scene& S, const pt_scene_params& P, uint32_t x, uint32_t y, Counters* c, HitRecorder* rec) {
    Rng rng{0u};
    Ray ray = camera_ray(P, x, y, &rng, nullptr);
    f3 throughput = V(1.0f), radiance = V(0.0f);
    float last_surface_pdf = 1.0f, last_weight_sum = 0.0f;
    MediumStack stack;
    c->paths++;
    const bool has_prims = S.has(PT_HAS_SPHERES | PT_HAS_CUBES | PT_HAS_CYLINDERS | PT_HAS_TORI | PT_HAS_MESHES);
    (void)has_prims;
    for (int bounce = 0; bounce < S.bounce_limit; ++bounce) {
        if (rec && rec->rays && uint32_t(bounce) < rec->n) {
            float* r = rec->rays + size_t(bounce) * 6;
            r[0] = ray.origin.x; r[1] = ray.origin.y; r[2] = ray.origin.z;
            r[3] = ray.direction.x; r[4] = ray.direction.y; r[5] = ray.direction.z;
        }
        SurfaceHit hit = trace_scene(S, ray, c);
        if (rec && uint32_t(bounce) < rec->n) store_hit(&rec->out[bounce], hit);
        pt_object obj;
        std::memset(&obj, 0, sizeof(obj));
        if (hit.o_idx >= 0) obj = S.objs[hit.o_idx];
        Medium cur = peek_medium(stack);
        if (cur.ior > 1.0f || length(cur.sigma) > 0.0f) {
            float dist = hit.m_idx == -1 ? 1000.0f : hit.t;
            f3 att = V(std::exp(-cur.sigma.x * dist), std::exp(-cur.sigma.y * dist), std::exp(-cur.sigma.z * dist));
            f3 sg = cur.sigma;
            float ex = sg.x > 1e-4f ? cur.emission.x * (1.0f - att.x) / sg.x : cur.emission.x * dist;
            float ey = sg.y > 1e-4f ? cur.emission.y * (1.0f - att.y) / sg.y : cur.emission.y * dist;
            float ez = sg.z > 1e-4f ? cur.emission.z * (1.0f - att.z) / sg.z : cur.emission.z * dist;
            radiance = radiance + throughput * clamp_firefly(V(ex, ey, ez), bounce);
            throughput = throughput * att;
        }
        if (hit.m_idx == -1) {
            if (S.has(PT_HAS_SKYBOX)) {
                f3 sky_color = sample_sky(S, ray.direction);
                float weight = 1.0f;
                if (bounce > 0 && S.has(PT_MIS_SKYBOX)) {
                    float sky_pdf = get_sky_pdf(S, P, ray.direction);
                    weight = mis_weight(last_surface_pdf, sky_pdf);
                }
                radiance = radiance + throughput * clamp_firefly(sky_color * weight, bounce);
            } else {
                radiance = radiance + throughput * clamp_firefly(V(0.02f, 0.03f, 0.05f), bounce);
            }
            break;
        }
        c->shade++;
        const pt_material& mat = S.mats[hit.m_idx];
        f2 final_uv = material_uv(hit, mat);
        float cur_height = 0.0f;
        if (S.has(PT_HAS_HEIGHTMAPS) && mat.height_idx >= 0) {  // shader.wgsl:2365-2373
            f3 view_ts = normalize(tangent_space(hit, -ray.direction));
            if (view_ts.z > 0.0f) {
                PomResult pom = calculate_pom(S, final_uv, view_ts, mat);
                final_uv = pom.uv;
                cur_height = pom.height;
            }
        }
        SurfaceContext ctx = get_surface_context(S, hit, mat, final_uv);
        if (length(ctx.emittance) > 0.0f) {
            float weight = 1.0f;
            if (S.has(PT_HAS_LIGHTS) && bounce > 0 && hit.o_idx >= 0 && obj.light_idx >= 0) {
                const pt_light& light = S.lights[obj.light_idx];
                float raw_light_pdf = get_light_pdf(S, light, obj, ray.origin, ray.direction, hit);
                f3 to_light = ld3(obj.world_position) - ray.origin;
                float d2 = fmx(dot(to_light, to_light), 0.001f);
                float light_importance = light.power / d2;
                float selection_pdf = last_weight_sum > 0.0f ? light_importance / last_weight_sum : 0.0f;
                float total_light_pdf = raw_light_pdf * selection_pdf;
                weight = mis_weight(last_surface_pdf, total_light_pdf);
            }
            radiance = radiance + throughput * clamp_firefly(ctx.emittance * weight, bounce);
            if (length(ctx.emittance) > 1.0f) break;
        }
        f3 hit_pos = ray.origin + ray.direction * hit.t;
        f3 Vv = -ray.direction;
        f3 n_orient = dot(ctx.normal, Vv) > 0.0f ? ctx.normal : -ctx.normal;
        if (S.has(PT_HAS_LIGHTS)) {
            int selected = -1;
            float weight_sum = 0.0f, selected_weight = 0.0f;
            for (size_t i = 0; i < S.lights.size(); ++i) {
                const pt_light& light = S.lights[i];
                if (light.obj_idx < 0) continue;
                const pt_object& lo = S.objs[light.obj_idx];
                f3 to_light = ld3(lo.world_position) - hit_pos;
                float d2 = fmx(dot(to_light, to_light), 0.001f);
                float w = light.power / d2;
                weight_sum += w;
                if (rng.next() < (w / weight_sum)) { selected = int(i); selected_weight = w; }
            }
            if (selected >= 0 && weight_sum > 0.0f) {
                const pt_light& light = S.lights[selected];
                const pt_object& lo = S.objs[light.obj_idx];
                LightSample ls = sample_light(S, light, lo, hit_pos, rng);
                if (ls.pdf > 0.0f) {
                    float selection_pdf = selected_weight / weight_sum;
                    ls.pdf *= selection_pdf;
                    f3 inc = evaluate_nee(S, ls, hit_pos, n_orient, Vv, mat, ctx, hit, cur_height, final_uv, stack,
                                          light.obj_idx, c);
                    radiance = radiance + throughput * clamp_firefly(inc, bounce);
                }
            }
            last_weight_sum = weight_sum;
        }
        if (S.has(PT_HAS_SKYBOX) && S.has(PT_MIS_SKYBOX)) {
            float ux = rng.next();
            float uy = rng.next();
            LightSample env = sample_env_cdf(S, P, f2{ux, uy});
            f3 inc = evaluate_nee(S, env, hit_pos, n_orient, Vv, mat, ctx, hit, cur_height, final_uv, stack, -1, c);
            radiance = radiance + throughput * clamp_firefly(inc, bounce);
        }
        if (!sample_surface(&ray, &throughput, mat, ctx, hit_pos, &stack, &last_surface_pdf, rng)) break;
        if (S.has(PT_HAS_HEIGHTMAPS) && mat.height_idx >= 0) {  // shader.wgsl:2458-2472
            f3 light_ts = normalize(tangent_space(hit, ray.direction));
            PomResult sr = calculate_shadow_pom(S, final_uv, cur_height, light_ts, mat);
            if (sr.hit) {
                final_uv = sr.uv;
                SurfaceContext ctx_pom = get_surface_context(S, hit, mat, final_uv);
                radiance = radiance + throughput * clamp_firefly(ctx_pom.emittance, bounce);
                if (length(ctx_pom.emittance) > 1.0f) break;
                if (!sample_surface(&ray, &throughput, mat, ctx_pom, hit_pos, &stack, &last_surface_pdf, rng)) break;
            }
        }
        if (bounce < 2) continue;
        float p = max_component(throughput);
        float survival_prob = clampf(p, 0.05f, 0.95f);
        if (rng.next() > survival_prob) break;
        throughput = throughput / survival_prob;
    }
    return radiance;
}

// Run fn(y) for the given rows on `threads` threads.
template <class F>
void parallel_rows(uint32_t n_rows, int threads, F fn) {
    int nt = threads > 0 ? threads : int(std::max(1u, std::thread::hardware_concurrency()));
    nt = std::min<int>(nt, int(std::max<uint32_t>(1, n_rows)));
    std::atomic<uint32_t> next{0};
    auto worker = [&](int tid) {
        for (;;) {
            uint32_t r = next.fetch_add(1);
            if (r >= n_rows) break;
            fn(r, tid);
        }
    };
    if (nt <= 1) { worker(0); return; }
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t) pool.emplace_back(worker, t);
    for (auto& th : pool) th.join();
}

f3 tonemap_store(f3 c, float exposure) {
    c = c * V(exposure);
    c = c / (c + V(0.3f));
    return V(std::pow(c.x, 0.4545f), std::pow(c.y, 0.4545f), std::pow(c.z, 0.4545f));
}

void fill_stats(pt_stats* st, const Counters& c) {
    if (!st) return;
    std::memset(st, 0, sizeof(*st));
    st->extension_rays = c.ext; st->shadow_rays = c.shadow; st->paths = c.paths; st->node_pops = c.pops;
    st->inner_pops = c.inner; st->object_visits = c.objs; st->tri_tests = c.tris; st->shade_events = c.shade;
}

}  // namespace

// =================================================================== C ABI ==
extern "C" {

int ref_scene_create(ref_scene** out, const pt_scene_desc* d) {
    if (!out || !d) return PT_ERR_ARG;
    auto* s = new (std::nothrow) ref_scene();
    if (!s) return PT_ERR_OOM;
    auto copy = [](auto& vec, const pt_array& a) {
        using T = typename std::remove_reference<decltype(vec)>::type::value_type;
        if (a.stride != sizeof(T) || (a.count && !a.ptr)) return false;
        vec.assign(static_cast<const T*>(a.ptr), static_cast<const T*>(a.ptr) + a.count);
        return true;
    };
    bool ok = copy(s->mats, d->materials) && copy(s->objs, d->objects) && copy(s->tlas, d->tlas) &&
              copy(s->meshes, d->meshes) && copy(s->bvh, d->bvh) && copy(s->tris, d->triangles) &&
              copy(s->planes, d->planes) && copy(s->lights, d->lights);
    if (!ok || !d->sky_rgba16f || !d->sky_cdf || d->sky_width == 0 || d->sky_height == 0) { delete s; return PT_ERR_LAYOUT; }
    s->sky_w = d->sky_width;
    s->sky_h = d->sky_height;
    s->sky.resize(size_t(s->sky_w) * s->sky_h * 4);
    for (size_t i = 0; i < s->sky.size(); ++i) s->sky[i] = half_to_float(d->sky_rgba16f[i]);
    s->cdf.assign(d->sky_cdf, d->sky_cdf + d->sky_cdf_count);
    s->flags = d->flags;
    s->bounce_limit = d->bounce_limit;
    if (d->texture_layers > 0) {
        if (!d->textures) { delete s; return PT_ERR_LAYOUT; }
        s->tex.assign(d->textures, d->textures + size_t(d->texture_layers) * kTexSize * kTexSize * 4);
        s->tex_layers = d->texture_layers;
    }
    *out = s;
    return PT_OK;
}

void ref_scene_destroy(ref_scene* s) { delete s; }

int ref_trace_rays(const ref_scene* s, const float* rays, uint32_t n, pt_hit* hits, pt_stats* stats) {
    if (!s || (n && (!rays || !hits))) return PT_ERR_ARG;
    Counters c;
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = rays + size_t(i) * 6;
        Ray ray{V(r[0], r[1], r[2]), V(r[3], r[4], r[5])};
        SurfaceHit h = trace_scene(*s, ray, &c);
        store_hit(&hits[i], h);
    }
    fill_stats(stats, c);
    return PT_OK;
}

int ref_trace_shadow_rays(const ref_scene* s, const float* rays, const int32_t* target_idx, const float* target_dist,
                          uint32_t n, float* trans) {
    if (!s || (n && (!rays || !target_idx || !target_dist || !trans))) return PT_ERR_ARG;
    Counters c;
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = rays + size_t(i) * 6;
        Ray ray{V(r[0], r[1], r[2]), V(r[3], r[4], r[5])};
        f3 t = trace_scene_shadow(*s, ray, target_idx[i], target_dist[i], &c);
        trans[3 * i] = t.x; trans[3 * i + 1] = t.y; trans[3 * i + 2] = t.z;
    }
    return PT_OK;
}

int ref_render(const ref_scene* s, const pt_scene_params* params, uint32_t section_w, uint32_t section_h,
               uint32_t n_passes, const uint32_t* seeds, float* accum, uint8_t* out_rgba8, pt_stats* stats,
               int threads, uint32_t row_stride) {
    if (!s || !params || !accum || params->width == 0 || params->height == 0) return PT_ERR_ARG;
    const pt_scene_params& base = *params;
    const uint32_t sx = base.section[0], sy = base.section[1];
    const uint32_t sw = section_w ? section_w : base.width, sh = section_h ? section_h : base.height;
    const uint32_t stride = row_stride ? row_stride : 1;
    const uint32_t n_rows = (sh + stride - 1) / stride;
    int nt = threads > 0 ? threads : int(std::max(1u, std::thread::hardware_concurrency()));
    std::vector<Counters> per(size_t(nt) + 1);
    for (uint32_t pass = 0; pass < n_passes; ++pass) {
        pt_scene_params P = base;
        P.sample_number = base.sample_number + pass;
        if (seeds) P.seed = seeds[pass];
        parallel_rows(n_rows, nt, [&](uint32_t r, int tid) {
            const uint32_t y = sy + r * stride;
            if (y >= P.height) return;
            for (uint32_t x = sx; x < sx + sw && x < P.width; ++x) {
                f3 rad = trace_path(*s, P, x, y, &per[tid], nullptr);
                const size_t i = size_t(y) * P.width + x;
                float w = 1.0f / float(P.sample_number + 1u);
                f3 old_c = V(accum[4 * i], accum[4 * i + 1], accum[4 * i + 2]);
                f3 fc = mix3(old_c, rad, w);
                accum[4 * i] = fc.x; accum[4 * i + 1] = fc.y; accum[4 * i + 2] = fc.z; accum[4 * i + 3] = 1.0f;
                if (out_rgba8) {
                    f3 o = tonemap_store(fc, P.exposure);
                    out_rgba8[4 * i] = to_unorm8(o.x); out_rgba8[4 * i + 1] = to_unorm8(o.y);
                    out_rgba8[4 * i + 2] = to_unorm8(o.z); out_rgba8[4 * i + 3] = 255;
                }
            }
        });
    }
    Counters total;
    for (auto& c : per) total.add(c);
    fill_stats(stats, total);
    return PT_OK;
}

int ref_trace_hits(const ref_scene* s, const pt_scene_params* params, uint32_t n_bounces, pt_hit* hits, int threads) {
    return ref_trace_hits_rays(s, params, n_bounces, hits, nullptr, threads);
}

int ref_trace_hits_rays(const ref_scene* s, const pt_scene_params* params, uint32_t n_bounces, pt_hit* hits,
                        float* rays, int threads) {
    if (!s || !params || !hits || n_bounces == 0) return PT_ERR_ARG;
    const pt_scene_params& P = *params;
    for (size_t i = 0; i < size_t(P.width) * P.height * n_bounces; ++i) {
        std::memset(&hits[i], 0, sizeof(pt_hit));
        hits[i].t = -1.0f; hits[i].o_idx = hits[i].m_idx = hits[i].prim_id = hits[i].plane_idx = -1;
    }
    if (rays) std::memset(rays, 0, size_t(P.width) * P.height * n_bounces * 6 * sizeof(float));
    parallel_rows(P.height, threads, [&](uint32_t y, int) {
        Counters c;
        for (uint32_t x = 0; x < P.width; ++x) {
            const size_t base = (size_t(y) * P.width + x) * n_bounces;
            HitRecorder rec{hits + base, n_bounces, rays ? rays + base * 6 : nullptr};
            trace_path(*s, P, x, y, &c, &rec);
        }
    });
    return PT_OK;
}

// shader.wgsl:2510-2603
int ref_gbuffer(const ref_scene* s, const pt_scene_params* params, pt_gbuffer_texel* gbuf, uint8_t* out, int threads) {
    if (!s || !params || !gbuf) return PT_ERR_ARG;
    const pt_scene_params& P = *params;
    const ref_scene& S = *s;
    parallel_rows(P.height, threads, [&](uint32_t y, int) {
        Counters c;
        for (uint32_t x = 0; x < P.width; ++x) {
            const uint32_t idx = y * P.width + x;
            Ray ray = camera_ray(P, x, y, nullptr, nullptr);
            SurfaceContext ctx;
            std::memset(&ctx, 0, sizeof(ctx));
            pt_material mat;
            std::memset(&mat, 0, sizeof(mat));
            float dist = 0.0f;
            f3 tint = V(1.0f);
            for (int bounce = 0; bounce < S.bounce_limit; ++bounce) {
                SurfaceHit hit = trace_scene(S, ray, &c);
                if (hit.m_idx == -1) {
                    ctx.normal = V(0.0f);
                    ctx.emittance = sample_sky(S, ray.direction);
                    ctx.albedo = V(0.0f);
                    break;
                }
                mat = S.mats[hit.m_idx];
                ctx = get_surface_context(S, hit, mat, material_uv(hit, mat));
                ctx.normal = dot(ctx.normal, -ray.direction) > 0.0f ? ctx.normal : -ctx.normal;
                dist += hit.t;
                if (ctx.alpha < 1.0f) {
                    tint = tint * mix3(V(1.0f), ctx.albedo, ctx.alpha);
                    ray.origin = ray.origin + ray.direction * hit.t - ctx.normal * 0.001f;
                    continue;
                }
                if (mat.metallic >= 0.99f && mat.roughness <= 0.01f) {
                    tint = tint * ctx.albedo;
                    ray.origin = ray.origin + ray.direction * hit.t + ctx.normal * 0.001f;
                    ray.direction = reflect(ray.direction, ctx.normal);
                    continue;
                }
                break;
            }
            f3 color = (ctx.albedo + ctx.emittance) * tint;
            pt_gbuffer_texel& g = gbuf[idx];
            g.normal[0] = ctx.normal.x; g.normal[1] = ctx.normal.y; g.normal[2] = ctx.normal.z;
            g.albedo[0] = color.x; g.albedo[1] = color.y; g.albedo[2] = color.z;
            g.dist = dist;
            g.transmission = mat.transmission;
            g.roughness = ctx.roughness;
            g.metallic = ctx.metallic;
            g.motion[0] = g.motion[1] = 0.0f;
            if (out) {
                f3 o;
                if (x < P.width * 1u / 4u) { float dd = dist / (dist + 3.0f); o = V(dd); }
                else if (x < P.width * 2u / 4u) o = ctx.normal * 0.5f + V(0.5f);
                else if (x < P.width * 3u / 4u) o = tonemap_store(color, P.exposure);
                else o = V(mat.transmission, ctx.roughness, ctx.metallic);
                out[4 * idx] = to_unorm8(o.x); out[4 * idx + 1] = to_unorm8(o.y);
                out[4 * idx + 2] = to_unorm8(o.z); out[4 * idx + 3] = 255;
            }
        }
    });
    return PT_OK;
}

// denoiser.wgsl:55-214 + Denoiser.execute tracer.js:2583-2637
int ref_denoise(const pt_denoise_params* u, int stride_mode, const float* accum, const pt_gbuffer_texel* gbuf,
                float* out_rgba, uint8_t* out_rgba8) {
    if (!u || !accum || !gbuf || u->width == 0 || u->height == 0) return PT_ERR_ARG;
    const int W = int(u->width), H = int(u->height);
    const size_t n = size_t(W) * H;
    std::vector<float> A(n * 4), B(n * 4);
    auto lumf = [](f3 c) { return dot(c, V(0.299f, 0.587f, 0.114f)); };
    auto tonemap = [](f3 c, float e) {
        f3 col = c * e;
        col = col / (col + V(0.3f));
        return V(std::pow(col.x, 0.4545f), std::pow(col.y, 0.4545f), std::pow(col.z, 0.4545f));
    };
    auto gn = [&](size_t i) { return ld3(gbuf[i].normal); };
    auto ga = [&](size_t i) { return ld3(gbuf[i].albedo); };
    // prepass (rows in parallel; each pixel's arithmetic is unchanged)
    parallel_rows(uint32_t(H), 0, [&](uint32_t yy, int) {
        const int y = int(yy);
        for (int x = 0; x < W; ++x) {
            const size_t idx = size_t(y) * W + x;
            float cc[4];
            if (u->flags & 1u) {
                float colors[9][4], lums[9];
                int k = 0;
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        int cx = std::min(std::max(x + dx, 0), W - 1), cy = std::min(std::max(y + dy, 0), H - 1);
                        const float* col = accum + (size_t(cy) * W + cx) * 4;
                        for (int q = 0; q < 4; ++q) colors[k][q] = col[q];
                        lums[k] = lumf(V(col[0], col[1], col[2]));
                        k++;
                    }
                for (int i = 0; i < 5; ++i)
                    for (int j = i + 1; j < 9; ++j)
                        if (lums[j] < lums[i]) {
                            std::swap(lums[i], lums[j]);
                            for (int q = 0; q < 4; ++q) std::swap(colors[i][q], colors[j][q]);
                        }
                for (int q = 0; q < 4; ++q) cc[q] = colors[4][q];
            } else {
                for (int q = 0; q < 4; ++q) cc[q] = accum[idx * 4 + q];
            }
            bool valid = cc[0] >= 0.0f && cc[1] >= 0.0f && cc[2] >= 0.0f && cc[0] <= 10000.0f && cc[1] <= 10000.0f &&
                         cc[2] <= 10000.0f;
            if (!valid) cc[0] = cc[1] = cc[2] = 0.0f;
            f3 mapped = tonemap(V(cc[0], cc[1], cc[2]), u->exposure);
            bool is_sky = length(gn(idx)) < 0.1f;
            if ((u->flags & 2u) && !is_sky) {
                f3 ma = fmx(tonemap(ga(idx), u->exposure), V(0.001f));
                mapped = mapped / ma;
                mapped = V(clampf(mapped.x, 0.0f, 10.0f), clampf(mapped.y, 0.0f, 10.0f), clampf(mapped.z, 0.0f, 10.0f));
            }
            A[idx * 4] = mapped.x; A[idx * 4 + 1] = mapped.y; A[idx * 4 + 2] = mapped.z; A[idx * 4 + 3] = cc[3];
        }
    });
    const float kernel[5] = {0.0625f, 0.25f, 0.375f, 0.25f, 0.0625f};
    for (uint32_t pass = 0; pass < u->passes; ++pass) {
        const uint32_t step = stride_mode == 0 ? (1u << pass) : (1u << (u->passes - 1));
        const std::vector<float>& in = (pass % 2 == 0) ? A : B;
        std::vector<float>& out = (pass % 2 == 0) ? B : A;
        parallel_rows(uint32_t(H), 0, [&](uint32_t yy, int) {
            const int y = int(yy);
            for (int x = 0; x < W; ++x) {
                const size_t ci = size_t(y) * W + x;
                const float* cc = &in[ci * 4];
                const pt_gbuffer_texel& cg = gbuf[ci];
                float sum[4] = {0, 0, 0, 0}, sum_w = 0.0f;
                float preserve = (1.0f - cg.roughness) * fmx(cg.metallic, cg.transmission);
                float attenuation = mixf(1.0f, 0.0f, preserve * u->glossy_protect);
                for (int dy = -2; dy <= 2; ++dy)
                    for (int dx = -2; dx <= 2; ++dx) {
                        int sxp = std::min(std::max(x + dx * int(step), 0), W - 1);
                        int syp = std::min(std::max(y + dy * int(step), 0), H - 1);
                        const size_t si = size_t(syp) * W + sxp;
                        const float* sc = &in[si * 4];
                        const pt_gbuffer_texel& sg = gbuf[si];
                        float h = kernel[dx + 2] * kernel[dy + 2];
                        f3 c_diff = V(cc[0], cc[1], cc[2]) - V(sc[0], sc[1], sc[2]);
                        float w_color = std::exp(-dot(c_diff, c_diff) / (u->sigma_color * u->sigma_color + 1e-5f));
                        float w_normal = 1.0f;
                        float cl = length(ld3(cg.normal)), sl = length(ld3(sg.normal));
                        if (cl > 0.1f || sl > 0.1f) {
                            float nd = fmx(0.0f, dot(ld3(cg.normal), ld3(sg.normal)));
                            w_normal = std::pow(nd, u->sigma_normal);
                        }
                        float cdm = cg.dist / (cg.dist + 3.0f), sdm = sg.dist / (sg.dist + 3.0f);
                        float d_diff = cdm - sdm;
                        float w_depth = std::exp(-(d_diff * d_diff) / (u->sigma_depth * u->sigma_depth + 1e-5f));
                        f3 a_diff = tonemap(ld3(cg.albedo), u->exposure) - tonemap(ld3(sg.albedo), u->exposure);
                        float w_albedo = std::exp(-dot(a_diff, a_diff) / (u->sigma_albedo * u->sigma_albedo + 1e-5f));
                        f3 m_diff = V(cg.transmission, cg.roughness, cg.metallic) - V(sg.transmission, sg.roughness, sg.metallic);
                        float w_material = std::exp(-dot(m_diff, m_diff) / (u->sigma_material * u->sigma_material + 1e-5f));
                        float weight = h * w_color * w_normal * w_depth * w_albedo * w_material;
                        if (dx != 0 || dy != 0) weight *= attenuation;
                        for (int q = 0; q < 4; ++q) sum[q] = sum[q] + sc[q] * weight;
                        sum_w += weight;
                    }
                for (int q = 0; q < 4; ++q) out[ci * 4 + q] = sum[q] / (sum_w + 1e-5f);
            }
        });
    }
    const std::vector<float>& fin = (u->passes % 2 == 0) ? A : B;
    for (size_t i = 0; i < n; ++i) {
        float fc[4] = {fin[4 * i], fin[4 * i + 1], fin[4 * i + 2], fin[4 * i + 3]};
        bool is_sky = length(gn(i)) < 0.1f;
        if ((u->flags & 2u) && !is_sky) {
            f3 ma = fmx(tonemap(ga(i), u->exposure), V(0.001f));
            fc[0] = fc[0] * ma.x; fc[1] = fc[1] * ma.y; fc[2] = fc[2] * ma.z;
        }
        if (out_rgba) for (int q = 0; q < 4; ++q) out_rgba[4 * i + q] = fc[q];
        if (out_rgba8) for (int q = 0; q < 4; ++q) out_rgba8[4 * i + q] = to_unorm8(fc[q]);
    }
    return PT_OK;
}

static SurfaceContext bsdf_ctx(const pt_material& m, const float n[3]) {
    SurfaceHit h = empty_hit();
    h.hit_n = V(n[0], n[1], n[2]);
    static const ref_scene kNoTextures;
    return get_surface_context(kNoTextures, h, m, material_uv(h, m));
}

int ref_bsdf_sample(const pt_material* m, const float n[3], const float v[3], uint32_t n_samples, uint32_t seed,
                    float* out) {
    if (!m || !n || !v || !out) return PT_ERR_ARG;
    const SurfaceContext ctx = bsdf_ctx(*m, n);
    for (uint32_t i = 0; i < n_samples; ++i) {
        Rng rng{(seed + i * 7919u) | 1u};
        Ray ray{V(v[0], v[1], v[2]), -V(v[0], v[1], v[2])};
        f3 tp = V(1.0f);
        MediumStack stack;
        float last_pdf = 1.0f;
        const bool ok = sample_surface(&ray, &tp, *m, ctx, V(0.f), &stack, &last_pdf, rng);
        float* o = out + size_t(i) * 8;
        o[0] = ray.direction.x; o[1] = ray.direction.y; o[2] = ray.direction.z;
        o[3] = tp.x; o[4] = tp.y; o[5] = tp.z;
        o[6] = last_pdf;
        o[7] = ok ? 1.0f : 0.0f;
    }
    return PT_OK;
}

int ref_bsdf_eval(const pt_material* m, const float n[3], const float v[3], const float* L, uint32_t k, float* f,
                  float* pdf) {
    if (!m || !n || !v || !L || !f || !pdf) return PT_ERR_ARG;
    const SurfaceContext ctx = bsdf_ctx(*m, n);
    MediumStack stack;
    const f3 Vv = V(v[0], v[1], v[2]);
    for (uint32_t i = 0; i < k; ++i) {
        const f3 l = V(L[3 * i], L[3 * i + 1], L[3 * i + 2]);
        const f3 e = eval_surface(Vv, l, *m, ctx, stack);
        f[3 * i] = e.x; f[3 * i + 1] = e.y; f[3 * i + 2] = e.z;
        pdf[i] = pdf_surface(Vv, l, *m, ctx, stack);
    }
    return PT_OK;
}

int ref_env_sample(const ref_scene* s, const pt_scene_params* params, uint32_t n, uint32_t seed, float* out) {
    if (!s || !params || !out) return PT_ERR_ARG;
    for (uint32_t i = 0; i < n; ++i) {
        Rng rng{(seed + i * 7919u) | 1u};
        float ux = rng.next();
        float uy = rng.next();
        LightSample ls = sample_env_cdf(*s, *params, f2{ux, uy});
        float* o = out + size_t(i) * 7;
        o[0] = ls.dir.x; o[1] = ls.dir.y; o[2] = ls.dir.z; o[3] = ls.pdf;
        o[4] = ls.color.x; o[5] = ls.color.y; o[6] = ls.color.z;
    }
    return PT_OK;
}

int ref_env_pdf(const ref_scene* s, const pt_scene_params* params, const float* dirs, uint32_t k, float* pdf) {
    if (!s || !params || !dirs || !pdf) return PT_ERR_ARG;
    for (uint32_t i = 0; i < k; ++i) pdf[i] = get_sky_pdf(*s, *params, V(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
    return PT_OK;
}

int ref_pcg(uint32_t state, uint32_t n, uint32_t* words, float* values) {
    if (!words && !values) return PT_ERR_ARG;
    Rng rng{state};
    for (uint32_t i = 0; i < n; ++i) {
        Rng probe = rng;
        const uint32_t w = rng.next_word();
        if (words) words[i] = w;
        if (values) values[i] = probe.next();
    }
    return PT_OK;
}

int ref_camera_rays(const pt_scene_params* params, uint32_t n, const uint32_t* pixels, int centre, float* rays,
                    float* jitter) {
    if (!params || !pixels || !rays) return PT_ERR_ARG;
    for (uint32_t i = 0; i < n; ++i) {
        Rng rng{0u};
        f2 j;
        const Ray r = camera_ray(*params, pixels[2 * i], pixels[2 * i + 1], centre ? nullptr : &rng, &j);
        float* o = rays + size_t(i) * 6;
        o[0] = r.origin.x; o[1] = r.origin.y; o[2] = r.origin.z;
        o[3] = r.direction.x; o[4] = r.direction.y; o[5] = r.direction.z;
        if (jitter) { jitter[2 * i] = j.x; jitter[2 * i + 1] = j.y; }
    }
    return PT_OK;
}

int ref_bsdf_terms(int term, const float* args, uint32_t n, float* out) {
    if (!args || !out) return PT_ERR_ARG;
    for (uint32_t i = 0; i < n; ++i) {
        const float a = args[3 * i], b = args[3 * i + 1], c = args[3 * i + 2];
        switch (term) {
            case 0: out[i] = D_GGX(a, b); break;                              // (NdotH, alpha2)
            case 1: out[i] = fresnel_dielectric(a, b, c); break;              // (cos, etai, etat)
            case 2: out[i] = G_Smith_GGX_div_NV(a, b, c); break;              // (NdotV, NdotL, alpha2)
            case 3: out[i] = G_Smith_GGX_Uncorrelated_div_NV(a, b, c); break;  // (NdotV, NdotL, alpha2)
            case 4: out[i] = G1_GGX_div_NV(a, b); break;                      // (NdotV, alpha2)
            case 5: out[i] = mis_weight(a, b); break;                         // (pdf_a, pdf_b)
            default: return PT_ERR_ARG;
        }
    }
    return PT_OK;
}

int ref_light_sample(const ref_scene* s, uint32_t light, const float pos[3], uint32_t n, uint32_t seed, float* out) {
    if (!s || !pos || !out || light >= s->lights.size()) return PT_ERR_ARG;
    const pt_light& L = s->lights[light];
    if (L.obj_idx < 0 || size_t(L.obj_idx) >= s->objs.size()) return PT_ERR_ARG;
    const pt_object& lo = s->objs[L.obj_idx];
    const f3 p = V(pos[0], pos[1], pos[2]);
    for (uint32_t i = 0; i < n; ++i) {
        Rng rng{(seed + i * 7919u) | 1u};
        const LightSample ls = sample_light(*s, L, lo, p, rng);
        float* o = out + size_t(i) * 8;
        o[0] = ls.dir.x; o[1] = ls.dir.y; o[2] = ls.dir.z; o[3] = ls.dist; o[4] = ls.pdf;
        o[5] = ls.color.x; o[6] = ls.color.y; o[7] = ls.color.z;
    }
    return PT_OK;
}

int ref_light_pdf(const ref_scene* s, uint32_t light, const float pos[3], const float* dirs, uint32_t k, float* pdf) {
    if (!s || !pos || !dirs || !pdf || light >= s->lights.size()) return PT_ERR_ARG;
    const pt_light& L = s->lights[light];
    if (L.obj_idx < 0 || size_t(L.obj_idx) >= s->objs.size()) return PT_ERR_ARG;
    const pt_object& lo = s->objs[L.obj_idx];
    for (uint32_t i = 0; i < k; ++i) {
        const Ray r{V(pos[0], pos[1], pos[2]), V(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2])};
        Counters c;
        const SurfaceHit h = trace_scene(*s, r, &c);
        pdf[i] = h.o_idx == L.obj_idx ? get_light_pdf(*s, L, lo, r.origin, r.direction, h) : 0.0f;
    }
    return PT_OK;
}

// tracer.js:135-204
int ref_build_sky_cdf(const uint16_t* data, uint32_t W, uint32_t H, float exposure, float* cdf, double* total_weight) {
    if (!data || !cdf || !W || !H) return PT_ERR_ARG;
    const double threshold = 65519.0;
    float* marg = cdf;
    float* cond = cdf + (H + 1);
    double total = 0.0;
    marg[0] = 0.0f;
    for (uint32_t y = 0; y < H; ++y) {
        const double sin_theta = std::sin(3.141592653589793 * (double(y) + 0.5) / double(H));
        float* row = cond + size_t(y) * (W + 1);
        row[0] = 0.0f;
        double acc = 0.0;
        for (uint32_t x = 0; x < W; ++x) {
            const uint16_t* t = data + (size_t(y) * W + x) * 4;
            double rr = std::min(double(half_to_float(t[0])) * exposure, threshold);
            double gg = std::min(double(half_to_float(t[1])) * exposure, threshold);
            double bb = std::min(double(half_to_float(t[2])) * exposure, threshold);
            acc += (0.2126 * rr + 0.7152 * gg + 0.0722 * bb) * sin_theta;
            row[x + 1] = float(acc);
        }
        for (uint32_t x = 1; x <= W; ++x)
            row[x] = acc > 0 ? float(double(row[x]) / acc) : float(double(x) / double(W));
        total += acc;
        marg[y + 1] = float(total);
    }
    for (uint32_t y = 1; y <= H; ++y)
        marg[y] = total > 0 ? float(double(marg[y]) / total) : float(double(y) / double(H));
    if (total_weight) *total_weight = total;
    return PT_OK;
}

}  // extern "C"

// ================================================================ FP64 =====
namespace {
struct d3 { double x, y, z; };
inline d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline d3 operator*(d3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline d3 dcross(d3 a, d3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline d3 dl3(const float* p) { return {p[0], p[1], p[2]}; }
d3 dmul(const float* m, d3 v, double w) {
    return {double(m[0]) * v.x + double(m[4]) * v.y + double(m[8]) * v.z + double(m[12]) * w,
            double(m[1]) * v.x + double(m[5]) * v.y + double(m[9]) * v.z + double(m[13]) * w,
            double(m[2]) * v.x + double(m[6]) * v.y + double(m[10]) * v.z + double(m[14]) * w};
}
double d_tri(d3 o, d3 d, const pt_triangle& t) {
    d3 v0 = dl3(t.v0), e1 = dl3(t.v1) - v0, e2 = dl3(t.v2) - v0;
    d3 h = dcross(d, e2);
    double a = ddot(e1, h);
    if (std::fabs(a) < 1e-12) return -1;
    double f = 1 / a;
    d3 s = o - v0;
    double u = f * ddot(s, h);
    if (u < 0 || u > 1) return -1;
    d3 q = dcross(s, e1);
    double v = f * ddot(d, q);
    if (v < 0 || u + v > 1) return -1;
    double tt = f * ddot(e2, q);
    return tt > 1e-6 ? tt : -1;
}
double d_sphere(d3 o, d3 d) {
    double a = ddot(d, d), hb = ddot(o, d), c = ddot(o, o) - 1, disc = hb * hb - a * c;
    if (disc < 0) return -1;
    double sq = std::sqrt(disc), r = (-hb - sq) / a;
    if (r < 0.001) r = (-hb + sq) / a;
    return r >= 0.001 ? r : -1;
}
double d_cube(d3 o, d3 d) {
    double tn = -INFINITY, tf = INFINITY;
    const double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    for (int k = 0; k < 3; ++k) {
        double inv = 1.0 / dd[k];
        double t0 = (-1 - oo[k]) * inv, t1 = (1 - oo[k]) * inv;
        tn = std::max(tn, std::min(t0, t1));
        tf = std::min(tf, std::max(t0, t1));
    }
    if (tn < tf && tf > 0) return tn > 0.001 ? tn : tf;
    return -1;
}
double d_cyl(d3 o, d3 d, double r1) {
    const double r0 = 1.0, k = r1 - r0, oe = r0 + k * o.y;
    double A = d.x * d.x + d.z * d.z - k * k * d.y * d.y;
    double B = 2 * (o.x * d.x + o.z * d.z - oe * k * d.y);
    double C = o.x * o.x + o.z * o.z - oe * oe;
    double best = 1e300;
    double disc = B * B - 4 * A * C;
    if (disc > 0) {
        double sq = std::sqrt(disc);
        for (double t : {(-B - sq) / (2 * A), (-B + sq) / (2 * A)}) {
            double y = o.y + t * d.y;
            if (t > 0.001 && t < best && y >= 0 && y <= 1) best = t;
        }
    }
    const double caps[2] = {0, 1}, radii[2] = {r0, r1};
    for (int i = 0; i < 2; ++i) {
        double t = (caps[i] - o.y) / d.y;
        if (t > 0.001 && t < best) {
            d3 p = o + d * t;
            if (p.x * p.x + p.z * p.z <= radii[i] * radii[i]) best = t;
        }
    }
    return best < 1e300 ? best : -1;
}
}  // namespace

// The torus (axis y, major radius 1, minor radius r) along o + t d in its
// local frame: the first root of (|p|^2 + 1 - r^2)^2 - 4 (p.x^2 + p.z^2) past
// the reference's 0.001 in local units along the normalised direction,
// bracketed by a sign change on a fine grid over the bounding sphere's chord
// and bisected to double precision (a grazing double root that does not
// change sign is missed, like any tangency).
double d_torus(d3 o, d3 d, double r) {
    const double dl = std::sqrt(ddot(d, d));
    const double b = ddot(o, d) / (dl * dl), c = (ddot(o, o) - (1.0 + r) * (1.0 + r)) / (dl * dl);
    const double disc = b * b - c;
    if (disc <= 0.0) return -1.0;
    const double sq = std::sqrt(disc);
    const double t0 = std::max(-b - sq, 0.001 / dl), t1 = -b + sq;
    if (t1 <= t0) return -1.0;
    auto f = [&](double t) {
        const d3 p = o + d * t;
        const double q = ddot(p, p) + 1.0 - r * r;
        return q * q - 4.0 * (p.x * p.x + p.z * p.z);
    };
    const int N = 8192;
    double a = t0, fa = f(a);
    for (int i = 1; i <= N; ++i) {
        double e = t0 + (t1 - t0) * double(i) / N, fe = f(e);
        if ((fa > 0.0) != (fe > 0.0)) {
            for (int k = 0; k < 200 && e - a > 0.0; ++k) {
                const double m = 0.5 * (a + e), fm = f(m);
                if ((fa > 0.0) != (fm > 0.0)) e = m; else { a = m; fa = fm; }
                if (m == a || m == e) break;
            }
            return 0.5 * (a + e);
        }
        a = e;
        fa = fe;
    }
    return -1.0;
}

extern "C" int ref64_closest(const ref_scene* s, const double* ray, pt_hit* hit, double* t_exact) {
    if (!s || !ray || !hit) return PT_ERR_ARG;
    const ref_scene& S = *s;
    d3 o{ray[0], ray[1], ray[2]}, d{ray[3], ray[4], ray[5]};
    double best = 1e10;
    std::memset(hit, 0, sizeof(*hit));
    hit->o_idx = hit->m_idx = hit->prim_id = hit->plane_idx = -1;
    if (S.has(PT_HAS_PLANES))
        for (size_t i = 0; i < S.planes.size(); ++i) {
            const pt_plane& p = S.planes[i];
            d3 n = dl3(p.normal);
            double den = ddot(n, d);
            if (std::fabs(den) > 1e-6) {
                double t = (double(p.d) - ddot(n, o)) / den;
                if (t > 0.001 && t < best) { best = t; hit->m_idx = p.material_idx; hit->plane_idx = int(i); hit->o_idx = -1; hit->prim_id = -1; }
            }
        }
    for (size_t i = 0; i < S.objs.size(); ++i) {
        const pt_object& ob = S.objs[i];
        const uint32_t type_flag[5] = {PT_HAS_MESHES, PT_HAS_SPHERES, PT_HAS_CUBES, PT_HAS_CYLINDERS, PT_HAS_TORI};
        if (ob.object_type < 0 || ob.object_type > 4 || !S.has(type_flag[ob.object_type])) continue;
        d3 lo = dmul(ob.inv_matrix, o, 1.0), ldir = dmul(ob.inv_matrix, d, 0.0);
        if (ob.object_type == 0) {
            uint32_t no, to;
            std::memcpy(&no, &ob.param0, 4);
            std::memcpy(&to, &ob.param1, 4);
            // triangles reachable from this BLAS root
            std::vector<uint32_t> st{no};
            const uint32_t root = no;
            while (!st.empty()) {
                uint32_t ni = st.back(); st.pop_back();
                const pt_bvh_node& nd = S.bvh[ni];
                if (nd.count > 0) {
                    for (uint32_t k = to + nd.next; k < to + nd.next + nd.count; ++k) {
                        double t = d_tri(lo, ldir, S.tris[k]);
                        if (t > 0 && t < best) { best = t; hit->m_idx = ob.material_idx; hit->o_idx = int(i); hit->prim_id = int(k); hit->plane_idx = -1; }
                    }
                } else {
                    st.push_back(ni + 1);
                    st.push_back(root + nd.next);
                }
            }
            continue;
        }
        double t = -1;
        if (ob.object_type == 1) t = d_sphere(lo, ldir);
        else if (ob.object_type == 2) t = d_cube(lo, ldir);
        else if (ob.object_type == 3) t = d_cyl(lo, ldir, ob.param0);
        else t = d_torus(lo, ldir, ob.param0);
        if (t > (ob.object_type == 4 ? 0.0 : 0.001) && t < best) { best = t; hit->m_idx = ob.material_idx; hit->o_idx = int(i); hit->prim_id = -1; hit->plane_idx = -1; }
    }
    hit->t = float(best);
    if (t_exact) *t_exact = best;
    return PT_OK;
}
