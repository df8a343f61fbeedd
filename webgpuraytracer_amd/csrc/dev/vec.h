// vec.h — f32 vector arithmetic for the gfx950 kernels.
//
// The kernels are compiled with -ffp-contract=off: every operator below is
// one IEEE-rounded f32 operation, and compound expressions are written in
// WGSL's left-to-right order, so that intersection distances and hit
// identities come out bit-identical to a CPU evaluation of the same WGSL
// expression.  min/max are fminf/fmaxf (v_min_f32 / v_max_f32: the non-NaN
// operand wins).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pt {

struct v2 { float x, y; };
struct v3 { float x, y, z; };

__device__ __forceinline__ v3 mk(float x, float y, float z) { return v3{x, y, z}; }
__device__ __forceinline__ v3 splat(float s) { return v3{s, s, s}; }
__device__ __forceinline__ v3 operator+(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ v3 operator-(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ v3 operator*(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ v3 operator/(v3 a, v3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
__device__ __forceinline__ v3 operator*(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ v3 operator*(float s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ v3 operator/(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ v3 operator-(v3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ v3 cross(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float len(v3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ v3 normalize(v3 a) { return a / len(a); }
__device__ __forceinline__ float mn(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ float mx(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ v3 vmin(v3 a, v3 b) { return mk(mn(a.x, b.x), mn(a.y, b.y), mn(a.z, b.z)); }
__device__ __forceinline__ v3 vmax(v3 a, v3 b) { return mk(mx(a.x, b.x), mx(a.y, b.y), mx(a.z, b.z)); }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return mn(mx(x, lo), hi); }
__device__ __forceinline__ float hmax(v3 v) { return mx(mx(v.x, v.y), v.z); }
__device__ __forceinline__ float hmin(v3 v) { return mn(mn(v.x, v.y), v.z); }
__device__ __forceinline__ float lerp(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ v3 lerp(v3 a, v3 b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ float sgn(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }
__device__ __forceinline__ v3 reflect(v3 i, v3 n) { return i - (2.0f * dot(n, i)) * n; }
__device__ __forceinline__ v3 refract(v3 i, v3 n, float eta) {
    const float d = dot(n, i);
    const float k = 1.0f - eta * eta * (1.0f - d * d);
    if (k < 0.0f) return splat(0.0f);
    return eta * i - (eta * d + sqrtf(k)) * n;
}
__device__ __forceinline__ v3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }

// Column-major 4x4 (16 floats) products, column terms summed in order.
__device__ __forceinline__ v3 xform_point(const float* m, v3 v) {
    return mk(m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * 1.0f,
              m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * 1.0f,
              m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * 1.0f);
}
__device__ __forceinline__ v3 xform_dir(const float* m, v3 v) {
    return mk(m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * 0.0f,
              m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * 0.0f,
              m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * 0.0f);
}
// transpose(M) * vec4(v, 0)
__device__ __forceinline__ v3 xform_tdir(const float* m, v3 v) {
    return mk(m[0] * v.x + m[1] * v.y + m[2] * v.z + m[3] * 0.0f,
              m[4] * v.x + m[5] * v.y + m[6] * v.z + m[7] * 0.0f,
              m[8] * v.x + m[9] * v.y + m[10] * v.z + m[11] * 0.0f);
}
// transpose(mat3(M)) * v
__device__ __forceinline__ v3 xform_normal(const float* m, v3 v) {
    return mk(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
              m[8] * v.x + m[9] * v.y + m[10] * v.z);
}
// mat3(M) * v
__device__ __forceinline__ v3 xform_mat3(const float* m, v3 v) {
    return mk(m[0] * v.x + m[4] * v.y + m[8] * v.z, m[1] * v.x + m[5] * v.y + m[9] * v.z,
              m[2] * v.x + m[6] * v.y + m[10] * v.z);
}

constexpr float kPI = 3.14159265359f;
constexpr float kTWO_PI = 6.28318530718f;
constexpr float kINF_DIST = 65504.0f;

// PCG-RXS-M-XS step and output (shader.wgsl:147-154).
struct Pcg {
    uint32_t s;
    __device__ __forceinline__ float next() {
        const uint32_t st = s;
        s = st * 747796405u + 2891336453u;
        uint32_t w = ((st >> ((st >> 28u) + 4u)) ^ st) * 277803737u;
        w = (w >> 22u) ^ w;
        return float(w) * 2.3283064365386963e-10f;
    }
};

__device__ __forceinline__ uint8_t unorm8(float x) {
    const float v = mn(mx(x, 0.0f), 1.0f);
    return uint8_t(v * 255.0f + 0.5f);
}

}  // namespace pt
