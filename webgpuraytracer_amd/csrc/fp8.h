// fp8.h — host encoding of OCP e4m3 (the "fn" variant: exponent bias 7,
// 3 mantissa bits, no infinities, S.1111.111 = NaN, largest finite 448).
// Rounds to nearest even and saturates finite values at +-448 — the result
// of clamping to +-448 and then converting with v_cvt_pk_fp8_f32, which is
// how the device quantizes activations.
#pragma once
#include <cmath>
#include <cstdint>

namespace pt {

inline uint8_t e4m3_encode(float x) {
    const uint8_t sign = std::signbit(x) ? 0x80u : 0x00u;
    const float a = std::fabs(x);
    if (std::isnan(a)) return sign | 0x7fu;
    if (a >= 448.0f) return sign | 0x7eu;
    if (a < 0.015625f) {  // below 2^-6: subnormal steps of 2^-9 (8 rounds up to the smallest normal, code 8)
        return sign | uint8_t(std::nearbyint(a * 512.0f));
    }
    int e;
    std::frexp(a, &e);  // a = f * 2^e, f in [0.5, 1): a in [2^(e-1), 2^e)
    int E = e - 1;
    int m = int(std::nearbyint(std::ldexp(a, 3 - E)));  // 8 .. 16
    if (m == 16) {
        m = 8;
        ++E;
    }
    return sign | uint8_t(((E + 7) << 3) | (m - 8));  // a < 448 keeps E + 7 <= 15 and the code below 0x7f
}

inline float e4m3_decode(uint8_t q) {
    const float s = (q & 0x80u) ? -1.0f : 1.0f;
    const int E = (q >> 3) & 15, m = q & 7;
    if (E == 15 && m == 7) return NAN;
    if (E == 0) return s * std::ldexp(float(m), -9);
    return s * std::ldexp(float(8 + m), E - 10);
}

}  // namespace pt
