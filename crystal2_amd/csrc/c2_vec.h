// Float vector math in the reference's operation order (glm 0.9.9.8, SIMD off).
//
// Every helper evaluates exactly the sequence of IEEE fp32 operations glm
// evaluates (left-to-right sums, no fused multiply-add: this header is always
// compiled with -ffp-contract=off), so host and device results are bit-identical
// to the reference's:
//   dot       glm/detail/func_geometric.inl:48-54  ((x*x' + y*y') + z*z')
//   cross     glm/detail/func_geometric.inl:68-79
//   normalize glm/detail/func_geometric.inl:82-90  (v * (1 / sqrt(dot(v, v))))
//   reflect   glm/detail/func_geometric.inl:104-110 (I - N * dot(N, I) * 2)
//   clamp/min/max  glm/detail/func_common.inl      ((y < x) ? y : x, NaN passes through)
//
// Transcendentals (sin, cos, tan, atan, acos, atan2, pow) are evaluated in
// double precision and rounded once to float. The double routines are accurate
// to about 1 ulp(double), so the float result is the correctly rounded value
// except when the exact result lies within ~2^-29 relative of a float rounding
// boundary; host (glibc) and device (ROCm ocml) therefore agree bit-for-bit in
// all but such cases, and agree with glibc's float functions wherever those
// are correctly rounded.
#pragma once
#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define C2_HD __host__ __device__ __forceinline__
#else
#define C2_HD inline
#endif

namespace c2 {

struct V3 {
    float x, y, z;
};
struct V2 {
    float x, y;
};

C2_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
C2_HD V3 v3(float s) { return V3{s, s, s}; }
C2_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
C2_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
C2_HD V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
C2_HD V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
C2_HD V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
C2_HD V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
C2_HD V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
C2_HD V3 operator/(V3 a, V3 b) { return V3{a.x / b.x, a.y / b.y, a.z / b.z}; }
C2_HD bool operator==(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
C2_HD bool operator!=(V3 a, V3 b) { return !(a == b); }
C2_HD bool is_zero(V3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }

C2_HD float dot(V3 a, V3 b) {
    float x = a.x * b.x, y = a.y * b.y, z = a.z * b.z;
    return x + y + z;
}
C2_HD V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
C2_HD float length(V3 a) { return sqrtf(dot(a, a)); }
C2_HD V3 normalize(V3 a) { return a * (1.f / sqrtf(dot(a, a))); }
C2_HD V3 reflect(V3 i, V3 n) { return i - n * dot(n, i) * 2.f; }
// std::max / std::min / glm::max / glm::min: first argument wins unless the
// comparison says otherwise (NaN in the second argument is ignored).
C2_HD float fmax_ref(float a, float b) { return (a < b) ? b : a; }
C2_HD float fmin_ref(float a, float b) { return (b < a) ? b : a; }
C2_HD float clampf_ref(float x, float lo, float hi) { return fmin_ref(fmax_ref(x, lo), hi); }
C2_HD V3 clamp_unit(V3 v) {  // fixVector, src/Core/Utils.h:163-166
    return V3{clampf_ref(v.x, -1.f, 1.f), clampf_ref(v.y, -1.f, 1.f), clampf_ref(v.z, -1.f, 1.f)};
}
C2_HD bool any_nan(V3 v) { return v.x != v.x || v.y != v.y || v.z != v.z; }
C2_HD bool is_inf(float x) { return x == INFINITY || x == -INFINITY; }
C2_HD bool any_inf(V3 v) { return is_inf(v.x) || is_inf(v.y) || is_inf(v.z); }

// glm::pi<float>(), two_pi, half_pi as float constants.
constexpr float kPi = 3.14159265358979323846264338327950288f;
constexpr float kTwoPi = 6.28318530717958647692528676655900576f;
constexpr float kHalfPi = 1.57079632679489661923132169163975144f;
constexpr float kOneMinusEpsilon = 0x1.fffffep-1f;

C2_HD float sin_r(float x) { return (float)sin((double)x); }
C2_HD float cos_r(float x) { return (float)cos((double)x); }
C2_HD float tan_r(float x) { return (float)tan((double)x); }
C2_HD float atan_r(float x) { return (float)atan((double)x); }
C2_HD float acos_r(float x) { return (float)acos((double)x); }
C2_HD float atan2_r(float y, float x) { return (float)atan2((double)y, (double)x); }
C2_HD float exp_r(float x) { return (float)exp((double)x); }
C2_HD float log_r(float x) { return (float)log((double)x); }
C2_HD float powf_r(float x, float y) { return (float)pow((double)x, (double)y); }
// x^5 for a float x: x^2 is exact in double, the two further products round once
// each (relative error < 2^-51), then one rounding to float. Deterministic IEEE
// operations, so host and device agree bit for bit.
C2_HD float pow5_r(float x) {
    double d = x;
    double d2 = d * d;
    return (float)(d2 * d2 * d);
}
C2_HD double powd(double x, double y) { return pow(x, y); }

}  // namespace c2
