// CPU restatement of Crystal2's PathTracer + MIS path -- TEST INFRASTRUCTURE ONLY.
//
// Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg, as the checker. The product never links or calls it.
//
// Each function restates the reference function cited beside it (paths
// relative to the Crystal2 repository), keeping its IEEE fp32 operation order
// (built with -ffp-contract=off). glm's operation order is restated from
// deps/TREngine/deps/glm (func_geometric.inl, func_matrix.inl, type_mat3x3.inl).
//
// Deliberate, documented substitutes for the reference's non-reproducible
// parts (identical in the GPU path):
//  * Halton dims >= 100 (HaltonSampler.cpp:285-288 draw from a shared
//    std::mt19937) and Blinn-Phong rejection redraws
//    (BlinnPhongReflection.cpp:31-53) use a 32-bit hash of (index, stream).
//    With orc_scene_set_reference_draws(s, 1) and one thread the oracle instead
//    continues the reference's own two generators in the reference's order
//    (SeqDraws below), which reproduces a "Threads": 1 reference render.
//  * Sphere uv (Sphere.cpp:54-55) is not computed: only image textures read it.
//  * Glass refraction that fails after FresnelDielectric reported no total
//    internal reflection (rounding) leaves the direction (0,0,0) instead of an
//    uninitialised vector (SpecularFresnel.cpp:42).
#include "c2_oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local std::string g_err;

// The reference's two file-static, default-seeded std::mt19937 generators:
// HaltonSampler.cpp:116 (consumed first by the permutation-table shuffles,
// :164-178 via :207-211, then by every SampleValue of a dimension >= 100,
// :285-288) and BlinnPhongReflection.cpp:31 (the rejection redraws, :45,52).
// A render with "Threads": 1 runs its 32x32 tiles one after another in
// submission order (SamplerIntegrator.cpp:30-67, FixedThreadPool's FIFO), so
// the sequence of draws is fixed; the oracle reproduces it when it renders on
// one thread with reference draws on. The two pair draws,
// Point2f(SampleValue(d), SampleValue(d + 1)) (GlobalSampler.cpp:25-26) and
// glm::vec2(uniformRandomFloat(mt), uniformRandomFloat(mt))
// (BlinnPhongReflection.cpp:45,52), evaluate their arguments in the build's
// order: left to right under clang, right to left under GCC and MSVC.
struct SeqDraws {
    std::mt19937 halton, phong;
    bool rtl = false;
};
thread_local SeqDraws* g_seq = nullptr;

// ------------------------------------------------------------------ transcendentals
#ifdef ORACLE_LIBM
inline float o_sin(float x) { return std::sin(x); }
inline float o_cos(float x) { return std::cos(x); }
inline float o_tan(float x) { return std::tan(x); }
inline float o_atan(float x) { return std::atan(x); }
inline float o_exp(float x) { return std::exp(x); }
inline float o_log(float x) { return std::log(x); }
inline float o_pow5(float x) { return std::pow(x, 5.f); }
inline float o_atan2(float y, float x) { return std::atan2(y, x); }
inline float o_acos(float x) { return std::acos(x); }
inline float o_powf(float x, float y) { return std::pow(x, y); }
// std::pow(float, int) promotes to double
inline float o_pow5d(float x) { return (float)std::pow((double)x, 5); }
#else
inline float o_sin(float x) { return (float)std::sin((double)x); }
inline float o_cos(float x) { return (float)std::cos((double)x); }
inline float o_tan(float x) { return (float)std::tan((double)x); }
inline float o_atan(float x) { return (float)std::atan((double)x); }
inline float o_exp(float x) { return (float)std::exp((double)x); }
inline float o_log(float x) { return (float)std::log((double)x); }
inline float o_atan2(float y, float x) { return (float)std::atan2((double)y, (double)x); }
inline float o_acos(float x) { return (float)std::acos((double)x); }
inline float o_powf(float x, float y) { return (float)std::pow((double)x, (double)y); }
// x^5 as three double products (x^2 exact), rounded once to float: the GPU's form.
inline float o_pow5(float x) {
    double d = x;
    double d2 = d * d;
    return (float)(d2 * d2 * d);
}
inline float o_pow5d(float x) { return o_pow5(x); }
#endif
// std::pow(float, int) and std::pow(float, double) promote to double in both variants.
inline double o_powd(double x, double y) { return std::pow(x, y); }

constexpr float PI = 3.14159265358979323846264338327950288f;
constexpr float TWO_PI = 6.28318530717958647692528676655900576f;
constexpr float HALF_PI = 1.57079632679489661923132169163975144f;
constexpr float ONE_MINUS_EPS = 0x1.fffffep-1f;
constexpr float EPS_UTIL = 1e-5f;  // src/Core/Utils.h:7
constexpr float EPS_SPAWN = 1e-4f; // src/Core/SurfaceInteraction.cpp:6
// DirectLightingIntegrator: transparent surfaces crossed before the first surface with a
// BSDF. The reference recurses without a bound; the GPU path runs this many wavefront
// steps, and the oracle stops at the same count.
constexpr int kDlChainLimit = 8;

// ------------------------------------------------------------------ glm-order vectors
struct vec3 {
    float x = 0, y = 0, z = 0;
    vec3() = default;
    vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    explicit vec3(float s) : x(s), y(s), z(s) {}
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
struct vec2 {
    float x = 0, y = 0;
    vec2() = default;
    vec2(float a, float b) : x(a), y(b) {}
};
inline vec2 operator+(vec2 a, vec2 b) { return {a.x + b.x, a.y + b.y}; }
inline vec2 operator*(float s, vec2 a) { return {s * a.x, s * a.y}; }
inline vec3 operator+(vec3 a, vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline vec3 operator-(vec3 a, vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline vec3 operator-(vec3 a) { return {-a.x, -a.y, -a.z}; }
inline vec3 operator*(vec3 a, vec3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline vec3 operator*(vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline vec3 operator*(float s, vec3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline vec3 operator/(vec3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline vec3& operator+=(vec3& a, vec3 b) { return a = a + b; }
inline vec3& operator*=(vec3& a, vec3 b) { return a = a * b; }
inline vec3& operator*=(vec3& a, float s) { return a = a * s; }
inline vec3& operator/=(vec3& a, float s) { return a = a / s; }
inline bool operator==(vec3 a, vec3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
inline bool operator!=(vec3 a, vec3 b) { return !(a == b); }
inline bool isZero(vec3 a) { return a == vec3(0.f); }
inline float dot(vec3 a, vec3 b) {  // func_geometric.inl:48-54
    vec3 t = a * b;
    return t.x + t.y + t.z;
}
inline vec3 cross(vec3 a, vec3 b) {  // func_geometric.inl:68-79
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
inline float length(vec3 a) { return std::sqrt(dot(a, a)); }
inline vec3 normalize(vec3 a) { return a * (1.f / std::sqrt(dot(a, a))); }  // :82-90
inline vec3 reflect(vec3 i, vec3 n) { return i - n * dot(n, i) * 2.f; }     // :104-110
inline float smax(float a, float b) { return (a < b) ? b : a; }             // std::max
inline float smin(float a, float b) { return (b < a) ? b : a; }             // std::min
// uniformRandomFloat, src/Core/Utils.h:12-15 (mt() is 32 bits; the product is float)
inline float uniformRandomFloat(std::mt19937& mt) { return smin(ONE_MINUS_EPS, (float)mt() * 2.3283064365386963e-10f); }
inline float clampf(float x, float lo, float hi) { return smin(smax(x, lo), hi); }
inline vec3 fixVector(vec3 v) {  // src/Core/Utils.h:163-166
    return {clampf(v.x, -1.f, 1.f), clampf(v.y, -1.f, 1.f), clampf(v.z, -1.f, 1.f)};
}
inline bool hasNan(vec3 v) { return std::isnan(v.x) || std::isnan(v.y) || std::isnan(v.z); }
inline bool hasInf(vec3 v) { return std::isinf(v.x) || std::isinf(v.y) || std::isinf(v.z); }
inline float square(float x) { return x * x; }

// mat3 with columns c[0..2] (glm column-major); mul per type_mat3x3.inl:468-474
struct mat3 {
    vec3 c[3];
};
inline vec3 mul(const mat3& m, vec3 v) {
    return {m.c[0].x * v.x + m.c[1].x * v.y + m.c[2].x * v.z, m.c[0].y * v.x + m.c[1].y * v.y + m.c[2].y * v.z,
            m.c[0].z * v.x + m.c[1].z * v.y + m.c[2].z * v.z};
}
// glm::inverse for mat3, func_matrix.inl:269-291
inline mat3 inverse(const mat3& M) {
    auto m = [&](int c, int r) { return M.c[c][r]; };
    float OneOverDeterminant = 1.f / (+m(0, 0) * (m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2)) -
                                      m(1, 0) * (m(0, 1) * m(2, 2) - m(2, 1) * m(0, 2)) +
                                      m(2, 0) * (m(0, 1) * m(1, 2) - m(1, 1) * m(0, 2)));
    mat3 I;
    I.c[0][0] = +(m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2)) * OneOverDeterminant;
    I.c[1][0] = -(m(1, 0) * m(2, 2) - m(2, 0) * m(1, 2)) * OneOverDeterminant;
    I.c[2][0] = +(m(1, 0) * m(2, 1) - m(2, 0) * m(1, 1)) * OneOverDeterminant;
    I.c[0][1] = -(m(0, 1) * m(2, 2) - m(2, 1) * m(0, 2)) * OneOverDeterminant;
    I.c[1][1] = +(m(0, 0) * m(2, 2) - m(2, 0) * m(0, 2)) * OneOverDeterminant;
    I.c[2][1] = -(m(0, 0) * m(2, 1) - m(2, 0) * m(0, 1)) * OneOverDeterminant;
    I.c[0][2] = +(m(0, 1) * m(1, 2) - m(1, 1) * m(0, 2)) * OneOverDeterminant;
    I.c[1][2] = -(m(0, 0) * m(1, 2) - m(1, 0) * m(0, 2)) * OneOverDeterminant;
    I.c[2][2] = +(m(0, 0) * m(1, 1) - m(1, 0) * m(0, 1)) * OneOverDeterminant;
    return I;
}

// ------------------------------------------------------------------ geometry
struct Ray {
    vec3 o, d;
};
struct Box {
    vec3 lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
    Box unite(const Box& b) const {  // glm::min / glm::max per component
        return Box{vec3(smin(lo.x, b.lo.x), smin(lo.y, b.lo.y), smin(lo.z, b.lo.z)),
                   vec3(smax(hi.x, b.hi.x), smax(hi.y, b.hi.y), smax(hi.z, b.hi.z))};
    }
    vec3 center() const { return (lo + hi) / 2.f; }
    int maxExtent() const {  // Geometry.h:77-91 (std::abs of extents)
        vec3 v = hi - lo;
        int axis = 0;
        float d = std::abs(v[0]);
        for (int i = 1; i < 3; i++)
            if (std::abs(v[i]) > d) { d = std::abs(v[i]); axis = i; }
        return axis;
    }
};
// RayBoxTest(ray, box, tMin, tMax), src/Core/Geometry.h:118-133: swaps by the sign of the
// direction (the KD-tree's form)
inline bool rayBoxDir(const Ray& r, const Box& b, float& tMin, float& tMax) {
    vec3 invD(1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z);
    bool inv[3] = {r.d.x < 0, r.d.y < 0, r.d.z < 0};
    vec3 minP = (b.lo - r.o) * invD;
    vec3 maxP = (b.hi - r.o) * invD;
    for (int i = 0; i < 3; i++) {
        if (inv[i]) std::swap(minP[i], maxP[i]);
        tMin = smax(tMin, minP[i]);
        tMax = smin(tMax, maxP[i]);
        if (tMax < tMin) return false;
    }
    return true;
}
// RayBoxTest with precomputed inverse, src/Core/Geometry.h:136-157
inline bool rayBox(const Ray& r, const bool inv[3], vec3 invD, const Box& b, float& tMin, float& tMax) {
    vec3 minP = (b.lo - r.o) * invD;
    vec3 maxP = (b.hi - r.o) * invD;
    for (int i = 0; i < 3; i++) {
        if (inv[i]) std::swap(minP[i], maxP[i]);
        tMin = smax(tMin, minP[i]);
        tMax = smin(tMax, maxP[i]);
        if (tMax < tMin) return false;
    }
    return true;
}

// SurfaceInteraction, src/Core/SurfaceInteraction.{h,cpp}
struct Isec {
    int entity = -1;
    float t = INFINITY;
    bool front = false;
    vec3 p, dir, dpdu, dpdv, n;
    vec2 uv;
    // SpawnRay, SurfaceInteraction.cpp:12-16
    Ray spawn(vec3 d) const {
        bool back = dot(d, n) < 0;
        return Ray{p + (back ? -n : n) * EPS_SPAWN, d};
    }
    // SpawnRayTo, :18-23
    Ray spawnTo(vec3 q) const {
        bool back = dot(q - p, n) < 0;
        vec3 o = p + (back ? -n : n) * EPS_SPAWN;
        return Ray{o, q - o};
    }
    vec3 toLocal(vec3 w) const {  // GetInverseTNB() * w
        return {dpdu.x * w.x + dpdu.y * w.y + dpdu.z * w.z, n.x * w.x + n.y * w.y + n.z * w.z,
                dpdv.x * w.x + dpdv.y * w.y + dpdv.z * w.z};
    }
    vec3 toWorld(vec3 l) const {  // GetTNB() * l
        return {dpdu.x * l.x + n.x * l.y + dpdv.x * l.z, dpdu.y * l.x + n.y * l.y + dpdv.y * l.z,
                dpdu.z * l.x + n.z * l.y + dpdv.z * l.z};
    }
};
struct SurfInfo {  // SurfaceInfo, SurfaceInteraction.h:6-17
    vec3 p, n;
    Ray spawn(vec3 d) const { return Ray{p + n * EPS_SPAWN, d}; }  // :58-61
};

// ------------------------------------------------------------------ shapes
struct Vert {
    vec3 p, n;
    vec2 uv;
};
struct Triangle {
    const Vert* v[3];
    vec3 dpdu, dpdv;
    Box box;
    // calculateDerivative, src/Shapes/Triangle.cpp:128-162
    void init() {
        float a00 = v[0]->uv.x - v[2]->uv.x, a01 = v[0]->uv.y - v[2]->uv.y;
        float a10 = v[1]->uv.x - v[2]->uv.x, a11 = v[1]->uv.y - v[2]->uv.y;
        float ood = 1.f / (a00 * a11 - a10 * a01);  // glm mat2 inverse
        float i00 = +a11 * ood, i01 = -a01 * ood, i10 = -a10 * ood, i11 = +a00 * ood;
        if (std::isinf(i00) || std::isinf(i01) || std::isnan(i00) || std::isnan(i01) || std::isnan(i10) ||
            std::isnan(i11)) {
            dpdu = dpdv = vec3(0.f);
        } else {
            vec3 a = v[0]->p - v[2]->p, b = v[1]->p - v[2]->p;
            // (mat2 * mat3x2)[j] = A * (a_j, b_j)
            vec3 r0{i00 * a.x + i10 * b.x, i00 * a.y + i10 * b.y, i00 * a.z + i10 * b.z};
            vec3 r1{i01 * a.x + i11 * b.x, i01 * a.y + i11 * b.y, i01 * a.z + i11 * b.z};
            dpdu = normalize(r0);
            dpdv = normalize(r1);
        }
        vec3 mn = v[0]->p, mx = v[0]->p;
        for (int i = 1; i < 3; i++) {
            mn = vec3(smin(mn.x, v[i]->p.x), smin(mn.y, v[i]->p.y), smin(mn.z, v[i]->p.z));
            mx = vec3(smax(mx.x, v[i]->p.x), smax(mx.y, v[i]->p.y), smax(mx.z, v[i]->p.z));
        }
        box = Box{mn, mx};
    }
    vec3 solve(const Ray& r) const {  // glm::inverse(mat3(e1, e2, -d)) * (o - v0)
        mat3 A{{v[1]->p - v[0]->p, v[2]->p - v[0]->p, -r.d}};
        return mul(inverse(A), r.o - v[0]->p);
    }
    // Triangle::Intersect, Triangle.cpp:32-81
    bool intersect(const Ray& r, Isec* is) const {
        vec3 res = solve(r);
        if (hasNan(res)) return false;
        if (res.x < 0.f || res.x > 1.f || res.y < 0.f || res.y > 1.f || res.x + res.y > 1.0f + EPS_UTIL || res.z < 0.f)
            return false;
        vec3 bary(1 - res.x - res.y, res.x, res.y);
        vec3 N = normalize(bary.x * v[0]->n + bary.y * v[1]->n + bary.z * v[2]->n);
        bool front = dot(r.d, N) < 0;
        N = front ? N : -N;
        vec3 du = dpdu;
        if (isZero(du)) {
            for (int i = 0; i < 3; i++) {
                vec3 e(0.f);
                e[i] = 1;
                vec3 t = cross(e, N);
                if (!isZero(t)) { du = normalize(t); break; }
            }
        }
        is->t = res.z;
        is->p = r.o + r.d * res.z;
        is->dir = r.d;
        is->n = N;
        is->front = front;
        is->dpdu = du;
        is->dpdv = cross(N, du);
        is->uv = bary.x * v[0]->uv + bary.y * v[1]->uv + bary.z * v[2]->uv;  // :43
        return true;
    }
    // Triangle::IntersectTest, Triangle.cpp:83-93
    bool intersectTest(const Ray& r, float tMin, float tMax) const {
        if (tMin > tMax) return false;
        vec3 res = solve(r);
        if (hasNan(res)) return false;
        if (res.x < 0 || res.x > 1 || res.y < 0 || res.y > 1 || res.x + res.y > 1 || res.z < 0) return false;
        return res.z >= tMin && res.z <= tMax;
    }
    float area() const {  // Triangle.cpp:95-98
        return length(cross(v[1]->p - v[0]->p, v[2]->p - v[0]->p)) / 2.f;
    }
    SurfInfo sample(vec2 s) const {  // SampleSurface + sampleTriangle, :100-125
        float u = 1 - std::sqrt(s.x);
        float w = s.y * std::sqrt(s.x);
        vec3 b(1.f - u - w, u, w);
        vec3 p = b.x * v[0]->p + b.y * v[1]->p + b.z * v[2]->p;
        vec3 N = normalize(b.x * v[0]->n + b.y * v[1]->n + b.z * v[2]->n);
        return SurfInfo{p, N};
    }
};

// Sphere, or with `disk` set, Disk (src/Shapes/Disk.cpp: the same record)
struct Sphere {
    vec3 pos;
    float r;
    mat3 l2w, w2l;
    bool disk = false;
    Box box() const {
        if (disk) {  // Disk::GetBoundingBox, Disk.cpp:34-43
            vec3 wn = mul(l2w, vec3(0, 1, 0));
            float sinTheta = std::sqrt(1.f - wn.y * wn.y);
            vec3 off(r, r * sinTheta, r);
            return Box{pos - off, pos + off};
        }
        return Box{pos - vec3(r), pos + vec3(r)};  // Sphere.cpp:31-33
    }
    // Disk::Intersect / IntersectTest, Disk.cpp:45-81: the plane y = 0 of the local frame
    bool diskT(const Ray& ray, vec3& P, vec3& d, float& t) const {
        P = mul(w2l, ray.o - pos);
        d = mul(w2l, ray.d);
        if (d.y == 0.f) return false;
        t = -P.y / d.y;
        return true;
    }
    bool quad(const Ray& ray, vec3& P, vec3& d, float& t1, float& t2) const {
        P = mul(w2l, ray.o - pos);
        d = mul(w2l, ray.d);
        float a = dot(d, d);
        float b = 2 * dot(d, P);
        float c = dot(P, P) - r * r;
        float disc = b * b - 4 * a * c;
        if (disc < 0) return false;
        disc = std::sqrt(disc);
        t1 = (-b + disc) / (2 * a);
        t2 = (-b - disc) / (2 * a);
        return true;
    }
    // Sphere::Intersect, Sphere.cpp:39-75
    bool intersect(const Ray& ray, Isec* is) const {
        if (disk) {
            vec3 P, d;
            float t;
            if (!diskT(ray, P, d, t)) return false;
            if (t < 0) return false;
            vec3 hit = P + d * t;
            if (dot(hit, hit) > r * r) return false;
            vec3 N = mul(l2w, vec3(0, 1, 0));
            bool front = d.y <= 0;
            N = front ? N : -N;
            vec3 hw = mul(l2w, hit) + pos;
            vec3 du = normalize(pos - hw);
            is->t = t;
            is->p = hw;
            is->dir = ray.d;
            is->n = N;
            is->front = front;
            is->dpdu = du;
            is->dpdv = cross(N, du);
            is->uv = vec2(length(hit) / r, (o_atan2(-hit.z, hit.x) + PI) / TWO_PI);  // Disk.cpp:61-62
            return true;
        }
        vec3 P, d;
        float t1, t2;
        if (!quad(ray, P, d, t1, t2)) return false;
        if (t1 > t2) std::swap(t1, t2);
        float t = t1;
        if (std::isnan(t1) || t1 < 0) t = t2;
        if (std::isnan(t) || t < 0) return false;
        vec3 hit = P + d * t;
        vec3 N = normalize(hit);
        float theta = o_atan2(-hit.z, hit.x) / PI;  // Sphere.cpp:57-58
        float phi = o_acos(hit.y / r) / PI + 1.0f;
        bool front = dot(d, N) < 0;
        N = front ? N : -N;
        vec3 real = mul(l2w, hit) + pos;
        vec3 du = normalize(mul(l2w, vec3(-N.z, 0, N.x)));
        N = mul(l2w, N);
        is->t = t;
        is->p = real;
        is->dir = ray.d;
        is->n = N;
        is->front = front;
        is->dpdu = du;
        is->dpdv = cross(N, du);
        is->uv = vec2(theta, phi);
        return true;
    }
    // Sphere::IntersectTest, :77-88
    bool intersectTest(const Ray& ray, float tMin, float tMax) const {
        if (disk) {
            vec3 P, d;
            float t;
            if (!diskT(ray, P, d, t)) return false;
            if (t < tMin || t > tMax) return false;
            vec3 hit = P + d * t;
            return !(dot(hit, hit) > r * r);
        }
        if (tMin > tMax) return false;
        vec3 P, d;
        float t1, t2;
        if (!quad(ray, P, d, t1, t2)) return false;
        return (t1 >= tMin && t1 <= tMax) || (t2 >= tMin && t2 <= tMax);
    }
    float area() const { return disk ? PI * r * r : 4 * PI * r * r; }  // Disk.cpp:83-86
    SurfInfo sample(vec2 s) const {  // NextUnitSphere, Utils.h:138-149
        if (disk) {  // Disk::SampleSurface, Disk.cpp:88-96
            float rr = std::sqrt(s.x);
            float phi = s.y * TWO_PI;
            vec3 dw = mul(l2w, vec3(o_cos(phi) * rr * r, 0.f, -o_sin(phi) * rr * r));
            vec3 nw = mul(l2w, vec3(0, 1, 0));
            return SurfInfo{pos + dw, nw};
        }
        float y = 2.f * s.x - 1.f;
        float phi = s.y * TWO_PI;
        float rr = std::sqrt(1 - y * y);
        vec3 dir(rr * o_cos(phi), y, rr * o_sin(phi));
        return SurfInfo{pos + dir * r, dir};
    }
};

// ------------------------------------------------------------------ BxDFs
enum BxDFType : int { REFL = 1, TRANS = 2, GLOSSY = 4, SPEC = 8, DIFF = 16, ALL = 0x1F };

enum FresnelKind { F_NOOP, F_DIELECTRIC, F_SCHLICK };
struct Fresnel {
    FresnelKind kind;
    vec3 r0;
    // Fresnel.cpp:5-24
    vec3 eval(float etaA, float etaB, float cosI) const {
        if (kind == F_NOOP) return vec3(1.f);
        float sinI = std::sqrt(smax(0.f, 1.f - cosI * cosI));
        float sinT = etaA / etaB * sinI;
        if (sinT >= 1.f) return vec3(1.f);
        if (kind == F_SCHLICK) return r0 + (vec3(1.f) - r0) * o_pow5(1.f - cosI);
        float cosT = std::sqrt(smax(0.f, 1.f - sinT * sinT));
        float Rparl = ((etaB * cosI) - (etaA * cosT)) / ((etaB * cosI) + (etaA * cosT));
        float Rperp = ((etaA * cosI) - (etaB * cosT)) / ((etaA * cosI) + (etaB * cosT));
        return vec3((Rparl * Rparl + Rperp * Rperp) / 2);
    }
};

inline vec2 cosSinPhi(vec3 w, float sinTheta) {  // MicrofacetDistribution.cpp:6-11
    if (sinTheta == 0.f) return vec2(1.f, 0.f);
    return vec2(clampf(w.x / sinTheta, -1, 1), clampf(-w.z / sinTheta, -1, 1));
}
inline vec3 unitFromCos(float cosTheta, float phi) {  // Utils.h:79-83
    float r = std::sqrt(1 - cosTheta * cosTheta);
    return vec3(r * o_cos(phi), cosTheta, -r * o_sin(phi));
}
// GGXDistribution (MicrofacetDistribution.cpp:15-82) or, with beckmann set,
// BeckmannDistribution (:84-161; alpha = roughness as given).
struct GGX {
    vec2 a;
    bool beckmann = false;
    float D(vec3 wh) const {
        if (beckmann) {
            float cos2 = wh.y * wh.y;
            float sin2 = 1.f - cos2;
            float tan2 = sin2 / cos2;
            if (std::isinf(tan2)) return 0.f;
            float cos4 = cos2 * cos2;
            vec2 v = cosSinPhi(wh, std::sqrt(sin2));
            float bot = PI * a.x * a.y * cos4;
            float top = o_exp(-tan2 * (square(v.x) / square(a.x) + square(v.y) / square(a.y)));
            return top / bot;
        }
        float cos2 = wh.y * wh.y;
        float sin2 = smax(0.f, 1.f - cos2);
        vec2 phi = cosSinPhi(wh, std::sqrt(sin2));
        float A = square(phi.x) / square(a.x) + square(phi.y) / square(a.y);
        return 1.f / (PI * a.x * a.y * square(A - cos2 * (A - 1)));
    }
    float lambda(vec3 w) const {
        if (beckmann) {
            float cosTheta = w.y;
            float sinTheta = std::sqrt(1.f - cosTheta * cosTheta);
            float tanTheta = std::abs(sinTheta / cosTheta);
            if (std::isinf(tanTheta)) return 0.f;
            vec2 v = cosSinPhi(w, sinTheta);
            float alpha = std::sqrt(square(v.x) * square(a.x) + square(v.y) * square(a.y));
            float aa = 1.f / (tanTheta * alpha);
            if (aa >= 1.6f) return 0.f;
            return (1 - 1.259f * aa + 0.396f * aa * aa) / (3.535f * aa + 2.181f * aa * aa);
        }
        float cos2 = w.y * w.y;
        float sin2 = 1.f - cos2;
        float tan2 = sin2 / cos2;
        if (std::isinf(tan2)) return 0.f;
        vec2 phi = cosSinPhi(w, std::sqrt(sin2));
        float alpha2 = square(phi.x) * square(a.x) + square(phi.y) * square(a.y);
        return (-1 + std::sqrt(1 + alpha2 * tan2)) / 2;
    }
    float G(vec3 wo, vec3 wi) const { return 1.f / (1.f + lambda(wo) + lambda(wi)); }
    vec3 sampleWh(vec3, vec2 s) const {
        if (beckmann) {
            float logSample = o_log(1.f - s.x);
            if (std::isinf(logSample)) logSample = 0;
            float tan2, phi;
            if (a.x == a.y) {
                tan2 = -square(a.x) * logSample;
                phi = s.y * TWO_PI;
            } else {
                phi = o_atan(a.y / a.x * o_tan(TWO_PI * s.y + HALF_PI));
                if (s.y > 0.5) phi += PI;
                float sinPhi = o_sin(phi), cosPhi = o_cos(phi);
                tan2 = -logSample / (cosPhi * cosPhi / square(a.x) + sinPhi * sinPhi / square(a.y));
            }
            float cosTheta = 1.f / std::sqrt(1.f + tan2);
            return unitFromCos(cosTheta, phi);
        }
        float cosTheta, phi;
        if (a.x == a.y) {
            cosTheta = std::sqrt((1 - s.x) / (s.x * (a.x * a.x - 1) + 1));
            phi = TWO_PI * s.y;
        } else {
            phi = o_atan(a.y / a.x * o_tan(TWO_PI * s.y + HALF_PI));
            if (s.y > 0.5) phi += PI;
            float sinPhi = o_sin(phi), cosPhi = o_cos(phi);
            float A = square(cosPhi) / square(a.x) + square(sinPhi) / square(a.y);
            float cos2 = A * (s.x - 1) / (s.x * (A - 1) - A);
            cosTheta = std::sqrt(cos2);
        }
        return unitFromCos(cosTheta, phi);
    }
    float pdf(vec3, vec3 wh) const { return D(wh) * std::abs(wh.y); }
};

// refract(wo, wh, eta, wt), Utils.h:102-112
inline bool refractH(vec3 wo, vec3 wh, float eta, vec3* wt) {
    float cosO = clampf(dot(wo, wh), -1.f, 1.f);
    float sin2O = smax(0.f, 1.f - cosO * cosO);
    float sin2I = eta * eta * sin2O;
    if (sin2I > 1.f) return false;
    float cosI = std::sqrt(1.f - sin2I);
    *wt = eta * -wo + (eta * cosO - cosI) * wh;
    return true;
}

enum BxKind { BX_LAMBERT, BX_SPECREFL, BX_SPECFRESNEL, BX_MFREFL, BX_MFTRANS, BX_MFBLEND, BX_PHONG, BX_ORENNAYAR,
              BX_FBLEND };

struct BxDF {
    BxKind kind;
    int type;
    vec3 R, T;
    float etaA = 1, etaB = 1;
    Fresnel F{F_NOOP, vec3(0.f)};
    GGX dist{vec2(0, 0)};
    int k = 0;
    uint32_t redrawIdx = 0, redrawDim = 0;  // Blinn-Phong redraw stream
    vec3 S{0.f};               // FresnelBlend: Ks
    float onA = 0, onB = 0;    // OrenNayarReflection: A, B

    // OrenNayarReflection, Reflections.cpp:7-56
    static float sinThetaOf(vec3 w) { return std::sqrt(smax(0.f, 1.f - w.y * w.y)); }
    vec3 orenNayarF(vec3 wo, vec3 wi) const {
        float sinThetaO = sinThetaOf(wo), sinThetaI = sinThetaOf(wi);
        float sinAlpha, tanBeta;
        if (wo.y < wi.y) {
            sinAlpha = sinThetaI;
            tanBeta = sinThetaO / wo.y;
        } else {
            sinAlpha = sinThetaO;
            tanBeta = sinThetaI / wi.y;
        }
        float dCos = 0.f;
        if (sinThetaI > EPS_UTIL && sinThetaO > EPS_UTIL) {
            vec2 dI = cosSinPhi(wi, sinThetaI), dO = cosSinPhi(wo, sinThetaO);
            dCos = smax(0.f, dI.x * dO.x + dI.y * dO.y);
        }
        return R / PI * (onA + onB * (dCos * sinAlpha * tanBeta));
    }
    // FresnelBlend, FresnelBlend.cpp:15-43
    float fblendPdf(vec3 wo, vec3 wi) const {
        vec3 H = normalize(wo + wi);
        return 0.5f * (smax(0.f, wi.y) / PI + dist.pdf(wo, H) / (4 * dot(wo, H)));
    }
    vec3 fblendF(vec3 wo, vec3 wi) const {
        float cosThetaI = smax(0.f, wi.y);
        float cosThetaO = smax(0.f, wo.y);
        vec3 diffuse = (28.f / (23.f * PI)) * R * (vec3(1.f) - S) * (1.f - o_pow5d(1.f - 0.5f * cosThetaO)) *
                       (1.f - o_pow5d(1.f - 0.5f * cosThetaI));
        vec3 H = normalize(wo + wi);
        vec3 schlick = S + (vec3(1.f) - S) * o_pow5(1.f - dot(wi, H));
        vec3 specular = dist.D(H) / (4 * dot(wi, H) * smax(cosThetaI, cosThetaO)) * schlick;
        return diffuse + specular;
    }

    bool matches(int t) const { return (type & t) == type; }
    bool contains(int t) const { return (type & t) != 0; }

    // MicrofacetReflection, Microfacet.cpp:17-55
    float mfReflPdf(vec3 wo, vec3 wi) const {
        vec3 H = normalize(wo + wi);
        return dist.pdf(wo, H) / (4 * dot(wo, H));
    }
    vec3 mfReflF(vec3 wo, vec3 wi) const {
        if (wo.y * wi.y < 0) return vec3(0.f);
        vec3 H = normalize(wo + wi);
        vec3 Fv = F.eval(etaA, etaB, smax(0.f, dot(wo, H)));
        float G = dist.G(wo, wi);
        float D = dist.D(H);
        return R * D * G * Fv / (4 * smax(0.f, wi.y) * smax(0.f, wo.y));
    }
    // MicrofacetTransmission, Microfacet.cpp:57-125
    float mfTransPdf(vec3 wo, vec3 wi) const {
        if (wo.y * wi.y > 0) return 0.f;
        float eta = etaB / etaA;
        vec3 H = normalize(wo + wi * eta);
        float absDot = std::abs(dot(H, wi));
        float dwh_dwi = eta * eta * absDot / square(dot(wo, H) + eta * dot(wi, H));
        return dist.pdf(wo, H) * dwh_dwi;
    }
    vec3 mfTransF(vec3 wo, vec3 wi) const {
        if (wi.y * wo.y > 0) return vec3(0);
        if (wo.y == 0.f) return vec3(0.f);
        float eta = etaB / etaA;
        vec3 H = normalize(wo + wi * eta);
        if (H.y < 0) H = -H;
        float bot = square(dot(wo, H) + eta * dot(wi, H));
        vec3 Fv = vec3(1.f) - F.eval(etaA, etaB, smax(0.f, dot(wo, H)));
        float G = dist.G(wo, wi);
        float D = dist.D(H);
        return R * D * G * Fv * std::abs(dot(H, wo)) * std::abs(dot(H, wi)) /
               (bot * smax(0.f, -wi.y) * smax(0.f, wo.y));
    }
    // BlinnPhongReflection.cpp:14-26
    float phongPdf(vec3 wo, vec3 wi) const {
        vec3 wh = normalize(wo + wi);
        return (float)(((k + 2) / TWO_PI) * o_powd(smax(0.f, wh.y), k + 1));
    }
    vec3 phongF(vec3 wo, vec3 wi) const {
        vec3 wh = normalize(wo + wi);
        float b = smax(0.f, wh.y);
        return R * ((float)o_powd(b, k) * (k + 2) / TWO_PI) / (4 * smax(0.f, wi.y) * smax(0.f, wo.y));
    }

    vec3 f(vec3 wo, vec3 wi) const {
        switch (kind) {
            case BX_LAMBERT: return R / PI;  // Lambertain.cpp:12-15
            case BX_SPECREFL:
            case BX_SPECFRESNEL: return vec3(0.f);
            case BX_MFREFL: return mfReflF(wo, wi);
            case BX_MFTRANS: return mfTransF(wo, wi);
            case BX_MFBLEND: return (wo.y * wi.y < 0) ? mfTransF(wo, wi) : mfReflF(wo, wi);  // :197-207
            case BX_PHONG: return phongF(wo, wi);
            case BX_ORENNAYAR: return orenNayarF(wo, wi);
            case BX_FBLEND: return fblendF(wo, wi);
        }
        return vec3(0.f);
    }
    float pdf(vec3 wo, vec3 wi) const {
        switch (kind) {
            case BX_LAMBERT: return smax(0.f, wi.y) / PI;  // :17-20
            case BX_SPECREFL:
            case BX_SPECFRESNEL: return 0.f;
            case BX_MFREFL: return mfReflPdf(wo, wi);
            case BX_MFTRANS: return mfTransPdf(wo, wi);
            case BX_MFBLEND: return (wo.y * wi.y < 0) ? mfTransPdf(wo, wi) : mfReflPdf(wo, wi);  // :185-195
            case BX_PHONG: return phongPdf(wo, wi);
            case BX_ORENNAYAR: return smax(0.f, wi.y) / PI;  // Reflections.cpp:28-31
            case BX_FBLEND: return fblendPdf(wo, wi);
        }
        return 0.f;
    }
    vec3 sample(vec2 s, vec3 wo, vec3* wi, float* pdfOut, int* sampled) const {
        switch (kind) {
            case BX_LAMBERT: {  // Lambertain.cpp:22-26 + NextCosineUnitHemiSphere, Utils.h:114-123
                float r = std::sqrt(s.x);
                float phi = s.y * TWO_PI;
                float x = r * o_cos(phi), z = r * o_sin(phi);
                float y = std::sqrt(1.0f - r * r);
                *pdfOut = y / PI;
                *wi = vec3(x, y, z);
                return f(wo, *wi);
            }
            case BX_SPECREFL:  // SpecularReflection.cpp:22-27
                *wi = vec3(-wo.x, wo.y, -wo.z);
                *pdfOut = 1.0f;
                return R * F.eval(etaA, etaB, wo.y);
            case BX_SPECFRESNEL: {  // SpecularFresnel.cpp:28-46
                vec3 fr = F.eval(etaA, etaB, wo.y);
                if (s.x < fr.x) {
                    *sampled = REFL | SPEC;
                    *wi = vec3(-wo.x, wo.y, -wo.z);
                    *pdfOut = fr.x;
                    return fr.x * R;
                }
                *sampled = TRANS | SPEC;
                if (!refractH(wo, vec3(0, 1, 0), etaA / etaB, wi)) *wi = vec3(0.f);
                *pdfOut = 1.f - fr.x;
                return (1.f - fr.x) * T;
            }
            case BX_MFREFL: {  // Microfacet.cpp:41-53
                vec3 wh = dist.sampleWh(wo, s);
                if (dot(wo, wh) <= 0) return vec3(0.f);
                *wi = reflect(-wo, wh);
                *pdfOut = mfReflPdf(wo, *wi);
                *wi = fixVector(*wi);
                return mfReflF(wo, *wi);
            }
            case BX_MFTRANS: {  // :103-123
                vec3 wh = dist.sampleWh(wo, s);
                if (dot(wo, wh) <= 0) return vec3(0.f);
                float eta = etaA / etaB;
                if (!refractH(wo, wh, eta, wi)) return vec3(0.f);
                *pdfOut = mfTransPdf(wo, *wi);
                *wi = fixVector(*wi);
                return mfTransF(wo, *wi);
            }
            case BX_MFBLEND: {  // :162-198
                vec3 wh = dist.sampleWh(wo, s);
                if (dot(wo, wh) <= 0) return vec3(0.f);
                float fr = F.eval(etaA, etaB, dot(wo, wh)).x;
                if (s.x < fr) {
                    *sampled = REFL | GLOSSY;
                    *wi = reflect(-wo, wh);
                    *pdfOut = pdf(wo, *wi) * fr;
                    *wi = fixVector(*wi);
                    return mfReflF(wo, *wi);
                }
                float eta = etaA / etaB;
                if (refractH(wo, wh, eta, wi)) {
                    *sampled = TRANS | GLOSSY;
                    *pdfOut = pdf(wo, *wi) * (1 - fr);
                    *wi = fixVector(*wi);
                    return mfTransF(wo, *wi);
                }
                *sampled = REFL | GLOSSY;
                *wi = reflect(-wo, wh);
                *pdfOut = pdf(wo, *wi) * (1 - fr);
                *wi = fixVector(*wi);
                return mfReflF(wo, *wi);
            }
            case BX_ORENNAYAR: {  // Reflections.cpp:58-61 + NextCosineUnitHemiSphere
                float r = std::sqrt(s.x);
                float phi = s.y * TWO_PI;
                float x = r * o_cos(phi), z = r * o_sin(phi);
                float y = std::sqrt(1.0f - r * r);
                *pdfOut = y / PI;
                *wi = vec3(x, y, z);
                return orenNayarF(wo, *wi);
            }
            case BX_FBLEND: {  // FresnelBlend.cpp:45-68
                float u = s.x;
                if (u < 0.5) {
                    s.x = u * 2.f;
                    float r = std::sqrt(s.x);
                    float phi = s.y * TWO_PI;
                    float x = r * o_cos(phi), z = r * o_sin(phi);
                    float y = std::sqrt(1.0f - r * r);
                    *pdfOut = y / PI;
                    *wi = vec3(x, y, z);
                } else {
                    s.x = (u - 0.5f) * 2.f;
                    vec3 wh = dist.sampleWh(wo, s);
                    vec3 L = reflect(-wo, wh);
                    if (L.y * wo.y <= 0) return vec3(0.f);
                    *wi = fixVector(L);
                }
                *pdfOut = fblendPdf(wo, *wi);
                return fblendF(wo, *wi);
            }
            case BX_PHONG: {  // BlinnPhongReflection.cpp:32-58 (deterministic redraws)
                // the reference loops until a draw succeeds; the hash stream stops after 1024 redraws
                for (uint32_t redraw = 0; redraw < (g_seq ? (1u << 24) : 1024u); redraw++) {
                    float cosTheta = (float)o_powd(s.x, 1.0 / (k + 2));
                    float phi = TWO_PI * s.y;
                    vec3 N = unitFromCos(cosTheta, phi);
                    bool ok = dot(N, wo) >= 0;
                    if (ok) {
                        *wi = reflect(-wo, N);
                        ok = wi->y >= 0;
                    }
                    if (!ok) {
                        if (g_seq) {  // the reference's generator, in the build's argument order
                            float a, b;
                            if (g_seq->rtl) { b = uniformRandomFloat(g_seq->phong); a = uniformRandomFloat(g_seq->phong); }
                            else { a = uniformRandomFloat(g_seq->phong); b = uniformRandomFloat(g_seq->phong); }
                            s = vec2(a, b);
                            continue;
                        }
                        uint32_t stream = 0x40000000u | (redrawDim << 12) | (redraw << 1);
                        s = vec2(hashFloat(redrawIdx, stream), hashFloat(redrawIdx, stream + 1));
                        continue;
                    }
                    *pdfOut = pdf(wo, *wi);
                    return f(wo, *wi);
                }
                *pdfOut = 0.f;
                return vec3(0.f);
            }
        }
        return vec3(0.f);
    }
    static float hashFloat(uint32_t a, uint32_t b) {
        uint32_t h = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u) * 0x85EBCA77u;
        h ^= h >> 16;
        h *= 0x7FEB352Du;
        h ^= h >> 15;
        h *= 0x846CA68Bu;
        h ^= h >> 16;
        return smin(ONE_MINUS_EPS, (float)h * 2.3283064365386963e-10f);
    }
};

// BSDF, src/BSDFs/BSDF.cpp:25-130
struct BSDF {
    const Isec* is = nullptr;
    BxDF b[8];
    int n = 0;
    bool empty() const { return n == 0; }
    float pdf(vec3 wo, vec3 wi, int flags) const {
        wo = fixVector(is->toLocal(wo));
        wi = fixVector(is->toLocal(wi));
        float p = 0.f;
        int matched = 0;
        bool reflect = wi.y * wo.y > 0;
        for (int i = 0; i < n; i++) {
            // operator precedence as written at BSDF.cpp:39-40
            if ((b[i].matches(flags) && (reflect && b[i].contains(REFL))) || (!reflect && b[i].contains(TRANS))) {
                p += b[i].pdf(wo, wi);
                matched++;
            }
        }
        return matched > 0 ? p / matched : 0;
    }
    vec3 f(vec3 wo, vec3 wi) const {
        wo = fixVector(is->toLocal(wo));
        wi = fixVector(is->toLocal(wi));
        vec3 L(0.f);
        bool reflect = wi.y * wo.y > 0;
        for (int i = 0; i < n; i++)
            if ((reflect && b[i].contains(REFL)) || (!reflect && b[i].contains(TRANS))) L += b[i].f(wo, wi);
        return L;
    }
    vec3 sample(float u1, vec2 u2, vec3 wo, vec3* wi, float* pdfOut, int flags, int* sampled) const {
        int tot = 0, idMap[10] = {0};
        for (int i = 0; i < n; i++)
            if (b[i].matches(flags)) idMap[tot++] = i;
        if (!tot) return vec3(0.f);
        int v = (int)(smin(0.9999999404f, u1) * tot);
        const BxDF& sel = b[idMap[v]];
        *sampled = sel.type;
        wo = fixVector(is->toLocal(wo));
        vec3 L = sel.sample(u2, wo, wi, pdfOut, sampled);
        if (isZero(L) || *pdfOut == 0.f) return vec3(0.f);
        if (!sel.contains(SPEC)) {
            for (int i = 0; i < n; i++)
                if (&b[i] != &sel && b[i].matches(flags)) *pdfOut += b[i].pdf(wo, *wi);
            bool reflect = wi->y * wo.y > 0;
            for (int i = 0; i < n; i++)
                if (&b[i] != &sel && b[i].matches(flags) &&
                    ((reflect && b[i].contains(REFL)) || (!reflect && b[i].contains(TRANS))))
                    L += b[i].f(wo, *wi);
        }
        *wi = fixVector(is->toWorld(*wi));
        *pdfOut /= tot;
        return L;
    }
};

// ------------------------------------------------------------------ Halton sampler
const int kPrimes100[100] = {
    2,   3,   5,   7,   11,  13,  17,  19,  23,  29,  31,  37,  41,  43,  47,  53,  59,  61,  67,  71,
    73,  79,  83,  89,  97,  101, 103, 107, 109, 113, 127, 131, 137, 139, 149, 151, 157, 163, 167, 173,
    179, 181, 191, 193, 197, 199, 211, 223, 227, 229, 233, 239, 241, 251, 257, 263, 269, 271, 277, 281,
    283, 293, 307, 311, 313, 317, 331, 337, 347, 349, 353, 359, 367, 373, 379, 383, 389, 397, 401, 409,
    419, 421, 431, 433, 439, 443, 449, 457, 461, 463, 467, 479, 487, 491, 499, 503, 509, 521, 523, 541};

const std::vector<std::vector<int>>& permTables() {  // HaltonSampler.cpp:164-178, 207-211
    static const std::vector<std::vector<int>> t = [] {
        std::vector<std::vector<int>> r;
        std::mt19937 mt;
        for (int i = 0; i < 100; i++) {
            std::vector<int> v(kPrimes100[i]);
            for (int j = 0; j < kPrimes100[i]; j++) v[j] = j;
            std::shuffle(v.begin(), v.end(), mt);
            r.push_back(v);
        }
        return r;
    }();
    return t;
}

uint32_t rev32(uint32_t n) {  // ReverseBits32, Utils.h:175-183
    n = (n << 16) | (n >> 16);
    n = ((n & 0x00ff00ff) << 8) | ((n & 0xff00ff00) >> 8);
    n = ((n & 0x0f0f0f0f) << 4) | ((n & 0xf0f0f0f0) >> 4);
    n = ((n & 0x33333333) << 2) | ((n & 0xcccccccc) >> 2);
    n = ((n & 0x55555555) << 1) | ((n & 0xaaaaaaaa) >> 1);
    return n;
}
uint64_t rev64(uint64_t n) { return ((uint64_t)rev32((uint32_t)n) << 32) | rev32((uint32_t)(n >> 32)); }

struct Halton {
    int scales[2], exps[2], stride, inv[2];
    Halton(int W, int H) {  // HaltonSampler.cpp:204-229
        int res[2] = {W, H};
        for (int i = 0; i < 2; i++) {
            int base = i == 0 ? 2 : 3, scale = 1, e = 0;
            while (scale < std::min(res[i], 128)) { scale *= base; ++e; }
            scales[i] = scale;
            exps[i] = e;
        }
        stride = scales[0] * scales[1];
        inv[0] = (int)invMod(scales[1], scales[0]);
        inv[1] = (int)invMod(scales[0], scales[1]);
    }
    static void extgcd(uint64_t a, uint64_t b, uint64_t& d, uint64_t& x, uint64_t& y) {
        if (!b) { d = a; x = 1; y = 0; }
        else { extgcd(b, a % b, d, y, x); y -= x * (a / b); }
    }
    static uint64_t invMod(uint64_t A, uint64_t M) {
        uint64_t d, x, y;
        extgcd(A, M, d, x, y);
        return d == 1 ? (x + M) % M : (uint64_t)-1;
    }
    static uint64_t revDigits(uint64_t v, int base, int n) {  // :240-249
        uint64_t r = 0;
        for (int i = 0; i < n; i++) { r = r * base + (v % base); v = v / base; }
        return r;
    }
    int index(int x, int y, int k) const {  // :251-270 (no % stride, as the reference)
        int64_t off = 0;
        if (stride > 1) {
            int pm[2] = {x % 128, y % 128};
            for (int i = 0; i < 2; i++) {
                uint64_t rem = revDigits((uint64_t)pm[i], i == 0 ? 2 : 3, exps[i]);
                off += rem * (uint64_t)(int64_t)scales[!i] * (uint64_t)(int64_t)inv[i];
            }
        }
        return (int)(off + k * stride);
    }
    float value(int idx, int dim) const {  // :272-291 + RadicalInverse*, :118-162
        if (dim == 0) return (float)((double)rev64((uint64_t)(int64_t)(idx >> exps[0])) * 0x1p-64);
        if (dim >= 100) return g_seq ? uniformRandomFloat(g_seq->halton) : BxDF::hashFloat((uint32_t)idx, (uint32_t)dim);
        int base = kPrimes100[dim];
        float invBase = 1.f / base, result = 0, invN = 1.0f;
        if (dim == 1) {
            uint64_t v = (uint64_t)(int64_t)(idx / scales[1]);
            while (v) {
                uint64_t digit = v % base;
                result = result * base + digit;
                invN *= invBase;
                v /= base;
            }
            return smin(result * invN, ONE_MINUS_EPS);
        }
        const std::vector<int>& perm = permTables()[dim];
        uint64_t v = (uint64_t)(int64_t)idx;
        while (v) {
            uint64_t digit = perm[v % base];
            result = result * base + digit;
            invN *= invBase;
            v /= base;
        }
        return smin((result + invBase * perm[0] / (1 - invBase)) * invN, ONE_MINUS_EPS);
    }
};

// GlobalSampler dimension bookkeeping, src/Samplers/GlobalSampler.cpp:10-69
struct Sampler {
    const Halton* h;
    int idx = 0, dim = 0;
    float get1D() { return h->value(idx, dim++); }
    vec2 get2D() {  // GlobalSampler::Get2D, GlobalSampler.cpp:19-29
        float a, b;
        if (g_seq && g_seq->rtl) { b = h->value(idx, dim + 1); a = h->value(idx, dim); }
        else { a = h->value(idx, dim); b = h->value(idx, dim + 1); }
        dim += 2;
        return vec2(a, b);
    }
};

// ------------------------------------------------------------------ scene + accelerators
enum { SH_TRI = 0, SH_SPHERE = 1 };
struct MatDesc {
    int type, ndf, k;
    vec3 color;
    vec2 rough;
    float ior;
    vec3 ks;
    float sigma;
    int kdTex;  // texture id (1-based), 0 = the constant colour
};
// ImageTexels, src/Textures/Texel.cpp:15-38: texels linearised at load, nearest lookup
struct Texture {
    int w = 0, h = 0;
    std::vector<vec3> data;
    // glm::clamp(uv, 0, .99999f) with glm's (x < y) ? y : x / (y < x) ? y : x; a NaN
    // coordinate (undefined in the reference's int cast) reads column / row 0
    static float clampUV(float x) {
        x = x < 0.f ? 0.f : x;
        x = .99999f < x ? .99999f : x;
        return x == x ? x : 0.f;
    }
    vec3 texel(vec2 uv) const {
        float u = clampUV(uv.x), v = clampUV(uv.y);
        int x = (int)(u * w);
        int y = h - (int)(v * h) - 1;
        return data[(size_t)y * w + x];
    }
};
struct Light {
    int type;  // 0 area, 1 point
    int entity;
    vec3 L, pos;
    float area = 0;
    vec3 flux() const {
        if (type == 1) return L;
        return L * PI * area;  // DiffusedAreaLight.cpp:41-44
    }
};

}  // namespace

struct orc_scene {
    std::vector<Vert> verts;
    std::vector<Triangle> tris;
    std::vector<Sphere> spheres;
    std::vector<int> kind, shape, mat, light;
    std::vector<MatDesc> mats;
    std::vector<Light> lights;
    std::vector<Texture> textures;
    int sky[6] = {0, 0, 0, 0, 0, 0};
    c2_camera cam;
    int accel = ORC_ACCEL_BVH;
    int tie = ORC_TIE_FIRST;
    int integrator = ORC_INTEGRATOR_PATH;
    bool refDraws = false;  // SeqDraws instead of the hash streams (one thread only)
    uint64_t refSkip = 0;   // draws of both generators discarded before the render
    // median-split BVH, src/Accelerators/BVH.cpp:13-34,45-54,156-188,269-355
    struct Node {
        Box box;
        int a = 0, b = 0;  // leaf: start, count | inner: children
        int axis = -1;
    };
    std::vector<Node> nodes;
    std::vector<int> order;
    int root = 0, tot = 0;
    // KDTree, src/Accelerators/KDTree.cpp:8-266 (node 0 = none)
    struct KdNode {
        std::vector<int> objs;
        int ch[2] = {0, 0};
        int axis = -1;
        float split = -1;
    };
    std::vector<KdNode> kd;
    Box kdMaster;
    int kdRoot = 0;

    Box entityBox(int e) const { return kind[e] == SH_TRI ? tris[shape[e]].box : spheres[shape[e]].box(); }

    static float surfaceArea(const Box& b) {  // BoundingBox::SurfaceArea, Geometry.h:101-109
        float area = 0;
        for (int i = 0; i < 3; i++) area += 2 * (b.hi[i] - b.lo[i]) * (b.hi[(i + 1) % 3] - b.lo[(i + 1) % 3]);
        return area;
    }
    // KDTree::_build, KDTree.cpp:97-193
    int kdBuild(const Box& outer, std::vector<int>& objs, int depth) {
        if (objs.size() <= 2 || depth == 20) {
            kd.push_back(KdNode());
            kd.back().objs = objs;
            return (int)kd.size() - 1;
        }
        int split = 0;
        float splitPos = 0;
        float smallestCost = INFINITY;
        float totalSA = surfaceArea(outer);
        vec3 diff = outer.hi - outer.lo;
        struct Edge {
            bool end;
            float pos;
            bool operator<(const Edge& e) const { return pos == e.pos ? end > e.end : pos < e.pos; }
        };
        for (int d = 0; d < 3; d++) {
            std::vector<Edge> edges;
            for (int o : objs) {
                edges.push_back(Edge{false, entityBox(o).lo[d]});
                edges.push_back(Edge{true, entityBox(o).hi[d]});
            }
            int nBelow = 0, nAbove = (int)objs.size();
            float L = outer.lo[d], R = outer.hi[d];
            std::sort(edges.begin(), edges.end());
            for (const Edge& e : edges) {
                if (e.end) --nAbove;
                if (e.pos > L && e.pos < R) {
                    float LA = diff[(d + 1) % 3];
                    float LB = diff[(d + 2) % 3];
                    float belowSA = 2 * (LA * LB + (e.pos - L) * (LA + LB));
                    float aboveSA = 2 * (LA * LB + (R - e.pos) * (LA + LB));
                    float eb = (nAbove == 0 || nBelow == 0) ? 0.2 : 0;
                    float pb = belowSA / totalSA, pa = aboveSA / totalSA;
                    float cost = 0.05 + (1 - eb) * (pb * nBelow + pa * nAbove);
                    if (cost < smallestCost) {
                        smallestCost = cost;
                        split = d;
                        splitPos = e.pos;
                    }
                }
                if (!e.end) ++nBelow;
            }
        }
        if (smallestCost >= objs.size()) {
            kd.push_back(KdNode());
            kd.back().objs = objs;
            return (int)kd.size() - 1;
        }
        Box left = outer, right = outer;
        left.hi[split] = splitPos;
        right.lo[split] = splitPos;
        std::vector<int> lo, hi;
        for (int o : objs) {
            if (entityBox(o).lo[split] <= splitPos) lo.push_back(o);
            if (entityBox(o).hi[split] > splitPos) hi.push_back(o);
        }
        kd.push_back(KdNode());
        int self = (int)kd.size() - 1;
        kd[self].axis = split;
        kd[self].split = splitPos;
        int a = kdBuild(left, lo, depth + 1);
        int b = kdBuild(right, hi, depth + 1);
        kd[self].ch[0] = a;
        kd[self].ch[1] = b;
        return self;
    }
    // KDTree::ray_test, KDTree.cpp:195-238
    bool kdRayTest(int p, const Ray& ray, Isec* info, float tMin, float tMax) const {
        if (!p || tMin > tMax) return false;
        if (tMin >= info->t) return false;
        const KdNode& nd = kd[p];
        if (nd.axis == -1) {
            bool hit = false;
            for (int o : nd.objs) {
                float t1 = tMin, t2 = smin(tMax, info->t);
                if (!rayBoxDir(ray, entityBox(o), t1, t2)) continue;
                Isec tmp;
                if (entityIntersect(o, ray, &tmp)) {
                    if (tmp.t < info->t) {
                        tmp.entity = o;
                        *info = tmp;
                    }
                    hit = true;
                }
            }
            return hit;
        }
        float tSplit = (nd.split - ray.o[nd.axis]) / ray.d[nd.axis];
        bool hit = false;
        int d = ray.d[nd.axis] < 0;
        if (tSplit >= tMin) hit |= kdRayTest(nd.ch[d], ray, info, tMin, smin(smin(tSplit, tMax), info->t));
        if (tSplit <= tMax) hit |= kdRayTest(nd.ch[!d], ray, info, smax(tSplit, 0.0f), smin(tMax, info->t));
        return hit;
    }
    // KDTree::ray_test_p, KDTree.cpp:240-266
    bool kdRayTestP(int p, const Ray& ray, int ignore, float tMin, float tMax) const {
        if (!p || tMin > tMax) return false;
        const KdNode& nd = kd[p];
        if (nd.axis == -1) {
            for (int o : nd.objs) {
                float t1 = tMin, t2 = tMax;
                if (!rayBoxDir(ray, entityBox(o), t1, t2)) continue;
                if (o != ignore && entityTest(o, ray, tMin, tMax)) return true;
            }
            return false;
        }
        float tSplit = (nd.split - ray.o[nd.axis]) / ray.d[nd.axis];
        bool hit = false;
        int d = ray.d[nd.axis] < 0;
        if (tSplit >= tMin) hit |= kdRayTestP(nd.ch[d], ray, ignore, tMin, smin(tSplit, tMax));
        if (hit) return true;
        if (tSplit <= tMax) hit |= kdRayTestP(nd.ch[!d], ray, ignore, smax(tSplit, 0.0f), tMax);
        return hit;
    }

    void build() {
        if (accel == ORC_ACCEL_KDTREE) {  // KDTree::Build, KDTree.cpp:59-71
            kd.assign(1, KdNode());
            kdMaster = Box();
            std::vector<int> all;
            for (int e = 0; e < (int)kind.size(); e++) {
                kdMaster = kdMaster.unite(entityBox(e));
                all.push_back(e);
            }
            kdRoot = kdBuild(kdMaster, all, 0);
            return;
        }
        int n = (int)kind.size();
        order.resize(n);
        for (int i = 0; i < n; i++) order[i] = i;
        nodes.assign((size_t)n * 2 + 1, Node());
        tot = 0;
        if (n > 0) buildRec(root, 0, n - 1);
    }
    void buildRec(int& p, int l, int r) {
        if (r - l + 1 <= 2) {  // createLeaf
            Box box = entityBox(order[l]);
            for (int i = l + 1; i <= r; i++) box = box.unite(entityBox(order[i]));
            p = ++tot;
            nodes[tot].a = l;
            nodes[tot].b = r - l + 1;
            nodes[tot].box = box;
            nodes[tot].axis = -1;
            return;
        }
        Box box = entityBox(order[l]);
        for (int i = l + 1; i <= r; i++) box = box.unite(entityBox(order[i]));
        int axis = box.maxExtent();
        int split = l + (r - l) / 2;
        std::nth_element(order.begin() + l, order.begin() + split, order.begin() + r + 1, [&](int x, int y) {
            return entityBox(x).center()[axis] < entityBox(y).center()[axis];
        });
        p = ++tot;  // createInternal
        nodes[tot].a = nodes[tot].b = 0;
        nodes[tot].box = box;
        nodes[tot].axis = axis;
        int self = tot;
        int c0 = 0, c1 = 0;
        buildRec(c0, l, split);
        buildRec(c1, split + 1, r);
        nodes[self].a = c0;
        nodes[self].b = c1;
    }

    bool entityIntersect(int e, const Ray& r, Isec* is) const {
        return kind[e] == SH_TRI ? tris[shape[e]].intersect(r, is) : spheres[shape[e]].intersect(r, is);
    }
    bool entityTest(int e, const Ray& r, float t0, float t1) const {
        return kind[e] == SH_TRI ? tris[shape[e]].intersectTest(r, t0, t1)
                                 : spheres[shape[e]].intersectTest(r, t0, t1);
    }
    bool better(float dis, int e, const Isec& best) const {
        if (dis < best.t) return true;
        return tie == ORC_TIE_LOWEST && dis == best.t && best.entity >= 0 && e < best.entity;
    }

    // BVH::Intersect, BVH.cpp:56-108 / Brute::Intersect, Brute.cpp:19-40
    bool intersect(const Ray& ray, Isec* info) const {
        bool hit = false;
        if (accel == ORC_ACCEL_KDTREE) {  // KDTree::Intersect, KDTree.cpp:73-77
            float tMin = 0, tMax = INFINITY;
            if (!rayBoxDir(ray, kdMaster, tMin, tMax)) return false;
            return kdRayTest(kdRoot, ray, info, tMin, tMax);
        }
        if (accel == ORC_ACCEL_BRUTE) {
            for (int e = 0; e < (int)kind.size(); e++) {
                Isec tmp;
                if (entityIntersect(e, ray, &tmp)) {
                    if (better(tmp.t, e, *info)) {
                        tmp.entity = e;
                        *info = tmp;
                    }
                    hit = true;
                }
            }
            return hit;
        }
        if (kind.empty()) return false;
        vec3 invDir(1.f / ray.d.x, 1.f / ray.d.y, 1.f / ray.d.z);
        bool neg[3] = {invDir.x < 0, invDir.y < 0, invDir.z < 0};
        Isec isec;
        int stack[64], top = 0, cur = root;
        const float tMin = 0, tMax = INFINITY;
        while (true) {
            const Node& nd = nodes[cur];
            float t1 = tMin, t2 = smin(tMax, info->t);
            if (rayBox(ray, neg, invDir, nd.box, t1, t2)) {
                if (nd.axis == -1) {
                    for (int i = 0; i < nd.b; i++) {
                        int e = order[nd.a + i];
                        t1 = tMin, t2 = smin(tMax, info->t);
                        if (!rayBox(ray, neg, invDir, entityBox(e), t1, t2)) continue;
                        if (!entityIntersect(e, ray, &isec)) continue;
                        float dis = isec.t;
                        if (dis < INFINITY) hit = true;
                        if (better(dis, e, *info)) {
                            isec.entity = e;
                            *info = isec;
                        }
                    }
                    if (!top) break;
                    cur = stack[--top];
                } else {
                    int init = neg[nd.axis] ? 1 : 0;
                    int ch[2] = {nd.a, nd.b};
                    stack[top++] = ch[!init];
                    cur = ch[init];
                }
            } else {
                if (!top) break;
                cur = stack[--top];
            }
        }
        return hit;
    }

    // BVH::IntersectTest, BVH.cpp:110-154 (ignore: entity whose shape is skipped,
    // GeometricEntity::IntersectTest, src/Core/Entities.cpp:26-31)
    bool intersectTest(const Ray& ray, int ignore, float tMin, float tMax) const {
        if (accel == ORC_ACCEL_KDTREE) return kdRayTestP(kdRoot, ray, ignore, tMin, tMax);  // KDTree.cpp:79-82
        if (accel == ORC_ACCEL_BRUTE) {
            for (int e = 0; e < (int)kind.size(); e++) {
                if (e == ignore) continue;
                if (entityTest(e, ray, tMin, tMax)) return true;
            }
            return false;
        }
        if (kind.empty()) return false;
        vec3 invDir(1.f / ray.d.x, 1.f / ray.d.y, 1.f / ray.d.z);
        bool neg[3] = {invDir.x < 0, invDir.y < 0, invDir.z < 0};
        int stack[64], top = 0, cur = root;
        while (true) {
            const Node& nd = nodes[cur];
            float t1 = tMin, t2 = tMax;
            if (rayBox(ray, neg, invDir, nd.box, t1, t2)) {
                if (nd.axis == -1) {
                    for (int i = 0; i < nd.b; i++) {
                        int e = order[nd.a + i];
                        t1 = tMin, t2 = tMax;
                        if (!rayBox(ray, neg, invDir, entityBox(e), t1, t2)) continue;
                        if (e == ignore) continue;
                        if (!entityTest(e, ray, t1, t2)) continue;
                        return true;
                    }
                    if (!top) break;
                    cur = stack[--top];
                } else {
                    int init = neg[nd.axis] ? 1 : 0;
                    int ch[2] = {nd.a, nd.b};
                    stack[top++] = ch[!init];
                    cur = ch[init];
                }
            } else {
                if (!top) break;
                cur = stack[--top];
            }
        }
        return false;
    }

    // Material::ComputeScatteringFunctions for the supported materials
    void buildBSDF(const Isec& is, BSDF& bsdf, int idx, int dimHint) const {
        bsdf.is = &is;
        int m = mat[is.entity];
        if (m < 0) return;
        const MatDesc& md = mats[m];
        BxDF b;
        switch (md.type) {
            case C2_MAT_DIFFUSE:  // DiffuseMaterial.cpp:16-20 (Kd texture: UVTexture::Evaluate)
                b.kind = BX_LAMBERT;
                b.type = DIFF | REFL;
                b.R = md.kdTex ? textures[md.kdTex - 1].texel(is.uv) : md.color;
                break;
            case C2_MAT_MIRROR:  // MirrorMaterial.cpp:11-13
                b.kind = BX_SPECREFL;
                b.type = SPEC | REFL;
                b.R = md.color;
                b.F = Fresnel{F_NOOP, vec3(0.f)};
                b.etaA = b.etaB = 1.f;
                break;
            case C2_MAT_GLASS: {  // Glass.cpp:16-34
                float etaA = 1.0f, etaB = md.ior;
                if (!is.front) std::swap(etaA, etaB);
                b.F = Fresnel{F_DIELECTRIC, vec3(0.f)};
                b.etaA = etaA;
                b.etaB = etaB;
                b.R = b.T = md.color;
                if (md.rough.x == 0.f && md.rough.y == 0.f) {
                    b.kind = BX_SPECFRESNEL;
                    b.type = TRANS | REFL | SPEC;
                } else {
                    b.kind = BX_MFBLEND;
                    b.type = TRANS | REFL | GLOSSY;
                    b.dist = GGX{md.rough};
                }
                break;
            }
            case C2_MAT_MICRO: {  // MicrofacetMaterial.cpp:18-42
                float etaA = 1.f, etaB = md.ior;
                if (!is.front) std::swap(etaA, etaB);
                b.kind = BX_MFREFL;
                b.type = REFL | GLOSSY;
                b.R = md.color;
                b.etaA = etaA;
                b.etaB = etaB;
                b.F = Fresnel{F_SCHLICK, vec3(0.99f)};
                b.dist = GGX{md.rough, md.ndf == C2_NDF_BECKMANN};
                break;
            }
            case C2_MAT_PLASTIC:  // Plastic.cpp:13-27
                b.kind = BX_FBLEND;
                b.type = REFL | GLOSSY;
                b.R = md.color;
                b.S = md.ks;
                b.dist = GGX{md.rough, md.ndf == C2_NDF_BECKMANN};
                break;
            case C2_MAT_ORENNAYAR: {  // DiffuseMaterial.cpp:29-34, Reflections.cpp:18-26
                b.kind = BX_ORENNAYAR;
                b.type = DIFF | REFL;
                b.R = md.color;
                float sigmaSq = md.sigma * md.sigma;
                b.onA = 1.f - sigmaSq / (2 * (sigmaSq + 0.33f));
                b.onB = (0.45f * sigmaSq) / (sigmaSq + 0.09f);
                break;
            }
            case C2_MAT_PHONG:  // PhongMaterial.cpp:13-22
                b.kind = BX_PHONG;
                b.type = GLOSSY | REFL;
                b.R = md.kdTex ? textures[md.kdTex - 1].texel(is.uv) : md.color;
                b.k = md.k;
                break;
            default: return;
        }
        b.redrawIdx = (uint32_t)idx;
        b.redrawDim = (uint32_t)dimHint;
        bsdf.b[bsdf.n++] = b;
    }

    // ImageCubemapTexture::Evaluate + XYZ2CubeUV, CubemapTexture.cpp:24-60
    vec3 skybox(vec3 p) const {
        float ax = std::abs(p.x), ay = std::abs(p.y), az = std::abs(p.z);
        int id;
        vec2 uv;
        if (ax >= ay && ax >= az) {
            id = p.x >= 0 ? 0 : 1;
            uv = vec2(-p.z / p.x, p.y / ax);
        } else if (ay >= az) {
            id = p.y >= 0 ? 2 : 3;
            uv = vec2(p.x / ay, -p.z / p.y);
        } else {
            id = p.z >= 0 ? 4 : 5;
            uv = vec2(p.x / p.z, p.y / az);
        }
        uv.x = uv.x * 0.5f + 0.5f;
        uv.y = uv.y * 0.5f + 0.5f;
        return textures[sky[id] - 1].texel(uv);
    }

    // DiffusedAreaLight::EvalEmitRadiance, DiffusedAreaLight.cpp:11-21
    vec3 emit(const Light& l, vec3 n, vec3 w) const { return dot(n, w) < 0 ? vec3(0.f) : l.L; }
    // SurfaceInteraction::Le, SurfaceInteraction.cpp:52-56
    vec3 Le(const Isec& is, vec3 w) const {
        int li = light[is.entity];
        if (li < 0) return vec3(0.f);
        vec3 n = is.front ? is.n : -is.n;
        return emit(lights[li], n, w);
    }
    // Shape::PdfLight, src/Shapes/Shape.cpp:4-12
    float pdfLight(int e, const SurfInfo& s, vec3 wi) const {
        Ray ray = s.spawn(wi);
        Isec il;
        if (!entityIntersect(e, ray, &il)) return 0.f;
        vec3 dd = s.p - il.p;
        float area = kind[e] == SH_TRI ? tris[shape[e]].area() : spheres[shape[e]].area();
        return dot(dd, dd) / (smax(0.f, dot(-normalize(wi), il.n)) * area);
    }
    // Light::Sample_Li (area: DiffusedAreaLight.cpp:25-34, point: PointLight.cpp:14-20)
    vec3 sampleLi(const Light& l, const SurfInfo& s, vec2 u, vec3* endpoint, float* pdf) const {
        if (l.type == 1) {
            *endpoint = l.pos;
            *pdf = 1.0f;
            vec3 I = l.L / (4 * PI);
            vec3 dd = *endpoint - s.p;
            return I / dot(dd, dd);
        }
        int e = l.entity;
        SurfInfo sl = kind[e] == SH_TRI ? tris[shape[e]].sample(u) : spheres[shape[e]].sample(u);
        *endpoint = sl.p;
        vec3 dir = *endpoint - s.p;
        *pdf = pdfLight(e, s, dir);
        return emit(l, sl.n, -dir);
    }
    static float powerHeuristic(int nf, float fPdf, int ng, float gPdf) {  // Utils.h:168-172
        float f = nf * fPdf, g = ng * gPdf;
        return (f * f) / (f * f + g * g);
    }

    // PathTracingIntegrator::EsimateDirect, PathTracingIntegrator.cpp:188-278
    vec3 estimateDirect(const Isec& is, const BSDF& bsdf, vec2 uL, vec2 uB, int li, Sampler& smp) const {
        const Light& light = lights[li];
        vec3 L(0.f);
        const int sampleType = ALL & ~SPEC;
        vec3 P = is.p, N = is.n, wo = -is.dir;
        SurfInfo surface{is.p, is.n};
        vec3 lightPos;
        float pdfL;
        vec3 Li = sampleLi(light, surface, uL, &lightPos, &pdfL);
        vec3 wi = normalize(lightPos - P);
        bool delta = light.type == 1;
        if (pdfL != 0.f && Li != vec3(0.f)) {
            vec3 f = bsdf.f(wo, wi) * smax(0.f, dot(N, wi));
            float pdfB = bsdf.pdf(wo, wi, sampleType);
            if (f != vec3(0.f)) {
                int ignore = delta ? -1 : light.entity;
                if (intersectTest(is.spawnTo(lightPos), ignore, 0, 1 - EPS_UTIL)) Li = vec3(0.f);
                if (Li != vec3(0.f)) {
                    if (delta) L += f * Li / pdfL;
                    else L += f * Li * powerHeuristic(1, pdfL, 1, pdfB) / pdfL;
                }
            }
        }
        if (!delta) {
            float pdfB = 0;
            int sampled = 0;
            vec3 f = bsdf.sample(smp.get1D(), uB, wo, &wi, &pdfB, sampleType, &sampled);
            bool specB = (sampled & SPEC) != 0;
            if (f == vec3(0.f)) return L;
            float weight = 1.0f;
            if (!specB) {
                f *= smax(0.f, dot(N, wi));
                pdfL = pdfLight(light.entity, surface, wi);
                if (pdfL == 0.f) return L;
                weight = powerHeuristic(1, pdfB, 1, pdfL);
            }
            Isec li2;
            vec3 Lb(0.f);
            if (intersect(is.spawn(wi), &li2)) {
                if (this->light[li2.entity] == li) Lb = Le(li2, -wi);
            }
            if (Lb != vec3(0.f)) L += f * Lb * weight / pdfB;
        }
        return L;
    }

    // PathTracingIntegrator::UniformSampleAllLights, :175-186
    vec3 sampleAllLights(const Isec& is, const BSDF& bsdf, Sampler& smp) const {
        vec3 L(0.f);
        vec2 uL = smp.get2D();
        vec2 uB = smp.get2D();
        int own = light[is.entity];
        for (int li = 0; li < (int)lights.size(); li++) {
            if (lights[li].flux() == vec3(0.f) || li == own) continue;
            L += estimateDirect(is, bsdf, uL, uB, li, smp);
        }
        return L;
    }

    // DirectLightingIntegrator::EsimateDirect, DirectLightingIntegrator.cpp:77-156. Part (i)
    // equals the path tracer's. Part (ii) has no early return for a zero BSDF sample,
    // so a sample that failed inside the BxDF still traces its direction: BSDF::
    // SampleDirection (BSDF.cpp:68-96) then leaves wi in the local frame and pdf at the
    // BxDF's value, and f = 0 adds f * Li * w / pdf = 0, or NaN when that pdf is 0 and
    // the local direction reaches this light (the sample is then discarded by
    // FrameBuffer::AddSample). Two failures leave pdf_bsdf unwritten, and the reference
    // then reads it uninitialised: no non-specular BxDF (SampleDirection returns before
    // writing pdf and sampledType), and a microfacet sample whose half vector faces
    // away or whose refraction fails (Microfacet.cpp:46,108,113-123,167 return before
    // writing). This restatement takes those terms as 0, which is what any finite,
    // nonzero leftover value gives; the reference's own films agree.
    vec3 estimateDirectDL(const Isec& is, const BSDF& bsdf, vec2 uL, vec2 uB, int li, Sampler& smp) const {
        const Light& light = lights[li];
        vec3 L(0.f);
        const int sampleType = ALL & ~SPEC;
        vec3 P = is.p, N = is.n, wo = -is.dir;
        SurfInfo surface{is.p, is.n};
        vec3 lightPos;
        float pdfL;
        vec3 Li = sampleLi(light, surface, uL, &lightPos, &pdfL);
        vec3 wi = normalize(lightPos - P);
        bool delta = light.type == 1;
        if (pdfL != 0.f && Li != vec3(0.f)) {
            vec3 f = bsdf.f(wo, wi) * smax(0.f, dot(N, wi));
            float pdfB = bsdf.pdf(wo, wi, sampleType);
            if (f != vec3(0.f)) {
                int ignore = delta ? -1 : light.entity;
                if (intersectTest(is.spawnTo(lightPos), ignore, 0, 1 - EPS_UTIL)) Li = vec3(0.f);
                if (Li != vec3(0.f)) {
                    if (delta) L += f * Li / pdfL;
                    else L += f * Li * powerHeuristic(1, pdfL, 1, pdfB) / pdfL;
                }
            }
        }
        if (!delta) {
            // sentinels: sampled stays -1 when no BxDF matches, pdfB stays -1 when the
            // BxDF returned before writing its pdf (the two uninitialised reads)
            float pdfB = -1.f;
            int sampled = -1;
            vec3 f = bsdf.sample(smp.get1D(), uB, wo, &wi, &pdfB, sampleType, &sampled);
            if (sampled == -1 || pdfB == -1.f) return L;
            bool specB = (sampled & SPEC) != 0;
            float weight = 1.0f;
            if (!specB) {
                f *= smax(0.f, dot(N, wi));
                pdfL = pdfLight(light.entity, surface, wi);
                if (pdfL == 0.f) return L;
                weight = powerHeuristic(1, pdfB, 1, pdfL);
            }
            Isec li2;
            vec3 Lb(0.f);
            if (intersect(is.spawn(wi), &li2)) {
                if (this->light[li2.entity] == li) Lb = Le(li2, -wi);
            }  // a miss: Light::Le of an area light is 0
            if (Lb != vec3(0.f)) L += f * Lb * weight / pdfB;
        }
        return L;
    }

    // DirectLightingIntegrator::UniformSampleAllLights, :62-73 (the hit's own light included)
    vec3 sampleAllLightsDL(const Isec& is, const BSDF& bsdf, Sampler& smp) const {
        vec3 L(0.f);
        vec2 uL = smp.get2D();
        vec2 uB = smp.get2D();
        for (int li = 0; li < (int)lights.size(); li++) {
            if (lights[li].flux() == vec3(0.f)) continue;
            L += estimateDirectDL(is, bsdf, uL, uB, li, smp);
        }
        return L;
    }

    // DirectLightingIntegrator::Evaluate / eval_rec, :22-58: emission at the hit, then direct
    // light at the first surface with a BSDF. A transparent hit recurses at the same
    // level (no depth is consumed) unless its entity carries an area light. Depth only
    // matters when it is 0 (level == maxDepth at level 0).
    vec3 evaluateDL(Ray ray, Sampler& smp, int maxDepth) const {
        if (maxDepth == 0) return vec3(0.f);
        for (int chain = 0; chain < kDlChainLimit; chain++) {
            Isec is;
            if (!intersect(ray, &is)) return sky[0] ? skybox(ray.d) : vec3(0.f);  // :34-38
            vec3 L = Le(is, -ray.d);
            BSDF bsdf;
            buildBSDF(is, bsdf, smp.idx, smp.dim + 2);  // uB's first dimension
            if (bsdf.empty()) {
                if (light[is.entity] >= 0) return L;
                ray = is.spawn(ray.d);  // eval_rec returns the deeper result alone
                continue;
            }
            return L + sampleAllLightsDL(is, bsdf, smp);
        }
        return vec3(0.f);
    }

    // PathTracingIntegrator::Evaluate, :28-111
    vec3 evaluate(Ray ray, Sampler& smp, int maxDepth, int argOrder) const {
        vec3 beta(1.f), L(0.f);
        bool lightPath = true;
        for (int bounces = 0; bounces < maxDepth; bounces++) {
            Isec is;
            if (!intersect(ray, &is)) {
                if (sky[0]) L += beta * skybox(ray.d);  // :39-46
                break;
            }
            vec3 wo = -ray.d;
            vec3 N = is.n;
            if (lightPath) L += beta * Le(is, -ray.d);
            BSDF bsdf;
            buildBSDF(is, bsdf, smp.idx, smp.dim + 2);  // uB's first dimension
            if (bsdf.empty()) {
                ray = is.spawn(ray.d);
                continue;
            }
            L += beta * sampleAllLights(is, bsdf, smp);
            // sampler->Get1D() and sampler->Get2D() as arguments of one call
            float u1;
            vec2 u2;
            int sampleDim;
            if (argOrder == 0) {
                sampleDim = smp.dim;
                u2 = smp.get2D();
                u1 = smp.get1D();
            } else {
                u1 = smp.get1D();
                sampleDim = smp.dim;
                u2 = smp.get2D();
            }
            bsdf.b[0].redrawDim = (uint32_t)sampleDim;
            vec3 wi;
            float pdf = 0;
            int type = 0;
            vec3 f = bsdf.sample(u1, u2, wo, &wi, &pdf, ALL, &type);
            if (std::abs(pdf) == 0.f || f == vec3(0)) break;
            bool specular = (type & SPEC) != 0;
            float cosine = specular ? 1.0f : smax(0.f, (type & TRANS) ? dot(-N, wi) : dot(N, wi));
            lightPath = specular || (type & TRANS);
            ray = is.spawn(wi);
            beta *= f * cosine / pdf;
            if (bounces > 3) {
                float q = smax(.05f, 1.f - length(beta));
                if (smp.get1D() < q) break;
                beta /= 1 - q;
            }
        }
        return L;
    }
};

namespace {

int setErr(const std::string& m) {
    g_err = m;
    return C2_ERR_INVALID;
}

struct Camera {  // PerspCamera, src/Cameras/PerspCamera.cpp:3-29
    vec3 eye, look, up, right;
    float W, H;
    Camera(const c2_camera& c, int width, int height) {
        float aspect = (float)width / height;
        eye = vec3(c.position[0], c.position[1], c.position[2]);
        vec3 at(c.look_at[0], c.look_at[1], c.look_at[2]);
        vec3 up0(c.up[0], c.up[1], c.up[2]);
        H = 1.0f * std::tan(c.fov / 2);  // host-side float libm, as the reference
        W = H * aspect;
        look = normalize(at - eye);
        vec3 p = dot(up0, look) * look;
        up = normalize(up0 - p);
        right = normalize(cross(look, up0));
    }
    Ray generate(vec2 pos) const {
        pos.x = pos.x * 2 - 1;
        pos.y = pos.y * 2 - 1;
        pos.x *= W, pos.y *= H;
        vec3 dir = normalize(right * pos.x + up * pos.y + look * 1.0f);
        return Ray{eye, dir};
    }
};

}  // namespace

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

int orc_scene_create(const c2_scene_desc* d, int accel, int tie, orc_scene** out) {
    if (!d || !out) return setErr("null argument");
    try {
        auto s = std::make_unique<orc_scene>();
        s->accel = accel;
        s->tie = tie;
        s->verts.resize((size_t)d->num_triangles * 3);
        for (size_t i = 0; i < s->verts.size(); i++) {
            const float* v = d->triangles + i * 8;
            s->verts[i] = Vert{vec3(v[0], v[1], v[2]), vec3(v[3], v[4], v[5]), vec2(v[6], v[7])};
        }
        s->tris.resize(d->num_triangles);
        for (int t = 0; t < d->num_triangles; t++) {
            Triangle& tr = s->tris[t];
            for (int k = 0; k < 3; k++) tr.v[k] = &s->verts[(size_t)t * 3 + k];
            tr.init();
        }
        for (int i = 0; i < d->num_spheres; i++) {
            const float* r = d->spheres + (size_t)i * 22;
            Sphere sp;
            sp.pos = vec3(r[0], r[1], r[2]);
            sp.r = r[3];
            for (int c = 0; c < 3; c++) {
                sp.l2w.c[c] = vec3(r[4 + c * 3], r[5 + c * 3], r[6 + c * 3]);
                sp.w2l.c[c] = vec3(r[13 + c * 3], r[14 + c * 3], r[15 + c * 3]);
            }
            s->spheres.push_back(sp);
        }
        for (int e = 0; e < d->num_entities; e++) {
            if (d->entity_shape_kind[e] == C2_SHAPE_DISK) s->spheres[d->entity_shape_index[e]].disk = true;
            s->kind.push_back(d->entity_shape_kind[e]);
            s->shape.push_back(d->entity_shape_index[e]);
            s->mat.push_back(d->entity_material[e]);
            s->light.push_back(d->entity_light[e]);
        }
        for (int m = 0; m < d->num_materials; m++) {
            const c2_material& cm = d->materials[m];
            s->mats.push_back(MatDesc{cm.type, cm.ndf, cm.phong_k, vec3(cm.color[0], cm.color[1], cm.color[2]),
                                      vec2(cm.roughness[0], cm.roughness[1]), cm.ior,
                                      vec3(cm.ks[0], cm.ks[1], cm.ks[2]), cm.sigma, cm.kd_texture});
        }
        for (int t = 0; t < d->num_textures; t++) {  // ImageTexels ctor, Texel.cpp:28-38
            const c2_texture& ct = d->textures[t];
            Texture tx;
            tx.w = ct.width;
            tx.h = ct.height;
            tx.data.resize((size_t)ct.width * ct.height);
            for (size_t i = 0; i < tx.data.size(); i++) {
                float c[3];
                for (int k = 0; k < 3; k++) c[k] = o_powf((float)ct.rgb[i * 3 + k] / 255.f, 2.2f);
                tx.data[i] = vec3(c[0], c[1], c[2]);
            }
            s->textures.push_back(std::move(tx));
        }
        for (int f = 0; f < 6; f++) s->sky[f] = d->skybox[f];
        for (int l = 0; l < d->num_lights; l++) {
            const c2_light& cl = d->lights[l];
            Light L;
            L.type = cl.type == C2_LIGHT_POINT ? 1 : 0;
            L.entity = cl.entity;
            L.L = vec3(cl.radiance[0], cl.radiance[1], cl.radiance[2]);
            L.pos = vec3(cl.position[0], cl.position[1], cl.position[2]);
            if (L.type == 0) {
                int e = L.entity;
                L.area = s->kind[e] == SH_TRI ? s->tris[s->shape[e]].area() : s->spheres[s->shape[e]].area();
            }
            s->lights.push_back(L);
        }
        s->cam = d->camera;
        s->build();
        *out = s.release();
        return C2_OK;
    } catch (const std::exception& e) {
        return setErr(e.what());
    }
}

void orc_scene_destroy(orc_scene* s) { delete s; }

int orc_scene_set_reference_draws(orc_scene* s, int enable, uint64_t skip) {
    if (!s) return setErr("null scene");
    s->refDraws = enable != 0;
    s->refSkip = skip;
    return C2_OK;
}

int orc_scene_set_integrator(orc_scene* s, int integrator) {
    if (!s || (integrator != ORC_INTEGRATOR_PATH && integrator != ORC_INTEGRATOR_DIRECT)) return setErr("bad integrator");
    s->integrator = integrator;
    return C2_OK;
}

static int renderImpl(orc_scene* s, int W, int H, int spp, int depth, int argOrder, int threads, int y0, int y1,
                      float* film, float* samples, double* seconds) {
    if (!s || W <= 0 || H <= 0 || spp <= 0) return setErr("bad arguments");
    if (s->refDraws && threads != 1) return setErr("reference draws need one thread");
    SeqDraws seq;
    if (s->refDraws) {  // the shuffles of GeneratePermutationTables come first (HaltonSampler.cpp:207-211)
        for (int i = 0; i < 100; i++) {
            std::vector<int> v(kPrimes100[i]);
            for (int j = 0; j < kPrimes100[i]; j++) v[j] = j;
            std::shuffle(v.begin(), v.end(), seq.halton);
        }
        seq.halton.discard(s->refSkip);
        seq.phong.discard(s->refSkip);
        seq.rtl = argOrder == 0;
    }
    y0 = std::max(0, y0);
    y1 = std::min(H, y1);
    Halton halton(W, H);
    Camera cam(s->cam, W, H);
    struct Pixel {
        vec3 c;
        float w = 0;
    };
    std::vector<Pixel> pix((size_t)W * H);
    const int TILE = 32;  // SamplerIntegrator.cpp:25-26
    int ntx = (W + TILE - 1) / TILE, nty = (H + TILE - 1) / TILE;
    std::atomic<int> next(0);
    auto worker = [&]() {
        g_seq = s->refDraws ? &seq : nullptr;
        while (true) {
            int task = next.fetch_add(1);
            if (task >= ntx * nty) break;
            int x = task / nty, y = task % nty;
            int xa = TILE * x, xb = std::min(xa + TILE, W), ya = TILE * y, yb = std::min(ya + TILE, H);
            ya = std::max(ya, y0);
            yb = std::min(yb, y1);
            for (int xx = xa; xx < xb; xx++)
                for (int yy = ya; yy < yb; yy++) {
                    for (int k = 0; k < spp; k++) {
                        Sampler smp{&halton, halton.index(xx, yy, k), 0};
                        vec2 sv = smp.get2D();
                        vec2 pos((float)xx + sv.x, (float)yy + sv.y);
                        pos.x = pos.x / W;
                        pos.y = pos.y / H;
                        vec3 c = s->integrator == ORC_INTEGRATOR_DIRECT ? s->evaluateDL(cam.generate(pos), smp, depth)
                                                                         : s->evaluate(cam.generate(pos), smp, depth, argOrder);
                        if (samples) {
                            float* o = samples + (((size_t)yy * W + xx) * spp + k) * 3;
                            o[0] = c.x; o[1] = c.y; o[2] = c.z;
                        }
                        // FrameBuffer::AddSample, FrameBuffer.cpp:20-32
                        if (hasNan(c)) c = vec3(0);
                        if (hasInf(c)) c = vec3(0);
                        Pixel& p = pix[(size_t)yy * W + xx];
                        p.c += c * 1.f;
                        p.w += 1.f;
                    }
                }
        }
    };
    auto t0 = std::chrono::steady_clock::now();
    int nt = std::max(1, threads);
    std::vector<std::thread> pool;
    for (int i = 0; i < nt; i++) pool.emplace_back(worker);
    for (auto& t : pool) t.join();
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    if (film) {  // GetImageDataRGB32F, FrameBuffer.cpp:60-75 (rows flipped)
        for (int yy = y0; yy < y1; yy++)
            for (int xx = 0; xx < W; xx++) {
                const Pixel& p = pix[(size_t)yy * W + xx];
                vec3 c = p.c / p.w;
                float* o = film + ((size_t)(H - yy - 1) * W + xx) * 3;
                o[0] = c.x; o[1] = c.y; o[2] = c.z;
            }
    }
    return C2_OK;
}

int orc_render(orc_scene* s, int W, int H, int spp, int depth, int argOrder, int threads, int y0, int y1,
               float* film, double* seconds) {
    return renderImpl(s, W, H, spp, depth, argOrder, threads, y0, y1, film, nullptr, seconds);
}

int orc_render_samples(orc_scene* s, int W, int H, int spp, int depth, int argOrder, int threads,
                       float* samples) {
    return renderImpl(s, W, H, spp, depth, argOrder, threads, 0, H, nullptr, samples, nullptr);
}

int orc_intersect(orc_scene* s, const float* rays, int64_t n, int32_t* prim, float* t) {
    if (!s || (!rays && n)) return setErr("null argument");
    for (int64_t i = 0; i < n; i++) {
        const float* r = rays + i * 8;
        Ray ray{vec3(r[0], r[1], r[2]), vec3(r[3], r[4], r[5])};
        Isec is;
        bool hit = s->intersect(ray, &is);
        prim[i] = hit && is.entity >= 0 ? is.entity : -1;
        t[i] = hit && is.entity >= 0 ? is.t : INFINITY;
    }
    return C2_OK;
}

int orc_intersect_test(orc_scene* s, const float* rays, const int32_t* ignore, int64_t n, uint8_t* hit) {
    if (!s || (!rays && n)) return setErr("null argument");
    for (int64_t i = 0; i < n; i++) {
        const float* r = rays + i * 8;
        Ray ray{vec3(r[0], r[1], r[2]), vec3(r[3], r[4], r[5])};
        hit[i] = s->intersectTest(ray, ignore ? ignore[i] : -1, r[6], r[7]) ? 1 : 0;
    }
    return C2_OK;
}

int orc_halton(int W, int H, const int32_t* q, int64_t n, int ndims, int32_t* index, float* values) {
    if (W <= 0 || H <= 0 || (!q && n)) return setErr("bad arguments");
    Halton h(W, H);
    for (int64_t i = 0; i < n; i++) {
        int idx = h.index(q[i * 3], q[i * 3 + 1], q[i * 3 + 2]);
        index[i] = idx;
        for (int d = 0; d < ndims; d++) values[i * ndims + d] = h.value(idx, d);
    }
    return C2_OK;
}

}  // extern "C"
