// Device shading: Halton sampler, camera, surface interaction, BxDFs, lights.
// Included by c2_device.hip only. Each function restates the reference function
// cited beside it in fp32 with the reference's operation order.
#pragma once
#include "c2_device.h"

namespace c2 {

// ------------------------------------------------------------------ Halton
struct HaltonDev {
    int32_t exp0;      // base exponent for x (base 2)
    int32_t scale1;    // base scale for y (base 3)
};

__device__ __forceinline__ uint32_t hash_u32_dev(uint32_t a, uint32_t b) {
    uint32_t h = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u) * 0x85EBCA77u;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return h;
}
__device__ __forceinline__ float hash_float(uint32_t a, uint32_t b) {
    float u = (float)hash_u32_dev(a, b) * 2.3283064365386963e-10f;
    return fmin_ref(kOneMinusEpsilon, u);
}

__constant__ int32_t c_primes[100] = {
    2,   3,   5,   7,   11,  13,  17,  19,  23,  29,  31,  37,  41,  43,  47,  53,  59,  61,  67,  71,
    73,  79,  83,  89,  97,  101, 103, 107, 109, 113, 127, 131, 137, 139, 149, 151, 157, 163, 167, 173,
    179, 181, 191, 193, 197, 199, 211, 223, 227, 229, 233, 239, 241, 251, 257, 263, 269, 271, 277, 281,
    283, 293, 307, 311, 313, 317, 331, 337, 347, 349, 353, 359, 367, 373, 379, 383, 389, 397, 401, 409,
    419, 421, 431, 433, 439, 443, 449, 457, 461, 463, 467, 479, 487, 491, 499, 503, 509, 521, 523, 541};

// HaltonSampler::SampleValue, src/Samplers/HaltonSampler.cpp:272-291 with
// RadicalInverse / RadicalInverseSpecialized / RadicalInverseScrambled (:118-162).
// The vector index is a non-negative int, so the uint64 digit loop runs on 32 bits.
__device__ __forceinline__ float halton_value(const DevScene& S, const HaltonDev& h, int32_t idx, int32_t dim) {
    if (dim == 0) {
        uint32_t v = (uint32_t)(idx >> h.exp0);
        // ReverseBits64(v) * 0x1p-64 for v < 2^32: bit-reversed 32-bit value / 2^32
        return (float)((double)__brev(v) * 0x1p-32);
    }
    if (dim >= 100) return hash_float((uint32_t)idx, (uint32_t)dim);
    uint32_t base = (uint32_t)c_primes[dim];
    float invBase = 1.f / (float)base;
    float result = 0.f, invN = 1.f;
    if (dim == 1) {
        uint32_t v = (uint32_t)(idx / h.scale1);
        while (v) {
            uint32_t q = v / base;
            uint32_t digit = v - q * base;
            result = result * (float)base + (float)digit;
            invN *= invBase;
            v = q;
        }
        return fmin_ref(result * invN, kOneMinusEpsilon);
    }
    const uint16_t* perm = S.perm + S.perm_off[dim];
    uint32_t v = (uint32_t)idx;
    while (v) {
        uint32_t q = v / base;
        uint32_t digit = perm[v - q * base];
        result = result * (float)base + (float)digit;
        invN *= invBase;
        v = q;
    }
    return fmin_ref((result + invBase * (float)perm[0] / (1 - invBase)) * invN, kOneMinusEpsilon);
}

// ------------------------------------------------------------------ interaction
struct Inter {
    V3 p, n, dpdu, dpdv, dir;  // n: shading normal flipped to face the ray
    bool front;
    int32_t entity, material, light;
};

// World <-> local (y-up TNB frame), SurfaceInteraction::GetTNB / GetInverseTNB
__device__ __forceinline__ V3 to_local(const Inter& it, V3 w) {
    return V3{it.dpdu.x * w.x + it.dpdu.y * w.y + it.dpdu.z * w.z, it.n.x * w.x + it.n.y * w.y + it.n.z * w.z,
              it.dpdv.x * w.x + it.dpdv.y * w.y + it.dpdv.z * w.z};
}
__device__ __forceinline__ V3 to_world(const Inter& it, V3 l) {
    return V3{it.dpdu.x * l.x + it.n.x * l.y + it.dpdv.x * l.z, it.dpdu.y * l.x + it.n.y * l.y + it.dpdv.y * l.z,
              it.dpdu.z * l.x + it.n.z * l.y + it.dpdv.z * l.z};
}
// SurfaceInteraction::SpawnRay / SpawnRayTo, SurfaceInteraction.cpp:12-23 (EPS = 1e-4)
__device__ __forceinline__ V3 spawn_origin(V3 p, V3 n, V3 dir) {
    bool back = dot(dir, n) < 0;
    return p + (back ? -n : n) * 1e-4f;
}

// Triangle::Intersect hit construction, Triangle.cpp:41-74
__device__ __forceinline__ void tri_frame(const DevScene& S, int tri, V3 d, float u, float v, V3& n, V3& du,
                                          bool& front) {
    const float* sh = S.tri_shade + (size_t)tri * 12;
    V3 n0{sh[0], sh[1], sh[2]}, n1{sh[3], sh[4], sh[5]}, n2{sh[6], sh[7], sh[8]};
    float b0 = 1 - u - v;
    n = normalize(b0 * n0 + u * n1 + v * n2);
    front = dot(d, n) < 0;
    n = front ? n : -n;
    du = V3{sh[9], sh[10], sh[11]};
    if (is_zero(du)) {
        V3 t = cross(V3{1, 0, 0}, n);
        if (!is_zero(t)) {
            du = normalize(t);
        } else {
            t = cross(V3{0, 1, 0}, n);
            if (!is_zero(t)) {
                du = normalize(t);
            } else {
                t = cross(V3{0, 0, 1}, n);
                if (!is_zero(t)) du = normalize(t);
            }
        }
    }
}

// Sphere::Intersect hit construction, Sphere.cpp:51-63 (uv is not needed)
__device__ __forceinline__ void sphere_frame(const float* sp, V3 o, V3 d, float t, V3& p, V3& n, V3& du,
                                             bool& front) {
    V3 c{sp[0], sp[1], sp[2]};
    V3 P = mat3_mul(sp + 13, o - c);
    V3 dl = mat3_mul(sp + 13, d);
    V3 hit = P + dl * t;
    V3 N = normalize(hit);
    front = dot(dl, N) < 0;
    N = front ? N : -N;
    p = mat3_mul(sp + 4, hit) + c;
    du = normalize(mat3_mul(sp + 4, V3{-N.z, 0, N.x}));
    n = mat3_mul(sp + 4, N);
}

// Disk::Intersect hit construction, Disk.cpp:54-64 (uv is not needed)
__device__ __forceinline__ void disk_frame(const float* sp, V3 o, V3 d, float t, V3& p, V3& n, V3& du,
                                           bool& front) {
    V3 c{sp[0], sp[1], sp[2]};
    V3 P = mat3_mul(sp + 13, o - c);
    V3 dl = mat3_mul(sp + 13, d);
    V3 hit = P + dl * t;
    V3 N = mat3_mul(sp + 4, V3{0, 1, 0});
    front = dl.y <= 0;
    n = front ? N : -N;
    p = mat3_mul(sp + 4, hit) + c;
    du = normalize(c - p);
}
// Hit frame of an analytic shape (spheres and disks share the record)
__device__ __forceinline__ void analytic_frame(int kind, const float* sp, V3 o, V3 d, float t, V3& p, V3& n,
                                               V3& du, bool& front) {
    if (kind == C2_SHAPE_DISK) disk_frame(sp, o, d, t, p, n, du, front);
    else sphere_frame(sp, o, d, t, p, n, du, front);
}

__device__ __forceinline__ Inter make_inter(const DevScene& S, V3 o, V3 d, const Hit& h) {
    Inter it;
    int4 e = S.entities[h.entity];
    it.entity = h.entity;
    it.material = e.z;
    it.light = e.w;
    it.dir = d;
    if (e.x == C2_SHAPE_TRIANGLE) {
        tri_frame(S, e.y, d, h.u, h.v, it.n, it.dpdu, it.front);
        it.p = o + d * h.t;
    } else {
        analytic_frame(e.x, S.spheres + (size_t)e.y * 22, o, d, h.t, it.p, it.n, it.dpdu, it.front);
    }
    it.dpdv = cross(it.n, it.dpdu);
    return it;
}

// ------------------------------------------------------------------------- textures
// glm::clamp(x, 0, .99999f) = min(max(x, 0), .99999f) with glm's (x < y) ? y : x and
// (y < x) ? y : x. A NaN coordinate passes the clamp and makes the reference's (int) cast
// undefined; here it reads column / row 0.
C2_HD float uv_clamp(float x) {
    x = x < 0.f ? 0.f : x;
    x = .99999f < x ? .99999f : x;
    return x == x ? x : 0.f;
}

// ImageTexels::GetTexel, Texel.cpp:15-19: nearest texel, rows flipped (v = 0 is the
// bottom row). Returns the flat index into texels.
__device__ __forceinline__ int32_t texel_index(const DevScene& S, int id, float u, float v) {
    const int4 t = S.tex[id - 1];
    u = uv_clamp(u);
    v = uv_clamp(v);
    int x = (int)(u * (float)t.x);
    int y = t.y - (int)(v * (float)t.y) - 1;
    return t.z + y * t.x + x;
}
// texel value, linearised as (b / 255)^2.2 (Texel.cpp:24-38)
__device__ __forceinline__ V3 texel_value(const DevScene& S, int32_t i) {
    const uchar4 c = S.texels[i];
    return V3{S.lin[c.x], S.lin[c.y], S.lin[c.z]};
}
__device__ __forceinline__ V3 tex_lookup(const DevScene& S, int id, float u, float v) {
    return texel_value(S, texel_index(S, id, u, v));
}

// ImageCubemapTexture::Evaluate / XYZ2CubeUV, CubemapTexture.cpp:24-60
__device__ __forceinline__ V3 sky_lookup(const DevScene& S, const int32_t* sky, V3 p) {
    float ax = fabsf(p.x), ay = fabsf(p.y), az = fabsf(p.z);
    int id;
    float u, v;
    if (ax >= ay && ax >= az) {
        id = p.x >= 0 ? 0 : 1;
        u = -p.z / p.x;
        v = p.y / ax;
    } else if (ay >= az) {
        id = p.y >= 0 ? 2 : 3;
        u = p.x / ay;
        v = -p.z / p.y;
    } else {
        id = p.z >= 0 ? 4 : 5;
        u = p.x / p.z;
        v = p.y / az;
    }
    u = u * 0.5f + 0.5f;
    v = v * 0.5f + 0.5f;
    return tex_lookup(S, sky[id], u, v);
}

// Surface uv of a hit: barycentric vertex uv (Triangle.cpp:41-43), Sphere.cpp:56-58
// (theta / pi, acos(y / r) / pi + 1 in the local frame), Disk.cpp:61-62.
__device__ __forceinline__ void surface_uv(const DevScene& S, int4 e, V3 o, V3 d, const Hit& h, float& u,
                                           float& v) {
    if (e.x == C2_SHAPE_TRIANGLE) {
        const float* t = S.tri_uv + (size_t)e.y * 6;
        float b0 = 1 - h.u - h.v;
        u = (b0 * t[0] + h.u * t[2]) + h.v * t[4];
        v = (b0 * t[1] + h.u * t[3]) + h.v * t[5];
        return;
    }
    const float* sp = S.spheres + (size_t)e.y * 22;
    V3 c{sp[0], sp[1], sp[2]};
    V3 P = mat3_mul(sp + 13, o - c);
    V3 dl = mat3_mul(sp + 13, d);
    V3 hit = P + dl * h.t;
    if (e.x == C2_SHAPE_DISK) {
        u = length(hit) / sp[3];
        v = (atan2_r(-hit.z, hit.x) + kPi) / kTwoPi;
    } else {
        u = atan2_r(-hit.z, hit.x) / kPi;
        v = acos_r(hit.y / sp[3]) / kPi + 1.0f;
    }
}

// Full single-shape intersection used by Shape::PdfLight (Shape.cpp:4-12):
// returns hit position and the facing shading normal.
__device__ __forceinline__ bool shape_hit(const DevScene& S, int kind, int shape, V3 o, V3 d, V3& p, V3& n) {
    if (kind == C2_SHAPE_TRIANGLE) {
        const float* tp = S.tri_pos + (size_t)shape * 9;
        V3 v0{tp[0], tp[1], tp[2]}, v1{tp[3], tp[4], tp[5]}, v2{tp[6], tp[7], tp[8]};
        V3 res = tri_solve(o, d, v0, v1, v2);
        if (!tri_accept_closest(res)) return false;
        V3 du;
        bool front;
        tri_frame(S, shape, d, res.x, res.y, n, du, front);
        p = o + d * res.z;
        return true;
    }
    const float* sp = S.spheres + (size_t)shape * 22;
    float t;
    if (!(kind == C2_SHAPE_DISK ? disk_closest(sp, o, d, t) : sphere_closest(sp, o, d, t))) return false;
    V3 du;
    bool front;
    analytic_frame(kind, sp, o, d, t, p, n, du, front);
    return true;
}

// ------------------------------------------------------------------ lights
// Shape::PdfLight, Shape.cpp:4-12 (SurfaceInfo::SpawnRay: origin p + n * 1e-4)
__device__ __forceinline__ float pdf_light(const DevScene& S, const DevLight& L, V3 sp, V3 sn, V3 wi) {
    V3 o = sp + sn * 1e-4f;
    V3 hp, hn;
    if (!shape_hit(S, L.shape_kind, L.shape, o, wi, hp, hn)) return 0.f;
    V3 dd = sp - hp;
    return dot(dd, dd) / (fmax_ref(0.f, dot(-normalize(wi), hn)) * L.area);
}

// Light::Sample_Li: DiffusedAreaLight.cpp:25-34, PointLight.cpp:14-20
__device__ __forceinline__ V3 sample_li(const DevScene& S, const DevLight& L, V3 sp, V3 sn, float u0, float u1,
                                        V3& endpoint, float& pdf) {
    if (L.type == C2_LIGHT_POINT) {
        endpoint = V3{L.pos[0], L.pos[1], L.pos[2]};
        pdf = 1.0f;
        V3 I = V3{L.L[0], L.L[1], L.L[2]} / (4 * kPi);
        V3 dd = endpoint - sp;
        return I / dot(dd, dd);
    }
    V3 lp, ln;
    if (L.shape_kind == C2_SHAPE_TRIANGLE) {  // Triangle::SampleSurface, Triangle.cpp:100-125
        const float* tp = S.tri_pos + (size_t)L.shape * 9;
        const float* sh = S.tri_shade + (size_t)L.shape * 12;
        float su = 1 - sqrtf(u0);
        float sv = u1 * sqrtf(u0);
        float b0 = 1.f - su - sv;
        lp = b0 * V3{tp[0], tp[1], tp[2]} + su * V3{tp[3], tp[4], tp[5]} + sv * V3{tp[6], tp[7], tp[8]};
        ln = normalize(b0 * V3{sh[0], sh[1], sh[2]} + su * V3{sh[3], sh[4], sh[5]} + sv * V3{sh[6], sh[7], sh[8]});
    } else {  // Sphere::SampleSurface + NextUnitSphere, Sphere.cpp:90-95, Utils.h:138-149
        const float* s3 = S.spheres + (size_t)L.shape * 22;
        if (L.shape_kind == C2_SHAPE_DISK) {  // Disk::SampleSurface, Disk.cpp:88-96
            float r = sqrtf(u0);
            float phi = u1 * kTwoPi;
            float R = s3[3];
            V3 dw = mat3_mul(s3 + 4, V3{cos_r(phi) * r * R, 0.f, -sin_r(phi) * r * R});
            lp = V3{s3[0], s3[1], s3[2]} + dw;
            ln = mat3_mul(s3 + 4, V3{0, 1, 0});
        } else {
        float y = 2.f * u0 - 1.f;
        float phi = u1 * kTwoPi;
        float r = sqrtf(1 - y * y);
        V3 dir{r * cos_r(phi), y, r * sin_r(phi)};
        lp = V3{s3[0], s3[1], s3[2]} + dir * s3[3];
        ln = dir;
        }
    }
    endpoint = lp;
    V3 dir = endpoint - sp;
    pdf = pdf_light(S, L, sp, sn, dir);
    // EvalEmitRadiance(surface_light, -dir), DiffusedAreaLight.cpp:11-21
    return dot(ln, -dir) < 0 ? V3{0, 0, 0} : V3{L.L[0], L.L[1], L.L[2]};
}

__device__ __forceinline__ float power_heuristic(float fPdf, float gPdf) {  // Utils.h:168-172
    float f = 1 * fPdf, g = 1 * gPdf;
    return (f * f) / (f * f + g * g);
}

// ------------------------------------------------------------------ BxDFs
enum : int { T_REFL = 1, T_TRANS = 2, T_GLOSSY = 4, T_SPEC = 8, T_DIFF = 16, T_ALL = 0x1F };
enum : int { BK_NONE, BK_LAMBERT, BK_SPECREFL, BK_SPECFRESNEL, BK_MFREFL, BK_MFBLEND, BK_PHONG, BK_ORENNAYAR, BK_FBLEND };
enum : int { FR_NOOP, FR_DIELECTRIC, FR_SCHLICK };

struct Bsdf {
    int kind, type, fresnel, k;
    int ndf;        // C2_NDF_* of the microfacet lobes
    V3 R;
    V3 S;           // FresnelBlend: specular reflectance Ks
    float etaA, etaB, ax, ay;
    float onA, onB; // OrenNayarReflection's A and B
    uint32_t ridx, rdim;  // Blinn-Phong redraw stream
};

// Fresnel::Eval (scalar: all components are equal), Fresnel.cpp:5-24
__device__ __forceinline__ float fresnel(const Bsdf& b, float etaA, float etaB, float cosI) {
    if (b.fresnel == FR_NOOP) return 1.f;
    float sinI = sqrtf(fmax_ref(0.f, 1.f - cosI * cosI));
    float sinT = etaA / etaB * sinI;
    if (sinT >= 1.f) return 1.f;
    if (b.fresnel == FR_SCHLICK) return 0.99f + (1.f - 0.99f) * pow5_r(1.f - cosI);  // std::pow(x, 5.f)
    float cosT = sqrtf(fmax_ref(0.f, 1.f - sinT * sinT));
    float Rparl = ((etaB * cosI) - (etaA * cosT)) / ((etaB * cosI) + (etaA * cosT));
    float Rperp = ((etaA * cosI) - (etaB * cosT)) / ((etaA * cosI) + (etaB * cosT));
    return (Rparl * Rparl + Rperp * Rperp) / 2;
}

// GGXDistribution, MicrofacetDistribution.cpp:6-82
__device__ __forceinline__ void cos_sin_phi(V3 w, float sinTheta, float& c, float& s) {
    if (sinTheta == 0.f) { c = 1.f; s = 0.f; return; }
    c = clampf_ref(w.x / sinTheta, -1.f, 1.f);
    s = clampf_ref(-w.z / sinTheta, -1.f, 1.f);
}
__device__ __forceinline__ float sq(float x) { return x * x; }
__device__ __forceinline__ float ggx_D(const Bsdf& b, V3 wh) {
    float cos2 = wh.y * wh.y;
    float sin2 = fmax_ref(0.f, 1.f - cos2);
    float pc, ps;
    cos_sin_phi(wh, sqrtf(sin2), pc, ps);
    float A = sq(pc) / sq(b.ax) + sq(ps) / sq(b.ay);
    return 1.f / (kPi * b.ax * b.ay * sq(A - cos2 * (A - 1)));
}
__device__ __forceinline__ float ggx_lambda(const Bsdf& b, V3 w) {
    float cos2 = w.y * w.y;
    float sin2 = 1.f - cos2;
    float tan2 = sin2 / cos2;
    if (is_inf(tan2)) return 0.f;
    float pc, ps;
    cos_sin_phi(w, sqrtf(sin2), pc, ps);
    float alpha2 = sq(pc) * sq(b.ax) + sq(ps) * sq(b.ay);
    return (-1 + sqrtf(1 + alpha2 * tan2)) / 2;
}
__device__ __forceinline__ float ggx_G(const Bsdf& b, V3 wo, V3 wi) {
    return 1.f / (1.f + ggx_lambda(b, wo) + ggx_lambda(b, wi));
}
__device__ __forceinline__ V3 unit_from_cos(float cosTheta, float phi) {  // Utils.h:79-83
    float r = sqrtf(1 - cosTheta * cosTheta);
    return V3{r * cos_r(phi), cosTheta, -r * sin_r(phi)};
}
__device__ __forceinline__ V3 ggx_sample_wh(const Bsdf& b, float sx, float sy) {
    float cosTheta, phi;
    if (b.ax == b.ay) {
        cosTheta = sqrtf((1 - sx) / (sx * (b.ax * b.ax - 1) + 1));
        phi = kTwoPi * sy;
    } else {
        phi = atan_r(b.ay / b.ax * tan_r(kTwoPi * sy + kHalfPi));
        if (sy > 0.5) phi += kPi;
        float sinPhi = sin_r(phi), cosPhi = cos_r(phi);
        float A = sq(cosPhi) / sq(b.ax) + sq(sinPhi) / sq(b.ay);
        float cos2 = A * (sx - 1) / (sx * (A - 1) - A);
        cosTheta = sqrtf(cos2);
    }
    return unit_from_cos(cosTheta, phi);
}
__device__ __forceinline__ float ggx_pdf(const Bsdf& b, V3 wh) { return ggx_D(b, wh) * fabsf(wh.y); }

// BeckmannDistribution, MicrofacetDistribution.cpp:84-161 (alpha = roughness as given:
// the constructor does not call RoughnessToAlpha)
__device__ __forceinline__ float beck_D(const Bsdf& b, V3 wh) {
    float cos2 = wh.y * wh.y;
    float sin2 = 1.f - cos2;
    float tan2 = sin2 / cos2;
    if (is_inf(tan2)) return 0.f;
    float cos4 = cos2 * cos2;
    float pc, ps;
    cos_sin_phi(wh, sqrtf(sin2), pc, ps);
    float bot = kPi * b.ax * b.ay * cos4;
    float top = exp_r(-tan2 * (sq(pc) / sq(b.ax) + sq(ps) / sq(b.ay)));
    return top / bot;
}
__device__ __forceinline__ float beck_lambda(const Bsdf& b, V3 w) {
    float cosTheta = w.y;
    float sinTheta = sqrtf(1.f - cosTheta * cosTheta);
    float tanTheta = fabsf(sinTheta / cosTheta);
    if (is_inf(tanTheta)) return 0.f;
    float pc, ps;
    cos_sin_phi(w, sinTheta, pc, ps);
    float alpha = sqrtf(sq(pc) * sq(b.ax) + sq(ps) * sq(b.ay));
    float a = 1.f / (tanTheta * alpha);
    if (a >= 1.6f) return 0.f;
    return (1 - 1.259f * a + 0.396f * a * a) / (3.535f * a + 2.181f * a * a);
}
__device__ __forceinline__ V3 beck_sample_wh(const Bsdf& b, float sx, float sy) {
    float logSample = log_r(1.f - sx);
    if (is_inf(logSample)) logSample = 0;
    float tan2, phi;
    if (b.ax == b.ay) {
        tan2 = -sq(b.ax) * logSample;
        phi = sy * kTwoPi;
    } else {
        phi = atan_r(b.ay / b.ax * tan_r(kTwoPi * sy + kHalfPi));
        if (sy > 0.5) phi += kPi;
        float sinPhi = sin_r(phi), cosPhi = cos_r(phi);
        tan2 = -logSample / (cosPhi * cosPhi / sq(b.ax) + sinPhi * sinPhi / sq(b.ay));
    }
    float cosTheta = 1.f / sqrtf(1.f + tan2);
    return unit_from_cos(cosTheta, phi);
}

// MicrofacetDistribution dispatch (GGX or Beckmann): D, G = 1 / (1 + lambda(wo) +
// lambda(wi)), Sample_wh and Pdf = D * |wh.y| have the same form for both.
__device__ __forceinline__ float mf_D(const Bsdf& b, V3 wh) {
    return b.ndf == C2_NDF_BECKMANN ? beck_D(b, wh) : ggx_D(b, wh);
}
__device__ __forceinline__ float mf_G(const Bsdf& b, V3 wo, V3 wi) {
    if (b.ndf == C2_NDF_BECKMANN) return 1.f / (1.f + beck_lambda(b, wo) + beck_lambda(b, wi));
    return ggx_G(b, wo, wi);
}
__device__ __forceinline__ V3 mf_sample_wh(const Bsdf& b, float sx, float sy) {
    return b.ndf == C2_NDF_BECKMANN ? beck_sample_wh(b, sx, sy) : ggx_sample_wh(b, sx, sy);
}
__device__ __forceinline__ float mf_pdf(const Bsdf& b, V3 wh) { return mf_D(b, wh) * fabsf(wh.y); }

// refract(wo, wh, eta, wt), Utils.h:102-112
__device__ __forceinline__ bool refract_h(V3 wo, V3 wh, float eta, V3& wt) {
    float cosO = clampf_ref(dot(wo, wh), -1.f, 1.f);
    float sin2O = fmax_ref(0.f, 1.f - cosO * cosO);
    float sin2I = eta * eta * sin2O;
    if (sin2I > 1.f) return false;
    float cosI = sqrtf(1.f - sin2I);
    wt = eta * -wo + (eta * cosO - cosI) * wh;
    return true;
}

// MicrofacetReflection, Microfacet.cpp:17-55
__device__ __forceinline__ float mfr_pdf(const Bsdf& b, V3 wo, V3 wi) {
    V3 H = normalize(wo + wi);
    return mf_pdf(b, H) / (4 * dot(wo, H));
}
__device__ __forceinline__ V3 mfr_f(const Bsdf& b, V3 wo, V3 wi) {
    if (wo.y * wi.y < 0) return V3{0, 0, 0};
    V3 H = normalize(wo + wi);
    float F = fresnel(b, b.etaA, b.etaB, fmax_ref(0.f, dot(wo, H)));
    float G = mf_G(b, wo, wi);
    float D = mf_D(b, H);
    return b.R * D * G * F / (4 * fmax_ref(0.f, wi.y) * fmax_ref(0.f, wo.y));
}
// MicrofacetTransmission, Microfacet.cpp:57-125
__device__ __forceinline__ float mft_pdf(const Bsdf& b, V3 wo, V3 wi) {
    if (wo.y * wi.y > 0) return 0.f;
    float eta = b.etaB / b.etaA;
    V3 H = normalize(wo + wi * eta);
    float absDot = fabsf(dot(H, wi));
    float dwh_dwi = eta * eta * absDot / sq(dot(wo, H) + eta * dot(wi, H));
    return mf_pdf(b, H) * dwh_dwi;
}
__device__ __forceinline__ V3 mft_f(const Bsdf& b, V3 wo, V3 wi) {
    if (wi.y * wo.y > 0) return V3{0, 0, 0};
    if (wo.y == 0.f) return V3{0, 0, 0};
    float eta = b.etaB / b.etaA;
    V3 H = normalize(wo + wi * eta);
    if (H.y < 0) H = -H;
    float bot = sq(dot(wo, H) + eta * dot(wi, H));
    float F = 1.f - fresnel(b, b.etaA, b.etaB, fmax_ref(0.f, dot(wo, H)));
    float G = mf_G(b, wo, wi);
    float D = mf_D(b, H);
    return b.R * D * G * F * fabsf(dot(H, wo)) * fabsf(dot(H, wi)) /
           (bot * fmax_ref(0.f, -wi.y) * fmax_ref(0.f, wo.y));
}
// BlinnPhongReflection, BlinnPhongReflection.cpp:14-26
__device__ __forceinline__ float phong_pdf(const Bsdf& b, V3 wo, V3 wi) {
    V3 wh = normalize(wo + wi);
    return (float)((double)((b.k + 2) / kTwoPi) * powd((double)fmax_ref(0.f, wh.y), (double)(b.k + 1)));
}
__device__ __forceinline__ V3 phong_f(const Bsdf& b, V3 wo, V3 wi) {
    V3 wh = normalize(wo + wi);
    float bb = fmax_ref(0.f, wh.y);
    return b.R * ((float)powd((double)bb, (double)b.k) * (b.k + 2) / kTwoPi) /
           (4 * fmax_ref(0.f, wi.y) * fmax_ref(0.f, wo.y));
}

// OrenNayarReflection, Reflections.cpp:28-61 (A and B from sigma, :18-26)
__device__ __forceinline__ V3 orennayar_f(const Bsdf& b, V3 wo, V3 wi) {
    float sinThetaO = sqrtf(fmax_ref(0.f, 1.f - wo.y * wo.y));
    float sinThetaI = sqrtf(fmax_ref(0.f, 1.f - wi.y * wi.y));
    float sinAlpha, tanBeta;
    if (wo.y < wi.y) {
        sinAlpha = sinThetaI;
        tanBeta = sinThetaO / wo.y;
    } else {
        sinAlpha = sinThetaO;
        tanBeta = sinThetaI / wi.y;
    }
    float dCos = 0.f;
    if (sinThetaI > 1e-5f && sinThetaO > 1e-5f) {  // EPS, src/Core/Utils.h:7
        float ci, si, co, so;
        cos_sin_phi(wi, sinThetaI, ci, si);
        cos_sin_phi(wo, sinThetaO, co, so);
        dCos = fmax_ref(0.f, ci * co + si * so);
    }
    return b.R / kPi * (b.onA + b.onB * (dCos * sinAlpha * tanBeta));
}
// FresnelBlend, FresnelBlend.cpp:15-68 (its own Schlick with R0 = Ks)
__device__ __forceinline__ float fblend_pdf(const Bsdf& b, V3 wo, V3 wi) {
    V3 H = normalize(wo + wi);
    return 0.5f * (fmax_ref(0.f, wi.y) / kPi + mf_pdf(b, H) / (4 * dot(wo, H)));
}
__device__ __forceinline__ V3 fblend_f(const Bsdf& b, V3 wo, V3 wi) {
    float cosThetaI = fmax_ref(0.f, wi.y);
    float cosThetaO = fmax_ref(0.f, wo.y);
    // (float)std::pow(x, 5) evaluates in double: pow5_r
    V3 diffuse = (28.f / (23.f * kPi)) * b.R * (V3{1.f, 1.f, 1.f} - b.S) * (1.f - pow5_r(1.f - 0.5f * cosThetaO)) *
                 (1.f - pow5_r(1.f - 0.5f * cosThetaI));
    V3 H = normalize(wo + wi);
    float c = dot(wi, H);
    V3 schlick = b.S + (V3{1.f, 1.f, 1.f} - b.S) * pow5_r(1.f - c);  // std::pow(1 - cos, 5.f)
    V3 specular = mf_D(b, H) / (4 * dot(wi, H) * fmax_ref(cosThetaI, cosThetaO)) * schlick;
    return diffuse + specular;
}

// BxDF::DistributionFunction / Pdf in local coordinates
__device__ __forceinline__ V3 bxdf_f(const Bsdf& b, V3 wo, V3 wi) {
    switch (b.kind) {
        case BK_LAMBERT: return b.R / kPi;
        case BK_MFREFL: return mfr_f(b, wo, wi);
        case BK_MFBLEND: return (wo.y * wi.y < 0) ? mft_f(b, wo, wi) : mfr_f(b, wo, wi);
        case BK_PHONG: return phong_f(b, wo, wi);
        case BK_ORENNAYAR: return orennayar_f(b, wo, wi);
        case BK_FBLEND: return fblend_f(b, wo, wi);
        default: return V3{0, 0, 0};
    }
}
__device__ __forceinline__ float bxdf_pdf(const Bsdf& b, V3 wo, V3 wi) {
    switch (b.kind) {
        case BK_LAMBERT: return fmax_ref(0.f, wi.y) / kPi;
        case BK_MFREFL: return mfr_pdf(b, wo, wi);
        case BK_MFBLEND: return (wo.y * wi.y < 0) ? mft_pdf(b, wo, wi) : mfr_pdf(b, wo, wi);
        case BK_PHONG: return phong_pdf(b, wo, wi);
        case BK_ORENNAYAR: return fmax_ref(0.f, wi.y) / kPi;
        case BK_FBLEND: return fblend_pdf(b, wo, wi);
        default: return 0.f;
    }
}
// BxDF::SampleDirection in local coordinates; `sampled` may be overwritten.
__device__ __forceinline__ V3 bxdf_sample(const Bsdf& b, float sx, float sy, V3 wo, V3& wi, float& pdf,
                                          int& sampled) {
    switch (b.kind) {
        case BK_LAMBERT: {  // Lambertain.cpp:22-26, NextCosineUnitHemiSphere Utils.h:114-123
            float r = sqrtf(sx);
            float phi = sy * kTwoPi;
            float x = r * cos_r(phi), z = r * sin_r(phi);
            float y = sqrtf(1.0f - r * r);
            pdf = y / kPi;
            wi = V3{x, y, z};
            return b.R / kPi;
        }
        case BK_SPECREFL:  // SpecularReflection.cpp:22-27
            wi = V3{-wo.x, wo.y, -wo.z};
            pdf = 1.0f;
            return b.R * 1.f;
        case BK_SPECFRESNEL: {  // SpecularFresnel.cpp:28-46
            float fr = fresnel(b, b.etaA, b.etaB, wo.y);
            if (sx < fr) {
                sampled = T_REFL | T_SPEC;
                wi = V3{-wo.x, wo.y, -wo.z};
                pdf = fr;
                return fr * b.R;
            }
            sampled = T_TRANS | T_SPEC;
            if (!refract_h(wo, V3{0, 1, 0}, b.etaA / b.etaB, wi)) wi = V3{0, 0, 0};
            pdf = 1.f - fr;
            return (1.f - fr) * b.R;
        }
        case BK_MFREFL: {  // Microfacet.cpp:41-53
            V3 wh = mf_sample_wh(b, sx, sy);
            if (dot(wo, wh) <= 0) return V3{0, 0, 0};
            wi = reflect(-wo, wh);
            pdf = mfr_pdf(b, wo, wi);
            wi = clamp_unit(wi);
            return mfr_f(b, wo, wi);
        }
        case BK_MFBLEND: {  // Microfacet.cpp:162-198
            V3 wh = mf_sample_wh(b, sx, sy);
            if (dot(wo, wh) <= 0) return V3{0, 0, 0};
            float fr = fresnel(b, b.etaA, b.etaB, dot(wo, wh));
            if (sx < fr) {
                sampled = T_REFL | T_GLOSSY;
                wi = reflect(-wo, wh);
                pdf = bxdf_pdf(b, wo, wi) * fr;
                wi = clamp_unit(wi);
                return mfr_f(b, wo, wi);
            }
            float eta = b.etaA / b.etaB;
            if (refract_h(wo, wh, eta, wi)) {
                sampled = T_TRANS | T_GLOSSY;
                pdf = bxdf_pdf(b, wo, wi) * (1 - fr);
                wi = clamp_unit(wi);
                return mft_f(b, wo, wi);
            }
            sampled = T_REFL | T_GLOSSY;
            wi = reflect(-wo, wh);
            pdf = bxdf_pdf(b, wo, wi) * (1 - fr);
            wi = clamp_unit(wi);
            return mfr_f(b, wo, wi);
        }
        case BK_ORENNAYAR: {  // Reflections.cpp:58-61: cosine-weighted hemisphere
            float r = sqrtf(sx);
            float phi = sy * kTwoPi;
            float x = r * cos_r(phi), z = r * sin_r(phi);
            float y = sqrtf(1.0f - r * r);
            pdf = y / kPi;
            wi = V3{x, y, z};
            return orennayar_f(b, wo, wi);
        }
        case BK_FBLEND: {  // FresnelBlend.cpp:45-68
            float u = sx;
            if (u < 0.5) {
                float s0 = u * 2.f;
                float r = sqrtf(s0);
                float phi = sy * kTwoPi;
                float x = r * cos_r(phi), z = r * sin_r(phi);
                float y = sqrtf(1.0f - r * r);
                pdf = y / kPi;  // NextCosineUnitHemiSphere writes it; replaced below
                wi = V3{x, y, z};
            } else {
                float s0 = (u - 0.5f) * 2.f;
                V3 wh = mf_sample_wh(b, s0, sy);
                V3 L = reflect(-wo, wh);
                if (L.y * wo.y <= 0) return V3{0, 0, 0};  // before writing wi and pdf
                wi = clamp_unit(L);
            }
            pdf = fblend_pdf(b, wo, wi);
            return fblend_f(b, wo, wi);
        }
        case BK_PHONG: {  // BlinnPhongReflection.cpp:32-58, deterministic redraw stream
            for (uint32_t redraw = 0; redraw < 1024; redraw++) {
                float cosTheta = (float)powd((double)sx, 1.0 / (b.k + 2));
                float phi = kTwoPi * sy;
                V3 N = unit_from_cos(cosTheta, phi);
                bool ok = dot(N, wo) >= 0;
                if (ok) {
                    wi = reflect(-wo, N);
                    ok = wi.y >= 0;
                }
                if (!ok) {
                    uint32_t stream = 0x40000000u | (b.rdim << 12) | (redraw << 1);
                    sx = hash_float(b.ridx, stream);
                    sy = hash_float(b.ridx, stream + 1);
                    continue;
                }
                pdf = phong_pdf(b, wo, wi);
                return phong_f(b, wo, wi);
            }
            pdf = 0.f;
            return V3{0, 0, 0};
        }
        default: return V3{0, 0, 0};
    }
}

// Material::ComputeScatteringFunctions -> one BxDF (all supported materials
// build exactly one), src/Materials/*.cpp
// texel: the Kd texel of the hit (flat index) for a textured Diffuse or Phong material
// MAXT: the largest C2_MAT_* type the scene holds (kMatAll: any). A kernel instantiated with
// C2_MAT_MIRROR keeps only the Lambertian and specular-reflection code after inlining, and
// with it far fewer registers; the caller guarantees that no other type reaches it.
constexpr int kMatAll = C2_MAT_ORENNAYAR;
template <int MAXT = kMatAll>
__device__ __forceinline__ Bsdf make_bsdf(const DevScene& S, const Inter& it, uint32_t idx, uint32_t rdim,
                                          int32_t texel) {
    Bsdf b;
    b.kind = BK_NONE;
    b.type = 0;
    b.fresnel = FR_NOOP;
    b.k = 0;
    b.ndf = C2_NDF_GGX;
    b.S = V3{0.f, 0.f, 0.f};
    b.etaA = b.etaB = 1.f;
    b.ax = b.ay = 0.f;
    b.onA = b.onB = 0.f;
    b.ridx = idx;
    b.rdim = rdim;
    if (it.material < 0) return b;
    const c2_material m = S.materials[it.material];
    if constexpr (MAXT < kMatAll) __builtin_assume(m.type >= 0 && m.type <= MAXT);
    b.R = V3{m.color[0], m.color[1], m.color[2]};
    // Diffuse / Phong with a "Kd" texture: R = texture(uv), DiffuseMaterial.cpp:16-20
    if (m.kd_texture > 0) b.R = texel_value(S, texel);
    switch (m.type) {
        case C2_MAT_DIFFUSE:
            b.kind = BK_LAMBERT;
            b.type = T_DIFF | T_REFL;
            break;
        case C2_MAT_MIRROR:
            b.kind = BK_SPECREFL;
            b.type = T_SPEC | T_REFL;
            break;
        case C2_MAT_GLASS: {
            float etaA = 1.0f, etaB = m.ior;
            if (!it.front) { float t = etaA; etaA = etaB; etaB = t; }
            b.etaA = etaA;
            b.etaB = etaB;
            b.fresnel = FR_DIELECTRIC;
            if (m.roughness[0] == 0.f && m.roughness[1] == 0.f) {
                b.kind = BK_SPECFRESNEL;
                b.type = T_TRANS | T_REFL | T_SPEC;
            } else {
                b.kind = BK_MFBLEND;
                b.type = T_TRANS | T_REFL | T_GLOSSY;
                b.ax = m.roughness[0];
                b.ay = m.roughness[1];
            }
            break;
        }
        case C2_MAT_MICRO: {
            float etaA = 1.f, etaB = m.ior;
            if (!it.front) { float t = etaA; etaA = etaB; etaB = t; }
            b.etaA = etaA;
            b.etaB = etaB;
            b.kind = BK_MFREFL;
            b.type = T_REFL | T_GLOSSY;
            b.fresnel = FR_SCHLICK;
            b.ax = m.roughness[0];
            b.ay = m.roughness[1];
            b.ndf = m.ndf;
            break;
        }
        case C2_MAT_PLASTIC:  // Plastic.cpp:13-27: one FresnelBlend(Kd, Ks, NDF)
            b.kind = BK_FBLEND;
            b.type = T_REFL | T_GLOSSY;
            b.S = V3{m.ks[0], m.ks[1], m.ks[2]};
            b.ax = m.roughness[0];
            b.ay = m.roughness[1];
            b.ndf = m.ndf;
            break;
        case C2_MAT_ORENNAYAR: {  // DiffuseMaterial.cpp:29-34, Reflections.cpp:18-26
            b.kind = BK_ORENNAYAR;
            b.type = T_DIFF | T_REFL;
            float sigmaSq = m.sigma * m.sigma;
            b.onA = 1.f - sigmaSq / (2 * (sigmaSq + 0.33f));
            b.onB = (0.45f * sigmaSq) / (sigmaSq + 0.09f);
            break;
        }
        case C2_MAT_PHONG:
            b.kind = BK_PHONG;
            b.type = T_GLOSSY | T_REFL;
            b.k = m.phong_k;
            break;
        default: break;
    }
    return b;
}

__device__ __forceinline__ bool matches(int type, int flags) { return (type & flags) == type; }

// BSDF::DistributionFunction with one BxDF, BSDF.cpp:47-62
__device__ __forceinline__ V3 bsdf_f(const Bsdf& b, const Inter& it, V3 woW, V3 wiW) {
    V3 wo = clamp_unit(to_local(it, woW)), wi = clamp_unit(to_local(it, wiW));
    bool reflect = wi.y * wo.y > 0;
    if ((reflect && (b.type & T_REFL)) || (!reflect && (b.type & T_TRANS))) return V3{0, 0, 0} + bxdf_f(b, wo, wi);
    return V3{0, 0, 0};
}
// BSDF::Pdf with one BxDF (operator precedence as at BSDF.cpp:39-40)
__device__ __forceinline__ float bsdf_pdf(const Bsdf& b, const Inter& it, V3 woW, V3 wiW, int flags) {
    V3 wo = clamp_unit(to_local(it, woW)), wi = clamp_unit(to_local(it, wiW));
    bool reflect = wi.y * wo.y > 0;
    if ((matches(b.type, flags) && (reflect && (b.type & T_REFL))) || (!reflect && (b.type & T_TRANS)))
        return (0.f + bxdf_pdf(b, wo, wi)) / 1;
    return 0.f;
}
// BSDF::SampleDirection with one BxDF, BSDF.cpp:64-130. Returns f; sets wi (world),
// pdf and sampled type; f = 0 when the BxDF does not match or sampling fails.
// status (optional): kSampleOk; kSampleNoMatch (no BxDF matches `flags`: the reference
// returns before writing wi, pdf and the sampled type); kSampleNoPdf (the BxDF returned
// before writing wi and pdf); kSampleZero (the BxDF wrote wi, in the local frame, and
// pdf, but f or pdf is 0: *wiFail gets that local wi).
enum : int { kSampleOk = 0, kSampleNoMatch = 1, kSampleNoPdf = 2, kSampleZero = 3 };
__device__ __forceinline__ V3 bsdf_sample(const Bsdf& b, const Inter& it, V3 woW, float sx, float sy, int flags,
                                          V3& wiW, float& pdf, int& sampled, int* status = nullptr,
                                          V3* wiFail = nullptr) {
    if (b.kind == BK_NONE || !matches(b.type, flags)) {
        if (status) *status = kSampleNoMatch;
        return V3{0, 0, 0};
    }
    sampled = b.type;
    V3 wo = clamp_unit(to_local(it, woW));
    V3 wi{0, 0, 0};
    pdf = -1.f;  // "not written" (a pdf is never negative)
    V3 L = bxdf_sample(b, sx, sy, wo, wi, pdf, sampled);
    if (pdf == -1.f) {
        pdf = 0.f;
        if (status) *status = kSampleNoPdf;
        return V3{0, 0, 0};
    }
    if (is_zero(L) || pdf == 0.f) {
        if (status) *status = kSampleZero;
        if (wiFail) *wiFail = wi;
        return V3{0, 0, 0};
    }
    if (status) *status = kSampleOk;
    wiW = clamp_unit(to_world(it, wi));
    pdf /= 1;
    return L;
}

}  // namespace c2
