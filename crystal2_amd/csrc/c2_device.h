// Device-side scene layout, ray/primitive intersection and 4-wide BVH traversal
// for gfx950 (wave64). Included by c2_device.hip only.
//
// Arithmetic follows the reference operation by operation (the file is compiled
// with -ffp-contract=off, divisions and square roots are IEEE correctly rounded):
//   triangle solve   src/Shapes/Triangle.cpp:32-93 with glm::inverse(mat3)
//                    (glm/detail/func_matrix.inl:269-291) and mat3*vec3
//                    (glm/detail/type_mat3x3.inl:468-474)
//   sphere           src/Shapes/Sphere.cpp:39-88
//   slab test        src/Core/Geometry.h:136-157 (std::max/std::min semantics)
//   traversal        src/Accelerators/BVH.cpp:56-154: per-entity box pre-clip with
//                    [tmin, min(tmax, best)] (closest) / [tmin, tmax] (any), closest by
//                    strict t with ties to the lowest entity index, any-hit over the
//                    box-clipped interval, the `ignore` entity skipped.
#pragma once
#include <type_traits>
#include <hip/hip_runtime.h>

#include "c2_bvh.h"
#include "c2_vec.h"
#include "../../include/crystal2.h"

namespace c2 {

struct DevLight {
    int32_t type;       // C2_LIGHT_*
    int32_t entity;     // AREA: attached entity
    int32_t shape_kind; // AREA: C2_SHAPE_*
    int32_t shape;      // AREA: triangle / sphere index
    float L[3];         // Lemit or Flux
    float pos[3];       // POINT position
    float area;         // AREA: Shape::SurfaceArea()
    int32_t eligible;   // Flux() != 0 (UniformSampleAllLights skips zero-flux lights)
};

struct DevScene {
    const Node4* nodes;
    const Node8* nodes8;       // the 8-wide tree (C2_ACCEL_BVH8), null otherwise
    const float4* prims;       // 3 float4 per slot
    const int4* entities;      // (kind, shape, material, light)
    const float* tri_pos;      // [ntri][9]
    const float* tri_shade;    // [ntri][12]: n0, n1, n2, dpdu (0 = derive from N)
    const float* spheres;      // [nsph][22]
    const c2_material* materials;
    const DevLight* lights;
    const uint16_t* perm;      // Halton permutation tables, dims 0..99
    const int32_t* perm_off;
    const uchar4* texels;      // every texture's texels (rgb + pad), rows top-down
    const int4* tex;           // texture id k: tex[k - 1] = (width, height, first texel, 0)
    const float* lin;          // [256]: (b / 255)^2.2, the texel linearisation
    const float* tri_uv;       // [ntri][6] vertex uv (allocated with textures only)
    const int4* kd_nodes;      // KD-tree accelerator (null with the BVH): nodes and, per leaf
    const float4* kd_rec;      // entry, 5 float4: (box lo, entity), (box hi, kind), then the
                               // triangle's 9 vertex floats or (sphere / disk record index)
    int32_t num_lights;
    int32_t num_entities;
};

// ------------------------------------------------------------------------- slab test
struct RayPre {
    V3 o, d, invD;
    bool neg[3];
};
C2_HD RayPre make_ray(V3 o, V3 d) {
    RayPre r;
    r.o = o;
    r.d = d;
    r.invD = V3{1.f / d.x, 1.f / d.y, 1.f / d.z};
    r.neg[0] = r.invD.x < 0;
    r.neg[1] = r.invD.y < 0;
    r.neg[2] = r.invD.z < 0;
    return r;
}
// RayBoxTest(ray, inv, invD, box, tMin, tMax), Geometry.h:136-157. The reference
// updates tMin = std::max(tMin, a) and tMax = std::min(tMax, b) axis by axis and
// returns false as soon as tMax < tMin. Here tMin and tMax are never NaN, so
// std::max(tMin, a) (= tMin < a ? a : tMin) and std::min(tMax, b) equal IEEE
// maxNum / minNum, which also ignore a NaN a or b; and later axes only shrink the
// interval, so the early returns do not change the outcome: one test after the
// three axes gives the same result and, when it passes, the same interval (up to
// the sign of a zero, which no comparison sees). Ray tmin / tmax must not be NaN.
C2_HD bool slab(const RayPre& r, float lx, float ly, float lz, float hx, float hy, float hz,
                                     float& tMin, float& tMax) {
    float ax = (lx - r.o.x) * r.invD.x, bx = (hx - r.o.x) * r.invD.x;
    float ay = (ly - r.o.y) * r.invD.y, by = (hy - r.o.y) * r.invD.y;
    float az = (lz - r.o.z) * r.invD.z, bz = (hz - r.o.z) * r.invD.z;
    if (r.neg[0]) { float t = ax; ax = bx; bx = t; }
    if (r.neg[1]) { float t = ay; ay = by; by = t; }
    if (r.neg[2]) { float t = az; az = bz; bz = t; }
    tMin = fmaxf(fmaxf(fmaxf(tMin, ax), ay), az);
    tMax = fminf(fminf(fminf(tMax, bx), by), bz);
    return !(tMax < tMin);
}

// ------------------------------------------------------------------------- triangle
// res = glm::inverse(mat3(v1 - v0, v2 - v0, -d)) * (o - v0)
C2_HD V3 tri_solve(V3 o, V3 d, V3 v0, V3 v1, V3 v2) {
    V3 e1 = v1 - v0, e2 = v2 - v0, md = -d, P = o - v0;
    float m00 = e1.x, m01 = e1.y, m02 = e1.z;
    float m10 = e2.x, m11 = e2.y, m12 = e2.z;
    float m20 = md.x, m21 = md.y, m22 = md.z;
    float c0 = m11 * m22 - m21 * m12;
    float c1 = m01 * m22 - m21 * m02;
    float c2 = m01 * m12 - m11 * m02;
    float ood = 1.f / (m00 * c0 - m10 * c1 + m20 * c2);
    float i00 = c0 * ood;
    float i10 = -(m10 * m22 - m20 * m12) * ood;
    float i20 = (m10 * m21 - m20 * m11) * ood;
    float i01 = -c1 * ood;
    float i11 = (m00 * m22 - m20 * m02) * ood;
    float i21 = -(m00 * m21 - m20 * m01) * ood;
    float i02 = c2 * ood;
    float i12 = -(m00 * m12 - m10 * m02) * ood;
    float i22 = (m00 * m11 - m10 * m01) * ood;
    V3 res;
    res.x = i00 * P.x + i10 * P.y + i20 * P.z;
    res.y = i01 * P.x + i11 * P.y + i21 * P.z;
    res.z = i02 * P.x + i12 * P.y + i22 * P.z;
    return res;
}
// Acceptance of Triangle::Intersect (closest hit, u+v slack EPS=1e-5), Triangle.cpp:39-40
C2_HD bool tri_accept_closest(V3 res) {
    if (res.x != res.x || res.y != res.y || res.z != res.z) return false;
    if (res.x < 0.f || res.x > 1.f || res.y < 0.f || res.y > 1.f || res.x + res.y > 1.0f + 1e-5f || res.z < 0.f)
        return false;
    return true;
}
// Triangle::IntersectTest on the box-clipped interval, Triangle.cpp:83-93
C2_HD bool tri_test(V3 o, V3 d, V3 v0, V3 v1, V3 v2, float tMin, float tMax) {
    if (tMin > tMax) return false;
    V3 res = tri_solve(o, d, v0, v1, v2);
    if (res.x != res.x || res.y != res.y || res.z != res.z) return false;
    if (res.x < 0 || res.x > 1 || res.y < 0 || res.y > 1 || res.x + res.y > 1 || res.z < 0) return false;
    return res.z >= tMin && res.z <= tMax;
}

// ------------------------------------------------------------------------- sphere
C2_HD V3 mat3_mul(const float* m, V3 v) {  // m: column-major 3x3
    return V3{m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z,
              m[2] * v.x + m[5] * v.y + m[8] * v.z};
}
C2_HD bool sphere_quad(const float* sp, V3 o, V3 d, V3& P, V3& dl, float& t1, float& t2) {
    V3 c{sp[0], sp[1], sp[2]};
    float r = sp[3];
    P = mat3_mul(sp + 13, o - c);
    dl = mat3_mul(sp + 13, d);
    float a = dot(dl, dl);
    float b = 2 * dot(dl, P);
    float cc = dot(P, P) - r * r;
    float disc = b * b - 4 * a * cc;
    if (disc < 0) return false;
    disc = sqrtf(disc);
    t1 = (-b + disc) / (2 * a);
    t2 = (-b - disc) / (2 * a);
    return true;
}
// Disk::Intersect / IntersectTest, Disk.cpp:45-81: the local plane y = 0 within the radius
C2_HD bool disk_plane(const float* sp, V3 o, V3 d, V3& P, V3& dl, float& t) {
    V3 c{sp[0], sp[1], sp[2]};
    P = mat3_mul(sp + 13, o - c);
    dl = mat3_mul(sp + 13, d);
    if (dl.y == 0.f) return false;
    t = -P.y / dl.y;
    return true;
}
C2_HD bool disk_closest(const float* sp, V3 o, V3 d, float& t) {
    V3 P, dl;
    float tt;
    if (!disk_plane(sp, o, d, P, dl, tt)) return false;
    if (tt < 0) return false;
    V3 hit = P + dl * tt;
    if (dot(hit, hit) > sp[3] * sp[3]) return false;
    t = tt;
    return true;
}
C2_HD bool disk_test(const float* sp, V3 o, V3 d, float tMin, float tMax) {
    V3 P, dl;
    float t;
    if (!disk_plane(sp, o, d, P, dl, t)) return false;
    if (t < tMin || t > tMax) return false;
    V3 hit = P + dl * t;
    return !(dot(hit, hit) > sp[3] * sp[3]);
}
// Sphere::Intersect distance selection, Sphere.cpp:44-49
C2_HD bool sphere_closest(const float* sp, V3 o, V3 d, float& t) {
    V3 P, dl;
    float t1, t2;
    if (!sphere_quad(sp, o, d, P, dl, t1, t2)) return false;
    if (t1 > t2) { float x = t1; t1 = t2; t2 = x; }
    t = t1;
    if (t1 != t1 || t1 < 0) t = t2;
    if (t != t || t < 0) return false;
    return true;
}
C2_HD bool sphere_test(const float* sp, V3 o, V3 d, float tMin, float tMax) {
    if (tMin > tMax) return false;
    V3 P, dl;
    float t1, t2;
    if (!sphere_quad(sp, o, d, P, dl, t1, t2)) return false;
    return (t1 >= tMin && t1 <= tMax) || (t2 >= tMin && t2 <= tMax);
}

// ------------------------------------------------------------------------- traversal
struct Hit {
    int32_t entity;  // -1: miss
    float t, u, v;
};

// Stack entry: >= 0 inner node; < 0 leaf = ~(slot | (count - 1) << 28).

__device__ __forceinline__ void load_prim(const DevScene& S, int slot, float4& a, float4& b, float4& c) {
    a = S.prims[3 * slot + 0];
    b = S.prims[3 * slot + 1];
    c = S.prims[3 * slot + 2];
}

// A float4 that every active lane of the wavefront reads at the same address `p` (derived
// from __builtin_amdgcn_readfirstlane): read through the constant address space, so the
// compiler issues scalar loads into SGPRs (the scalar data cache) instead of one vector
// load per lane through the address unit. Only for memory no kernel writes while it runs
// (the BVH nodes, the KD-tree's nodes and leaf records); the scalar cache is invalidated at
// every dispatch.
typedef float c2f4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 load_uniform4(const char* p) {
    typedef const __attribute__((address_space(4))) c2f4_t* ConstPtr;
    const c2f4_t v = *(ConstPtr)p;  // an address-space cast: C++ casts do not convert address spaces
    return make_float4(v.x, v.y, v.z, v.w);
}

// Closest hit (ANY = false) or any hit (ANY = true). Closest hit: tmin = 0 and
// tmax = ray tmax (the reference passes 0, +inf). Returns true on a hit.
//
// The traversal stack: the top K entries of each lane live in LDS (a [K][block size]
// slab per block, slot-major, so the lanes of a wavefront touch consecutive words
// and no two lanes share a bank beyond the wave64 halves); entries deeper than K
// move to a private overflow array, which only deep traversals touch. The stack
// never holds more than kStackSize entries (at most 3 pushes per level of a tree
// at most kMaxBvh4Depth deep).
constexpr int kTravBlock = 256;  // block size of the wavefront traversal kernels (the stack default)
#ifndef C2_TRAV_K_ANY
#define C2_TRAV_K_ANY 16
#endif
#ifndef C2_TRAV_K_CLOSEST
#define C2_TRAV_K_CLOSEST 8
#endif
template <bool ANY>
struct TravK {
    static constexpr int value = ANY ? C2_TRAV_K_ANY : C2_TRAV_K_CLOSEST;
};
// LDS backing of the stack for one block of BS threads; declare as
// `__shared__ TravLds<ANY> lds;` (BS = kTravBlock) or `TravLds<ANY, BS>`.
template <bool ANY, int BS = kTravBlock>
struct TravLds {
    static constexpr int kBS = BS;
    int32_t idx[TravK<ANY>::value * BS];
    float t[ANY ? 1 : TravK<ANY>::value * BS];
};

// Overflow part of a lane's stack, in private memory: the entries pushed beyond the K
// that the LDS ring holds.
template <bool ANY, int SS = kStackSize>
struct TravOverflow {
    static constexpr int K = TravK<ANY>::value;
    int32_t i[SS - K];
    float t[ANY ? 1 : SS - K];
};
// The stack proper holds only scalars and pointers, so that the stack pointer and lane
// stay in registers (an array member would put the whole object in private memory).
template <bool ANY, int BS = kTravBlock, int SS = kStackSize>
struct TravStack {
    static constexpr int K = TravK<ANY>::value;
    static_assert((K & (K - 1)) == 0 && K > 0 && K <= SS, "LDS stack depth must be a power of two");
    int32_t* li;
    float* lt;
    int32_t* oi;
    float* ot;
    int lane;
    int sp;
    __device__ __forceinline__ TravStack(TravLds<ANY, BS>& L, TravOverflow<ANY, SS>& O)
        : li(L.idx), lt(L.t), oi(O.i), ot(O.t), lane((int)threadIdx.x), sp(0) {}
    __device__ __forceinline__ void push(int32_t e, float t) {
        int off = (sp & (K - 1)) * BS + lane;
        if (sp >= K) {  // evict the oldest resident entry (same ring slot) to the overflow array
            oi[sp - K] = li[off];
            if (!ANY) ot[sp - K] = lt[off];
        }
        li[off] = e;
        if (!ANY) lt[off] = t;
        sp++;
    }
    __device__ __forceinline__ void pop(int32_t& e, float& t) {
        sp--;
        int off = (sp & (K - 1)) * BS + lane;
        e = li[off];
        t = ANY ? 0.f : lt[off];
        if (sp >= K) {  // refill the freed ring slot with the entry K below the new top
            li[off] = oi[sp - K];
            if (!ANY) lt[off] = ot[sp - K];
        }
    }
};

// The primitives of one leaf child word `cur` (< 0), with the per-entity box pre-clip
// [tmin, min(tmax, best)] (closest) / [tmin, tmax] (any). Any hit: returns true when one
// occludes (hit.entity set). Closest hit: updates (best, bestE, bestU, bestV), ties to the
// lowest entity index.
template <bool ANY>
__device__ __forceinline__ bool leaf_prims(const DevScene& S, const RayPre& r, V3 o, V3 d, float tmin, float tmax,
                                           int ignore, int32_t cur, float& best, int32_t& bestE, float& bestU,
                                           float& bestV, Hit& hit) {
    uint32_t enc = (uint32_t)(~cur);
    int slot = (int)(enc & ((1u << kSlotBits) - 1u));
    int n = (int)(enc >> kSlotBits) + 1;
    for (int i = 0; i < n; i++) {
        float4 a, b, c;
        load_prim(S, slot + i, a, b, c);
        int32_t ent = __float_as_int(c.y);
        int32_t kind = __float_as_int(c.z);
        float t1 = tmin, t2 = ANY ? tmax : fmin_ref(tmax, best);
        if (kind == C2_SHAPE_TRIANGLE) {
            V3 v0{a.x, a.y, a.z}, v1{a.w, b.x, b.y}, v2{b.z, b.w, c.x};
            // entity box (component-wise min / max of finite vertices: the IEEE
            // forms equal std::min / std::max up to the sign of a zero, which
            // no comparison of the slab test sees)
            float lxx = fminf(fminf(v0.x, v1.x), v2.x), lyy = fminf(fminf(v0.y, v1.y), v2.y);
            float lzz = fminf(fminf(v0.z, v1.z), v2.z), hxx = fmaxf(fmaxf(v0.x, v1.x), v2.x);
            float hyy = fmaxf(fmaxf(v0.y, v1.y), v2.y), hzz = fmaxf(fmaxf(v0.z, v1.z), v2.z);
            if (!slab(r, lxx, lyy, lzz, hxx, hyy, hzz, t1, t2)) continue;
            if (ANY) {
                if (ent == ignore) continue;
                if (tri_test(o, d, v0, v1, v2, t1, t2)) {
                    hit.entity = ent;
                    return true;
                }
            } else {
                V3 res = tri_solve(o, d, v0, v1, v2);
                if (!tri_accept_closest(res)) continue;
                float dis = res.z;
                if (dis < best || (dis == best && ent < bestE)) {
                    best = dis;
                    bestE = ent;
                    bestU = res.x;
                    bestV = res.y;
                }
            }
        } else if (kind == C2_SHAPE_DISK) {
            // box: centre -+ (r, r * sinTheta, r), Disk::GetBoundingBox (b.y holds r * sinTheta)
            const float* sp3 = S.spheres + (size_t)__float_as_int(b.x) * 22;
            float rr = a.w, oy = b.y;
            if (!slab(r, a.x - rr, a.y - oy, a.z - rr, a.x + rr, a.y + oy, a.z + rr, t1, t2)) continue;
            if (ANY) {
                if (ent == ignore) continue;
                if (disk_test(sp3, o, d, t1, t2)) {
                    hit.entity = ent;
                    return true;
                }
            } else {
                float dis;
                if (!disk_closest(sp3, o, d, dis)) continue;
                if (dis < best || (dis == best && ent < bestE)) {
                    best = dis;
                    bestE = ent;
                    bestU = 0.f;
                    bestV = 0.f;
                }
            }
        } else {
            const float* sp3 = S.spheres + (size_t)__float_as_int(b.x) * 22;
            float rr = a.w;
            if (!slab(r, a.x - rr, a.y - rr, a.z - rr, a.x + rr, a.y + rr, a.z + rr, t1, t2)) continue;
            if (ANY) {
                if (ent == ignore) continue;
                if (sphere_test(sp3, o, d, t1, t2)) {
                    hit.entity = ent;
                    return true;
                }
            } else {
                float dis;
                if (!sphere_closest(sp3, o, d, dis)) continue;
                if (dis < best || (dis == best && ent < bestE)) {
                    best = dis;
                    bestE = ent;
                    bestU = 0.f;
                    bestV = 0.f;
                }
            }
        }
    }
    return false;
}

// root: Node4 index where the traversal starts (a batch keeps every scene's tree in
// one node array); a negative root is a scene without entities (always a miss).
template <bool ANY, int BS = kTravBlock>
__device__ __forceinline__ bool traverse(const DevScene& S, int32_t root, V3 o, V3 d, float tmin, float tmax,
                                         int ignore, Hit& hit, TravLds<ANY, BS>& lds) {
    hit.entity = -1;
    if (root < 0) return false;
    RayPre r = make_ray(o, d);
    // byte offset of each axis's near plane within its (lo, hi) pair
    const uint32_t oX = r.neg[0] ? 16u : 0u, oY = r.neg[1] ? 16u : 0u, oZ = r.neg[2] ? 16u : 0u;
    TravOverflow<ANY> ovf;
    TravStack<ANY, BS> st(lds, ovf);
    int32_t cur = root;
    float best = INFINITY;
    int32_t bestE = -1;
    float bestU = 0.f, bestV = 0.f;
    while (true) {
        if (cur >= 0) {
            // The same slab arithmetic for the four children at once. The near plane
            // of an axis is the child's hi plane when the direction is negative: it is
            // picked by address (float4 index 2 * axis + neg), which is the reference's
            // swap of the two plane distances. Empty slots hold inverted boxes
            // (+inf / -inf), whose interval is empty for every direction.
            // 32-bit byte offsets from the array base (the node array stays below
            // 4 GiB): the loads take a scalar base plus a 32-bit lane offset
            //
            // When every active lane of the wavefront is at the same node (23% of the
            // inner-node steps of shadow-ray waves and 35% of extension-ray waves on the
            // 1080p sphere field), the node comes in through three scalar loads and each
            // lane picks its near and far planes with selects: the same floats reach the
            // same arithmetic. This took the 1080p sphere-field render from 405 to 419-422
            // Msamples/s; scalar loads of leaf primitives as well cost 1-2%.
            const char* nb = reinterpret_cast<const char*>(S.nodes);
            float4 nx, fx, ny, fy, nz, fz;
            int4 ch;
            const int32_t cu = __builtin_amdgcn_readfirstlane(cur);
            if (__ballot(cur != cu) == 0) {
                const char* un = nb + (uint32_t)cu * 128u;
                const float4 lx = load_uniform4(un), hx = load_uniform4(un + 16);
                const float4 ly = load_uniform4(un + 32), hy = load_uniform4(un + 48);
                const float4 lz = load_uniform4(un + 64), hz = load_uniform4(un + 80);
                const float4 c4 = load_uniform4(un + 96);
                nx = r.neg[0] ? hx : lx;
                fx = r.neg[0] ? lx : hx;
                ny = r.neg[1] ? hy : ly;
                fy = r.neg[1] ? ly : hy;
                nz = r.neg[2] ? hz : lz;
                fz = r.neg[2] ? lz : hz;
                ch = make_int4(__float_as_int(c4.x), __float_as_int(c4.y), __float_as_int(c4.z), __float_as_int(c4.w));
            } else {
                const uint32_t no = (uint32_t)cur * 128u;
                nx = *reinterpret_cast<const float4*>(nb + (no + oX));
                fx = *reinterpret_cast<const float4*>(nb + (no + (16u - oX)));
                ny = *reinterpret_cast<const float4*>(nb + (no + 32u + oY));
                fy = *reinterpret_cast<const float4*>(nb + (no + (48u - oY)));
                nz = *reinterpret_cast<const float4*>(nb + (no + 64u + oZ));
                fz = *reinterpret_cast<const float4*>(nb + (no + (80u - oZ)));
                ch = *reinterpret_cast<const int4*>(nb + (no + 96u));
            }
            float lim = ANY ? tmax : fmin_ref(tmax, best);
            float k0 = INFINITY, k1 = INFINITY, k2 = INFINITY, k3 = INFINITY;
            int32_t e0 = 0, e1 = 0, e2 = 0, e3 = 0;
#define C2_CHILD(I, C, CH, KOUT, EOUT)                                                                \
    {                                                                                                \
        float tn = fmaxf(fmaxf(fmaxf(tmin, (nx.C - r.o.x) * r.invD.x), (ny.C - r.o.y) * r.invD.y),      \
                         (nz.C - r.o.z) * r.invD.z);                                                 \
        float tf = fminf(fminf(fminf(lim, (fx.C - r.o.x) * r.invD.x), (fy.C - r.o.y) * r.invD.y),       \
                         (fz.C - r.o.z) * r.invD.z);                                                 \
        if (!(tf < tn)) {                                                                            \
            KOUT = tn;                                                                               \
            EOUT = CH;                                                                               \
        }                                                                                            \
    }
            C2_CHILD(0, x, ch.x, k0, e0)
            C2_CHILD(1, y, ch.y, k1, e1)
            C2_CHILD(2, z, ch.z, k2, e2)
            C2_CHILD(3, w, ch.w, k3, e3)
#undef C2_CHILD
            // sort the four (key, entry) pairs ascending; misses carry +inf keys
// (as selects: no branch, no exec-mask change)
#define C2_CSWAP(KA, EA, KB, EB)                                                 \
    {                                                                            \
        const bool sw = KB < KA;                                                 \
        const float lo = sw ? KB : KA, hi = sw ? KA : KB;                        \
        const int32_t elo = sw ? EB : EA, ehi = sw ? EA : EB;                    \
        KA = lo;                                                                 \
        KB = hi;                                                                 \
        EA = elo;                                                                \
        EB = ehi;                                                                \
    }
            // near-first also for any hit: measured faster than a fixed order
            C2_CSWAP(k0, e0, k1, e1)
            C2_CSWAP(k2, e2, k3, e3)
            C2_CSWAP(k0, e0, k2, e2)
            C2_CSWAP(k1, e1, k3, e3)
            C2_CSWAP(k1, e1, k2, e2)
#undef C2_CSWAP
            if (k3 != INFINITY) st.push(e3, k3);
            if (k2 != INFINITY) st.push(e2, k2);
            if (k1 != INFINITY) st.push(e1, k1);
            if (k0 != INFINITY) {
                cur = e0;
                continue;
            }
        } else {
            if (leaf_prims<ANY>(S, r, o, d, tmin, tmax, ignore, cur, best, bestE, bestU, bestV, hit)) return true;
        }
        // pop; a closest-hit entry whose entry distance is already beyond the best hit
        // cannot contain a closer one (its slab interval starts after `best`).
        while (true) {
            if (st.sp == 0) {
                if (ANY) return false;
                hit.entity = bestE;
                hit.t = best;
                hit.u = bestU;
                hit.v = bestV;
                return bestE >= 0;
            }
            int32_t e;
            float et;
            st.pop(e, et);
            if (!ANY && et > best) continue;
            cur = e;
            break;
        }
    }
}

// The same traversal over the 8-wide tree (Node8, C2_ACCEL_BVH8): the eight children's slab
// tests from 14 loads (two 128-byte lines), the hit children sorted by entry distance with
// Batcher's 19-comparator network (as selects), the nearest taken next and the others pushed
// farthest first. Leaves and ties as in traverse: the result does not depend on the order.
template <bool ANY, int BS = kTravBlock>
__device__ __forceinline__ bool traverse8(const DevScene& S, int32_t root, V3 o, V3 d, float tmin, float tmax,
                                          int ignore, Hit& hit, TravLds<ANY, BS>& lds) {
    hit.entity = -1;
    if (root < 0) return false;
    RayPre r = make_ray(o, d);
    // byte offset of each axis's near plane within its (lo, hi) pair of 32-byte planes
    const uint32_t oX = r.neg[0] ? 32u : 0u, oY = r.neg[1] ? 32u : 0u, oZ = r.neg[2] ? 32u : 0u;
    TravOverflow<ANY, kStackSize8> ovf;
    TravStack<ANY, BS, kStackSize8> st(lds, ovf);
    int32_t cur = root;
    float best = INFINITY;
    int32_t bestE = -1;
    float bestU = 0.f, bestV = 0.f;
    while (true) {
        if (cur >= 0) {
            const char* nb = reinterpret_cast<const char*>(S.nodes8);
            const uint32_t no = (uint32_t)cur * 256u;  // the node array stays below 4 GiB
#define C2_LD4(OFF) (*reinterpret_cast<const float4*>(nb + (no + (OFF))))
            const float4 nx0 = C2_LD4(oX), nx1 = C2_LD4(oX + 16u);
            const float4 fx0 = C2_LD4(32u - oX), fx1 = C2_LD4(48u - oX);
            const float4 ny0 = C2_LD4(64u + oY), ny1 = C2_LD4(80u + oY);
            const float4 fy0 = C2_LD4(96u - oY), fy1 = C2_LD4(112u - oY);
            const float4 nz0 = C2_LD4(128u + oZ), nz1 = C2_LD4(144u + oZ);
            const float4 fz0 = C2_LD4(160u - oZ), fz1 = C2_LD4(176u - oZ);
#undef C2_LD4
            const int4 ch0 = *reinterpret_cast<const int4*>(nb + (no + 192u));
            const int4 ch1 = *reinterpret_cast<const int4*>(nb + (no + 208u));
            const float lim = ANY ? tmax : fmin_ref(tmax, best);
            float k[8];
            int32_t e[8];
#define C2_CHILD8(J, NX, NY, NZ, FX, FY, FZ, C, CH)                                                     \
    {                                                                                                \
        const float tn = fmaxf(fmaxf(fmaxf(tmin, (NX.C - r.o.x) * r.invD.x), (NY.C - r.o.y) * r.invD.y), \
                               (NZ.C - r.o.z) * r.invD.z);                                           \
        const float tf = fminf(fminf(fminf(lim, (FX.C - r.o.x) * r.invD.x), (FY.C - r.o.y) * r.invD.y),  \
                               (FZ.C - r.o.z) * r.invD.z);                                           \
        const bool in = !(tf < tn);                                                                  \
        k[J] = in ? tn : INFINITY;                                                                   \
        e[J] = CH;                                                                                   \
    }
            C2_CHILD8(0, nx0, ny0, nz0, fx0, fy0, fz0, x, ch0.x)
            C2_CHILD8(1, nx0, ny0, nz0, fx0, fy0, fz0, y, ch0.y)
            C2_CHILD8(2, nx0, ny0, nz0, fx0, fy0, fz0, z, ch0.z)
            C2_CHILD8(3, nx0, ny0, nz0, fx0, fy0, fz0, w, ch0.w)
            C2_CHILD8(4, nx1, ny1, nz1, fx1, fy1, fz1, x, ch1.x)
            C2_CHILD8(5, nx1, ny1, nz1, fx1, fy1, fz1, y, ch1.y)
            C2_CHILD8(6, nx1, ny1, nz1, fx1, fy1, fz1, z, ch1.z)
            C2_CHILD8(7, nx1, ny1, nz1, fx1, fy1, fz1, w, ch1.w)
#undef C2_CHILD8
#define C2_CSWAP8(A, B)                                              \
    {                                                                \
        const bool sw = k[B] < k[A];                                 \
        const float lo = sw ? k[B] : k[A], hi = sw ? k[A] : k[B];    \
        const int32_t elo = sw ? e[B] : e[A], ehi = sw ? e[A] : e[B]; \
        k[A] = lo;                                                   \
        k[B] = hi;                                                   \
        e[A] = elo;                                                  \
        e[B] = ehi;                                                  \
    }
            C2_CSWAP8(0, 1) C2_CSWAP8(2, 3) C2_CSWAP8(4, 5) C2_CSWAP8(6, 7)
            C2_CSWAP8(0, 2) C2_CSWAP8(1, 3) C2_CSWAP8(4, 6) C2_CSWAP8(5, 7)
            C2_CSWAP8(1, 2) C2_CSWAP8(5, 6)
            C2_CSWAP8(0, 4) C2_CSWAP8(1, 5) C2_CSWAP8(2, 6) C2_CSWAP8(3, 7)
            C2_CSWAP8(2, 4) C2_CSWAP8(3, 5)
            C2_CSWAP8(1, 2) C2_CSWAP8(3, 4) C2_CSWAP8(5, 6)
#undef C2_CSWAP8
#pragma unroll
            for (int j = 7; j >= 1; j--)
                if (k[j] != INFINITY) st.push(e[j], k[j]);
            if (k[0] != INFINITY) {
                cur = e[0];
                continue;
            }
        } else {
            if (leaf_prims<ANY>(S, r, o, d, tmin, tmax, ignore, cur, best, bestE, bestU, bestV, hit)) return true;
        }
        while (true) {
            if (st.sp == 0) {
                if (ANY) return false;
                hit.entity = bestE;
                hit.t = best;
                hit.u = bestU;
                hit.v = bestV;
                return bestE >= 0;
            }
            int32_t e;
            float et;
            st.pop(e, et);
            if (!ANY && et > best) continue;
            cur = e;
            break;
        }
    }
}

// ------------------------------------------------------------------------- KD-tree
// RayBoxTest(ray, box, tMin, tMax) (src/Core/Geometry.h:118-133): the swap follows the
// sign of the direction itself (not of its reciprocal, as in the BVH's form).
C2_HD bool ray_box_ref(V3 o, V3 d, V3 inv, const float* lo, const float* hi, float& t0, float& t1) {
    const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z}, ii[3] = {inv.x, inv.y, inv.z};
    for (int i = 0; i < 3; i++) {
        float a = (lo[i] - oo[i]) * ii[i], b = (hi[i] - oo[i]) * ii[i];
        if (dd[i] < 0) {
            float t = a;
            a = b;
            b = t;
        }
        t0 = fmax_ref(t0, a);
        t1 = fmin_ref(t1, b);
        if (t1 < t0) return false;
    }
    return true;
}

constexpr int kKdStack = 24;  // > the tree's depth limit of 20: at most one entry per level

// The KD traversal's stack as the BVH's: the top kKdRing entries (node, window start,
// window end) of each lane in an LDS ring at [slot][threadIdx], deeper entries in a
// private overflow array. A register array indexed by the stack pointer would hold all
// kKdStack entries in VGPRs (one wavefront per SIMD). Four resident entries (12 KB per
// 256-thread block) leave the KD kernels at 8 waves per SIMD, which their 45-61 VGPRs
// allow; eight (24 KB) held them at 6 and rendered the 100-sphere field at 1080p 7% slower
// (42.4 against 45.3 Msamples/s, identical films).
constexpr int kKdRing = 4;
struct KdLds {  // 12 KB per block; declare as `__shared__ KdLds lds;`
    int32_t n[kKdRing * kTravBlock];
    float a[kKdRing * kTravBlock], b[kKdRing * kTravBlock];
};
struct KdOverflow {
    int32_t n[kKdStack - kKdRing];
    float a[kKdStack - kKdRing], b[kKdStack - kKdRing];
};
struct KdStack {
    int32_t* ln;
    float* la;
    float* lb;
    int32_t* on;
    float* oa;
    float* ob;
    int lane;
    int sp;
    __device__ __forceinline__ KdStack(KdLds& L, KdOverflow& O)
        : ln(L.n), la(L.a), lb(L.b), on(O.n), oa(O.a), ob(O.b), lane((int)threadIdx.x), sp(0) {}
    __device__ __forceinline__ void push(int32_t n, float a, float b) {
        const int off = (sp & (kKdRing - 1)) * kTravBlock + lane;
        if (sp >= kKdRing) {  // evict the oldest resident entry (same ring slot)
            on[sp - kKdRing] = ln[off];
            oa[sp - kKdRing] = la[off];
            ob[sp - kKdRing] = lb[off];
        }
        ln[off] = n;
        la[off] = a;
        lb[off] = b;
        sp++;
    }
    __device__ __forceinline__ void pop(int32_t& n, float& a, float& b) {
        sp--;
        const int off = (sp & (kKdRing - 1)) * kTravBlock + lane;
        n = ln[off];
        a = la[off];
        b = lb[off];
        if (sp >= kKdRing) {  // refill the freed slot with the entry kKdRing below the new top
            ln[off] = on[sp - kKdRing];
            la[off] = oa[sp - kKdRing];
            lb[off] = ob[sp - kKdRing];
        }
    }
};
// Traversal kinds of the kernels (their template argument ACC): the 4-wide BVH, the KD-tree,
// the 8-wide BVH.
constexpr int kAccBvh4 = 0, kAccKd = 1, kAccBvh8 = 2;
// LDS of a traversal kernel: the KD-tree's stack or the BVH's
template <int ACC, bool ANY>
using TravLdsFor = std::conditional_t<ACC == kAccKd, KdLds, TravLds<ANY>>;

// KD-tree traversal with the reference's windows (KDTree.cpp:73-81, 195-266), recursion
// turned into an explicit stack in the same visiting order:
//  * inner node: t_split = (split - o[axis]) / d[axis]; the near child (below, or above
//    for a negative direction) gets [tMin, min(t_split, tMax, best)] when t_split >= tMin,
//    the far child [max(t_split, 0), min(tMax, best after the near child)] when
//    t_split <= tMax; a node is skipped when tMin > tMax, or when tMin >= best (closest);
//  * closest-hit leaf: each entity's box is tested against [tMin, min(tMax, best)], then
//    the full shape intersection replaces the best hit when strictly closer (first found
//    wins ties); the ray is first clipped to the union box of all entities;
//  * any-hit leaf: the box test as above, then the shape's own test over the node's
//    window [tMin, tMax], skipping the ignored entity; the first hit ends the search.
template <bool ANY>
__device__ __forceinline__ bool kd_traverse(const DevScene& S, int32_t root, const float* mlo, const float* mhi,
                                            V3 o, V3 d, float tMin, float tMax, int ignore, Hit& hit, KdLds& lds) {
    hit.entity = -1;
    if (root <= 0) return false;
    const V3 inv{1.f / d.x, 1.f / d.y, 1.f / d.z};
    if (!ANY && !ray_box_ref(o, d, inv, mlo, mhi, tMin, tMax)) return false;
    float best = INFINITY, bestU = 0.f, bestV = 0.f;
    int32_t bestE = -1;
    KdOverflow ovf;
    KdStack st(lds, ovf);
    int32_t p = root;
    float t0 = tMin, t1 = tMax;
    while (true) {
        if (p != 0 && !(t0 > t1) && (ANY || !(t0 >= best))) {
            // When every active lane is at the same node, the node and its leaf records
            // come in through scalar loads (load_uniform4): the 100-sphere field at 1080p
            // went from 45 to 57 Msamples/s with the same film.
            const int32_t pu = __builtin_amdgcn_readfirstlane(p);
            const bool uni = __ballot(p != pu) == 0;
            int4 nd;
            if (uni) {
                const float4 u = load_uniform4(reinterpret_cast<const char*>(S.kd_nodes + pu));
                nd = make_int4(__float_as_int(u.x), __float_as_int(u.y), __float_as_int(u.z), __float_as_int(u.w));
            } else {
                nd = S.kd_nodes[p];
            }
            if (nd.x < 0) {
                const int32_t ubase = __builtin_amdgcn_readfirstlane(nd.z);  // used when uni
                for (int k = 0; k < nd.w; k++) {
                    // one record per leaf entry: its five loads are independent
                    float4 blo, bhi, c0, c1, c2;
                    if (uni) {
                        const char* ur = reinterpret_cast<const char*>(S.kd_rec + (size_t)(ubase + k) * 5);
                        blo = load_uniform4(ur);
                        bhi = load_uniform4(ur + 16);
                        c0 = load_uniform4(ur + 32);
                        c1 = load_uniform4(ur + 48);
                        c2 = load_uniform4(ur + 64);
                    } else {
                        const float4* rec = S.kd_rec + (size_t)(nd.z + k) * 5;
                        blo = rec[0];
                        bhi = rec[1];
                        c0 = rec[2];
                        c1 = rec[3];
                        c2 = rec[4];
                    }
                    const int32_t e = __float_as_int(blo.w), kind = __float_as_int(bhi.w);
                    const float lo[3] = {blo.x, blo.y, blo.z}, hi[3] = {bhi.x, bhi.y, bhi.z};
                    float a = t0, b = ANY ? t1 : fmin_ref(t1, best);
                    if (!ray_box_ref(o, d, inv, lo, hi, a, b)) continue;
                    if (ANY && e == ignore) continue;
                    if (kind == C2_SHAPE_TRIANGLE) {
                        V3 v0{c0.x, c0.y, c0.z}, v1{c0.w, c1.x, c1.y}, v2{c1.z, c1.w, c2.x};
                        if (ANY) {
                            if (tri_test(o, d, v0, v1, v2, t0, t1)) {
                                hit.entity = e;
                                return true;
                            }
                        } else {
                            V3 res = tri_solve(o, d, v0, v1, v2);
                            if (tri_accept_closest(res) && res.z < best) {
                                best = res.z;
                                bestE = e;
                                bestU = res.x;
                                bestV = res.y;
                            }
                        }
                    } else {
                        const float* sp3 = S.spheres + (size_t)__float_as_int(c0.x) * 22;
                        const bool disk = kind == C2_SHAPE_DISK;
                        if (ANY) {
                            if (disk ? disk_test(sp3, o, d, t0, t1) : sphere_test(sp3, o, d, t0, t1)) {
                                hit.entity = e;
                                return true;
                            }
                        } else {
                            float dis;
                            bool h = disk ? disk_closest(sp3, o, d, dis) : sphere_closest(sp3, o, d, dis);
                            if (h && dis < best) {
                                best = dis;
                                bestE = e;
                                bestU = 0.f;
                                bestV = 0.f;
                            }
                        }
                    }
                }
            } else {
                const int axis = nd.x;
                const float split = __int_as_float(nd.y);
                const float oa = axis == 0 ? o.x : (axis == 1 ? o.y : o.z);
                const float da = axis == 0 ? d.x : (axis == 1 ? d.y : d.z);
                const float ts = (split - oa) / da;
                const bool neg = da < 0;
                const int32_t nearC = neg ? nd.w : nd.z, farC = neg ? nd.z : nd.w;
                if (ts <= t1 && st.sp < kKdStack) st.push(farC, fmax_ref(ts, 0.f), t1);  // after the near subtree
                if (ts >= t0) {
                    p = nearC;
                    t1 = ANY ? fmin_ref(ts, t1) : fmin_ref(fmin_ref(ts, t1), best);
                    continue;
                }
            }
        }
        if (st.sp == 0) break;
        float w1;
        st.pop(p, t0, w1);
        t1 = ANY ? w1 : fmin_ref(w1, best);
    }
    if (ANY) return false;
    hit.entity = bestE;
    hit.t = best;
    hit.u = bestU;
    hit.v = bestV;
    return bestE >= 0;
}

// One traversal of kind ACC from `root` (KD-tree: its root and union box kdlo / kdhi).
template <int ACC, bool ANY, class Lds>
__device__ __forceinline__ bool accel_trace(const DevScene& S, int32_t root, const float* kdlo, const float* kdhi, V3 o,
                                            V3 d, float tmin, float tmax, int ignore, Hit& h, Lds& lds) {
    if constexpr (ACC == kAccKd)
        return kd_traverse<ANY>(S, root, kdlo, kdhi, o, d, tmin, tmax, ignore, h, lds);
    else if constexpr (ACC == kAccBvh8)
        return traverse8<ANY>(S, root, o, d, tmin, tmax, ignore, h, lds);
    else
        return traverse<ANY>(S, root, o, d, tmin, tmax, ignore, h, lds);
}

}  // namespace c2
