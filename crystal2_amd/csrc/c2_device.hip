// gfx950 kernels and the device half of the C-ABI (include/crystal2.h):
// Accelerator (4-wide BVH closest / any hit) and the wavefront PathTracer + MIS.
//
// Wavefront chunk: Halton samples k .. k + nsamp - 1 of a range of pixels (paths in
// flight; the single renderer puts four samples of every pixel in one chunk):
//   generate  camera ray + sampler state           SamplerIntegrator.cpp:42-52, PerspCamera.cpp:23-29
//   per bounce b < Depth:
//     extend   closest hit of every queued path     BVH::Intersect
//     (b >= 1: the queue grouped by shading class, k_class_count / k_class_scatter)
//     shade    emission, BSDF, BSDF-sampled MIS direction (one per vertex, see below),
//              continuation + Russian roulette      PathTracingIntegrator.cpp:36-109
//     nee      light samples -> shadow slots, and whether the MIS ray is needed
//                                                   PathTracingIntegrator.cpp:175-278
//     occlude  any hit of every shadow ray          BVH::IntersectTest
//     mis      closest hit of every MIS ray (side stream, beside occlude)
//     resolve  L += beta * sum over lights (in the reference's light order) of the
//              unshadowed light samples and the MIS term, then beta <- beta'
//   accumulate film += L (NaN/Inf -> 0), one thread per pixel adding its samples in
//              sample order                         FrameBuffer.cpp:20-32
// Consecutive chunks alternate between two lanes (path buffers + stream each); the
// film adds stay in chunk order through an event chain (enqueue_passes).
//
// MIS ray: every supported material has a single BxDF, so the BSDF-sampled
// direction in EsimateDirect (PathTracingIntegrator.cpp:240) is the same for all
// lights (same sampleBSDF, BxDF index 0). The reference traces that ray once per
// light whose own shape it hits (Pdf_Li != 0); here it is traced once and only
// the light owning the hit entity can receive the term, which is exactly the
// reference's sum.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "c2_bvh.h"
#include "c2_kdtree.h"
#include "c2_sched.h"
#include "c2_device.h"
#include "c2_host.h"
#include "c2_nrc.h"
#include "c2_shade.h"

namespace c2 {

#define HIP_CHECK(x)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess)                                                                         \
            fail(C2_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_));                      \
    } while (0)

constexpr int kBlock = 256;
static_assert(kBlock == kTravBlock, "traversal kernels index their LDS stack by threadIdx.x");
// blocks for the MIS-ray kernel (grid-stride; about 0.005 rays per sample, a few ten thousand
// per launch): a small grid is dispatched and drained early instead of holding block slots
// beside the shadow-ray kernel for its whole duration
constexpr int kMisGrid = 64;
// Chunks of fewer paths trace their shadow and MIS rays in line on the lane's stream rather
// than forking the shadow rays to the side stream (direct lighting always does). Measured on
// the sphere field (Msamples/s, in line against forked): 1280 x 720, 4.6 M paths per chunk,
// 383.7 against 376-380; 960 x 540, 2.6 M, 334-335 both; 1920 x 1080, 10.4 M, 421-423 against
// 424-425. The Cornell box (1 M paths): 601 against 591, its direct lighting 2,057 against
// 1,523, and direct lighting on the 1080p sphere field 1,400-1,406 against 1,384-1,396.
constexpr int64_t kInlineShadowPaths = 6ll << 20;
// Direct lighting: wavefront steps per path (transparent surfaces crossed + 1).
constexpr int kDlChainLimit = 8;
// Default grid cap of the per-bounce queue kernels. Their live counts are only known
// on the device, so they are launched for the upper bound with grid-stride loops;
// the cap keeps a shrinking queue from paying for tens of thousands of empty blocks.
constexpr int kQueueBlocksDefault = 8192;
// Path-buffer lanes: consecutive chunks (sample passes) alternate between two sets of
// path buffers on two streams, so the long, thinly populated last bounces of one pass
// share the GPU with the first bounces of the next. Film accumulation stays in chunk
// order (an event chain), so the films equal those of a single lane bit for bit.
#ifndef C2_LANES
#define C2_LANES 2
#endif
#ifndef C2_MIS_SIDE
#define C2_MIS_SIDE 1
#endif
constexpr int kLanes = C2_LANES;
static_assert(kLanes >= 1 && kLanes <= 4, "1 to 4 lanes");
// Samples per chunk: a launch carries up to this many consecutive samples of every pixel
// (more rays per launch in the thinly populated late bounces).
#ifndef C2_SAMPLES_PER_CHUNK
#define C2_SAMPLES_PER_CHUNK 4
#endif
constexpr int kSamplesPerChunk = C2_SAMPLES_PER_CHUNK;
// Small frames: a launch carries more samples, up to kSmallFrameSamples, so that it holds
// about kSmallFramePaths paths (a 256 x 256 frame: 16 samples per launch instead of 4;
// frames from 512 x 512 up keep kSamplesPerChunk). Launches of a few hundred thousand paths
// are bound by launch latency, not by the GPU.
#ifndef C2_SMALL_FRAME_PATHS
#define C2_SMALL_FRAME_PATHS (1 << 20)
#endif
#ifndef C2_SMALL_FRAME_SAMPLES
#define C2_SMALL_FRAME_SAMPLES 16
#endif
constexpr int64_t kSmallFramePaths = C2_SMALL_FRAME_PATHS;
constexpr int kSmallFrameSamples = C2_SMALL_FRAME_SAMPLES;
// Samples per launch of a single renderer whose largest pass has npix paths per sample
// (`base` samples for frames of kSmallFramePaths / base paths and more).
inline int samples_per_launch(int base, int64_t npix) {
    if (npix <= 0) return base;
    const int64_t want = (kSmallFramePaths + npix - 1) / npix;
    return (int)std::max<int64_t>(base, std::min<int64_t>(want, kSmallFrameSamples));
}
#ifndef C2_BATCH_SAMPLES_PER_CHUNK
#define C2_BATCH_SAMPLES_PER_CHUNK 16
#endif
constexpr int kBatchSamplesPerChunk = C2_BATCH_SAMPLES_PER_CHUNK;  // a batch (one lane)
#ifndef C2_DIRECT_SAMPLES_PER_CHUNK
#define C2_DIRECT_SAMPLES_PER_CHUNK 8
#endif
constexpr int kDirectSamplesPerChunk = C2_DIRECT_SAMPLES_PER_CHUNK;  // direct lighting (short paths)
// With the radiance cache: samples per launch, each with its own training pixels; the
// launch's training rows are split into one Adam step per sample, taken in order
// after the launch (the next launch's queries see all of them).
#ifndef C2_CACHE_SAMPLES_PER_CHUNK
#define C2_CACHE_SAMPLES_PER_CHUNK 4
#endif
constexpr int kCacheSamplesPerChunk = C2_CACHE_SAMPLES_PER_CHUNK;
// Shading classes (k_class_count) and the first bounce whose queue is grouped by them
constexpr int kClasses = 32;
#ifndef C2_SORT_FROM
#define C2_SORT_FROM 1
#endif
constexpr int kSortFrom = C2_SORT_FROM;
// MIS rays on a side stream of each lane, concurrent with its shadow rays (else in
// line on the lane's stream)
constexpr bool kMisSide = C2_MIS_SIDE != 0;

// ============================================================================ accelerator kernels

// Master box of a KD-tree scene (KDTree::Intersect clips the ray to it first)
struct KdMaster {
    float lo[3], hi[3];
    int32_t root;
};

template <int ACC>
__global__ void __launch_bounds__(kBlock) k_intersect(DevScene S, const float* __restrict__ rays, int64_t n,
                                                      int32_t* __restrict__ prim, float* __restrict__ tout, KdMaster M) {
    __shared__ TravLdsFor<ACC, false> lds;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float4* r4 = reinterpret_cast<const float4*>(rays + i * 8);
        float4 a = r4[0], b = r4[1];
        Hit h;
        bool hit;
        hit = accel_trace<ACC, false>(S, ACC == kAccKd ? M.root : 0, M.lo, M.hi, V3{a.x, a.y, a.z},
                                      V3{a.w, b.x, b.y}, 0.f, b.w, -1, h, lds);
        prim[i] = hit ? h.entity : -1;
        tout[i] = hit ? h.t : INFINITY;
    }
}

template <int ACC>
__global__ void __launch_bounds__(kBlock) k_intersect_test(DevScene S, const float* __restrict__ rays,
                                                           const int32_t* __restrict__ ignore, int64_t n,
                                                           uint8_t* __restrict__ out, KdMaster M) {
    __shared__ TravLdsFor<ACC, true> lds;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float4* r4 = reinterpret_cast<const float4*>(rays + i * 8);
        float4 a = r4[0], b = r4[1];
        Hit h;
        const int ig = ignore ? ignore[i] : -1;
        bool hit;
        hit = accel_trace<ACC, true>(S, ACC == kAccKd ? M.root : 0, M.lo, M.hi, V3{a.x, a.y, a.z},
                                     V3{a.w, b.x, b.y}, b.z, b.w, ig, h, lds);
        out[i] = hit ? 1 : 0;
    }
}

// Results of rays cast into an empty scene: no hit (prim -1, t +inf), no occluder.
// Kernels rather than hipMemsetAsync, so callers may capture these entry points into
// a graph.
__global__ void k_fill_miss(int32_t* __restrict__ prim, float* __restrict__ tout, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        prim[i] = -1;
        tout[i] = INFINITY;
    }
}

__global__ void k_fill_u8(uint8_t* __restrict__ p, int64_t n, uint8_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ============================================================================ wavefront state

enum : int32_t {
    F_LIGHTPATH = 1,  // add emission at the next hit (specular or transmission bounce)
    F_CONTINUE = 2,   // path continues after this vertex
    F_NEE = 4,        // this vertex ran UniformSampleAllLights (resolve adds beta * U)
    F_MIS = 8,        // MIS ray queued for this vertex
    F_TRAIN = 16,     // radiance-cache training path (index in bits 8..31)
    F_QUERY = 32,     // the path ended in the radiance cache; meta.z holds its query slot
};
// Flags of a bounce's vertex record (vtx_i.w)
enum : int32_t {
    V_FRONT = 1,     // the ray hit the front face
    V_NEE = 2,       // the light loop runs at this vertex
    V_CONTINUE = 4,  // the path continues (resolve sets beta <- beta2)
};

// Path state (ray, throughput, radiance, sampler state, hit) is indexed by path p and
// lives across bounces. A bounce's vertex records are indexed by the vertex's position i
// in the bounce's queue (the order k_surface, k_shade, k_nee and k_resolve all walk), so
// those kernels read and write them coalesced although p is scattered once the queue is
// grouped by shading class. Shadow rays are packed per wavefront: the 64 vertices
// i = 64w .. 64w + 63 own slots [64w * lmax, (64w + 64) * lmax), and their rays fill them
// from the start, light by light, each light's rays in lane order (see wave_slot).
struct PathBuffers {
    float4* ray_o;   // [p] o.xyz
    float4* ray_d;   // [p] d.xyz
    float4* beta;    // [p] beta at the current vertex
    float4* L;       // [p] path radiance
    int4* meta;      // [p] film pixel (all views' films concatenated), halton index, dimension, flags
    int32_t* vid;    // [p] view (scene of a batch) of the path
    float4* hit;     // [p] t, u, v, entity bits
    float4* beta2;   // [i] beta after the continuation (applied by resolve)
    float4* vtx_p;   // [i] vertex position, uL.x
    float4* vtx_n;   // [i] facing shading normal, uL.y
    float4* vtx_du;  // [i] dpdu, (BSDF-sampled f != 0) bits
    float4* vtx_d;   // [i] incoming ray direction, pdf of the BSDF-sampled direction
    int4* vtx_i;     // [i] entity, material, attached light, V_* flags (every queued vertex)
    int2* vtx_h;     // [i] Halton index, view
    int32_t* texel;  // [i] Kd texel of the extension hit (textured materials only)
    float4* mis_w;   // [i] BSDF-sampled direction (shared by all lights)
    float4* mis_f;   // [i] f * cos for that direction
    float4* mis_L;   // [i] resolved MIS term
    int32_t* mis_hit;  // [i] light owning the MIS hit, -1 none
    int32_t* nee_out;  // [i] 1 = the MIS ray is traced
    uint32_t* sh_mask; // [g * cap + i] bit j: light 32g + j of vertex i emitted a shadow ray
    const int32_t* elig;  // eligible lights (Flux != 0) in light order
    float4* sh_o;    // [slot] shadow ray origin, ignore entity bits
    float4* sh_d;    // [slot] shadow ray direction, BVH root bits
    float4* sh_c;    // [slot] unshadowed contribution
    uint8_t* sh_vis; // [slot] 1 = unoccluded, 0 = occluded
    int32_t* queue[2];
    // radiance cache
    float4* nrc_feat;   // [query][kNrcRec4]: packed query features, then the throughput (replaced by
                        // throughput x prediction by the inference), by query slot
    int32_t* tlist;     // training path -> path
    float4* trec;       // [training path][depth][kNrcQ4 + 2]: packed query, beta, L after emission
    uint8_t* trec_valid;
    float4* bfeat;      // training batch features [cap][kNrcQ4]
    float4* btarget;    // training batch targets [cap]
    int32_t* sh_list;   // slots of the bounce's shadow rays
    int32_t* mis_list;  // vertices (queue positions) tracing the MIS ray
    int32_t* queue_s;   // this bounce's queue grouped by shading class (see k_class_count)
    uint8_t* cls;       // shading class per queue index
    int32_t* ccount;    // [kClasses] paths per class, [kClasses] scatter cursors
    int32_t cap;        // paths the buffers hold (sh_mask's group stride)
    int32_t* counters;  // [0..1] queue counts, [2] shadow, [3] mis, [4] cache queries, [5] training
                        // paths, [6] captured rays, [7] training rows
    unsigned long long* totals;  // extension, shadow, mis rays (64-bit running totals)
};

// One rendered image: a scene with its camera, film and sampler. A single renderer
// has one view; a batch has one per scene, all sharing one concatenated scene.
struct View {
    int32_t width, height;
    int32_t spp;             // samples per pixel of the view
    int32_t root;            // Node4 index of the scene's BVH root, -1 = no entities
    int32_t root8;           // Node8 index of its 8-wide root (renderers with the 8-wide tree)
    int32_t elig_begin;      // eligible lights (Flux != 0): elig[elig_begin, + num_elig)
    int32_t num_elig;
    int32_t nd_elig;         // eligible non-delta lights (one Get1D each per vertex)
    int32_t stride0, exp0, scale1, mult0, mult1, bscale0, bexp1;  // Halton constants
    int32_t win_x0, win_y0, win_w, win_h;  // pixel window receiving samples
    int64_t film_off;        // first pixel of this view's film in the concatenated film
    V3 eye, look, up, right;
    float camW, camH;
    int32_t sky[6];          // skybox face texture ids (posx, negx, posy, negy, posz, negz), 0 = none
    float kd_lo[3], kd_hi[3];  // KD-tree scenes: union box of the entities (root is the KD root)
};

struct PassParams {
    int32_t sample;          // Halton sample number k of the chunk's first sample
    int32_t pix_begin, pix_count;  // linear pass indices of this chunk
    int32_t nsamp;           // samples k .. k + nsamp - 1 of those indices: path i is sample
                             // k + i / pix_count of linear index pix_begin + i % pix_count
    int32_t max_depth;
    int32_t arg_order;
    int32_t lmax;            // shadow slots per path: the largest eligible-light count (>= 1)
    int32_t integrator;      // C2_INTEGRATOR_PATH or C2_INTEGRATOR_DIRECT
    const View* views;
    const int2* pv;          // views of this pass: (view, first linear index), ascending
    int32_t npv;
    int32_t cache_bounce;            // radiance cache: -1 off
    int32_t train_ratio;
    int32_t train_cap;               // training rows per Adam step (the launch collects nsamp x this)
    int32_t train_paths;             // training path record capacity
    int32_t wide8;                   // 1: traversals use the 8-wide tree (View::root8)
};

// Radiance-cache query in the MLP kernels' packed layout (kNrcQ4 float4: position and
// roughness; azimuth and cos theta of the outgoing direction and of the facing normal;
// albedo and 1). Roughness: diffuse 1, specular 0, GGX sqrt(ax ay), Phong sqrt(2 / (K + 2)).
struct CacheQuery {
    float4 q[kNrcQ4];
};
__device__ __forceinline__ CacheQuery cache_query(const DevScene& S, const Inter& it, V3 d, const int32_t* texel) {
    float rough = 1.f;
    V3 alb{0.f, 0.f, 0.f};
    if (it.material >= 0) {
        const c2_material m = S.materials[it.material];
        alb = V3{m.color[0], m.color[1], m.color[2]};
        if (m.kd_texture > 0) alb = texel_value(S, *texel);
        if (m.type == C2_MAT_MIRROR) rough = 0.f;
        else if (m.type == C2_MAT_GLASS || m.type == C2_MAT_MICRO || m.type == C2_MAT_PLASTIC)
            rough = sqrtf(m.roughness[0] * m.roughness[1]);
        else if (m.type == C2_MAT_PHONG) rough = sqrtf(2.f / (m.phong_k + 2.f));
    }
    CacheQuery c;
    c.q[0] = make_float4(it.p.x, it.p.y, it.p.z, rough);
    c.q[1] = make_float4(atan2f(-d.z, -d.x) * 0.15915494f + 0.5f, -d.y * 0.5f + 0.5f,
                         atan2f(it.n.z, it.n.x) * 0.15915494f + 0.5f, it.n.y * 0.5f + 0.5f);
    c.q[2] = make_float4(alb.x, alb.y, alb.z, 1.f);
    return c;
}

__device__ __forceinline__ int32_t halton_index_dev(const View& P, int x, int y, int k) {
    // HaltonSampler::GetVectorIndexForSample, HaltonSampler.cpp:251-270
    int64_t off = 0;
    if (P.stride0 > 1) {
        uint64_t v = (uint64_t)(x % 128), r0 = 0;
        for (int i = 0; i < P.exp0; i++) { r0 = r0 * 2 + (v % 2); v /= 2; }
        off += (int64_t)(r0 * (uint64_t)(int64_t)P.scale1 * (uint64_t)(int64_t)P.mult0);
        v = (uint64_t)(y % 128);
        uint64_t r1 = 0;
        for (int i = 0; i < P.bexp1; i++) { r1 = r1 * 3 + (v % 3); v /= 3; }
        off += (int64_t)(r1 * (uint64_t)(int64_t)P.bscale0 * (uint64_t)(int64_t)P.mult1);
    }
    return (int32_t)(off + (int64_t)(k * P.stride0));
}

// ---------------------------------------------------------------------------- path steps
// The arithmetic of one path's steps, shared by the wavefront kernels (which keep a
// path's state in the path buffers between launches) and the persistent kernel (which
// keeps it in registers), so both form the same floats in the same order.

// Pixel of linear index lin of a view's window. Bands of 8 rows walked column by
// column, so the 64 paths of a wavefront start in one 8 x 8 tile (coherent primary and
// shadow rays). A bijection onto the window for any size; the last band holds the
// leftover rows.
__device__ __forceinline__ void window_pixel(const View& V, int lin, int& x, int& y) {
    const int winH = V.win_h;
    const int fb = winH >> 3;
    const int band = lin / (V.win_w * 8);
    const int hb = band < fb ? 8 : winH - fb * 8;
    const int idxb = lin - band * V.win_w * 8;
    x = V.win_x0 + idxb / hb;
    y = V.win_y0 + band * 8 + idxb % hb;
}

// Camera ray direction of Halton sample k at pixel (x, y), and the sample's Halton index:
// SamplerIntegrator.cpp:42-52 with PerspCamera::GenerateRay (PerspCamera.cpp:23-29).
__device__ __forceinline__ V3 camera_dir(const DevScene& S, const View& V, int x, int y, int k, int32_t& idx) {
    const HaltonDev hd{V.exp0, V.scale1};
    idx = halton_index_dev(V, x, y, k);
    float s0 = halton_value(S, hd, idx, 0), s1 = halton_value(S, hd, idx, 1);
    float px = (float)x + s0, py = (float)y + s1;
    px = px / V.width;
    py = py / V.height;
    px = px * 2 - 1;
    py = py * 2 - 1;
    px *= V.camW, py *= V.camH;
    return normalize(V.right * px + V.up * py + V.look * 1.0f);
}

// Radiance the hit's own light emits back along the ray (0 from its back side), added
// when the previous bounce was specular or transmissive, PathTracingIntegrator.cpp:54-57.
__device__ __forceinline__ V3 hit_emission(const DevScene& S, const Inter& it, V3 d) {
    if (it.light < 0) return V3{0, 0, 0};
    const DevLight& Ld = S.lights[it.light];
    V3 nact = it.front ? it.n : -it.n;
    return dot(nact, -d) < 0 ? V3{0, 0, 0} : V3{Ld.L[0], Ld.L[1], Ld.L[2]};
}

// A surface vertex (a hit with a material) after shade_surface: what the light loop
// needs, and the continuation.
struct SurfVertex {
    float uL0, uL1;     // UniformSampleAllLights' light-sample dimensions
    float pdfB;         // pdf of the BSDF-sampled MIS direction
    int fBnz;           // that direction carries f != 0 (the MIS ray may be needed)
    V3 wiB, misF;       // the MIS direction and f * cos along it
    int32_t dim;        // sampler dimension after the vertex
    bool cont, lp;      // the path continues; through a specular / transmissive sample
    V3 beta2, no, wiC;  // throughput after the continuation, and the next ray
};

// Vertex stage of a surface hit: the light-sample and MIS-direction dimensions
// (UniformSampleAllLights, :175-278: uL = Get2D(), uB = Get2D(), one Get1D per eligible
// non-delta light), the BSDF-sampled MIS direction (shared by all lights), and for the
// path tracer the continuation sample and Russian roulette (:79-102).
template <int MAXT = kMatAll>
__device__ __forceinline__ void shade_surface(const DevScene& S, const View& V, const Inter& it, V3 d, V3 beta,
                                              int32_t idx, int32_t dim, int bounce, int32_t texel, int integrator,
                                              int arg_order, SurfVertex& sv) {
    const HaltonDev hd{V.exp0, V.scale1};
    Bsdf bs = make_bsdf<MAXT>(S, it, (uint32_t)idx, (uint32_t)(dim + 2), texel);
    V3 wo = -d;
    float uL0 = halton_value(S, hd, idx, dim), uL1 = halton_value(S, hd, idx, dim + 1);
    float uB0 = halton_value(S, hd, idx, dim + 2), uB1 = halton_value(S, hd, idx, dim + 3);
    const bool direct = integrator == C2_INTEGRATOR_DIRECT;
    int own = it.light;
    // the path tracer skips the hit's own light (:181-184), direct lighting does not
    // (DirectLightingIntegrator.cpp:68-71)
    int ownElig = (!direct && own >= 0 && S.lights[own].eligible) ? 1 : 0;
    int nGet1D = V.nd_elig - ownElig;  // one Get1D per eligible non-delta light, :240
    V3 misF{0, 0, 0}, wiB{0, 0, 0};
    float pdfB = 0.f;
    int fBnz = 0;
    if (nGet1D > 0) {
        int typeB = 0, stB = kSampleOk;
        V3 wiFail{0, 0, 0};
        V3 fB = bsdf_sample(bs, it, wo, uB0, uB1, T_ALL & ~T_SPEC, wiB, pdfB, typeB, &stB, &wiFail);
        if (!is_zero(fB)) {
            fBnz = 1;
            misF = fB * fmax_ref(0.f, dot(it.n, wiB));
        } else if (direct && stB == kSampleZero && pdfB == 0.f) {
            // Direct lighting keeps going after a zero sample
            // (DirectLightingIntegrator.cpp:120-150): the BxDF's local-frame wi is traced
            // as it is, and f * Li * w / pdf is 0 / 0 when the ray reaches the light. With
            // a nonzero pdf the term is 0, and with an unwritten pdf the reference reads an
            // uninitialised value: both add nothing and are not traced.
            fBnz = 1;
            wiB = wiFail;
            misF = fB * fmax_ref(0.f, dot(it.n, wiB));
        }
    }
    sv.uL0 = uL0;
    sv.uL1 = uL1;
    sv.pdfB = pdfB;
    sv.fBnz = fBnz;
    sv.wiB = wiB;
    sv.misF = misF;
    dim += 4 + nGet1D;
    sv.cont = false;
    sv.lp = false;
    sv.beta2 = beta;
    sv.no = V3{0, 0, 0};
    sv.wiC = V3{0, 0, 0};
    if (!direct) {
        // the path tracer continues (the direct-lighting path ends here):
        // sampler->Get1D() and sampler->Get2D() as arguments of one call, :79
        float u2x, u2y;
        uint32_t sdim;
        if (arg_order == C2_ARGORDER_RTL) {
            sdim = (uint32_t)dim;
            u2x = halton_value(S, hd, idx, dim);
            u2y = halton_value(S, hd, idx, dim + 1);
        } else {
            sdim = (uint32_t)(dim + 1);
            u2x = halton_value(S, hd, idx, dim + 1);
            u2y = halton_value(S, hd, idx, dim + 2);
        }
        // (the Get1D value only selects the BxDF; every supported material has one)
        dim += 3;
        bs.rdim = sdim;
        V3 wiC{0, 0, 0};
        float pdfC = 0.f;
        int typeC = 0;
        V3 fC = bsdf_sample(bs, it, wo, u2x, u2y, T_ALL, wiC, pdfC, typeC);
        if (!(fabsf(pdfC) == 0.f || is_zero(fC))) {
            bool specular = (typeC & T_SPEC) != 0;
            float cosine = specular ? 1.0f : fmax_ref(0.f, (typeC & T_TRANS) ? dot(-it.n, wiC) : dot(it.n, wiC));
            V3 beta2 = beta * (fC * cosine / pdfC);
            bool cont = true;
            if (bounce > 3) {  // Russian roulette, :96-102
                float qq = fmax_ref(.05f, 1.f - length(beta2));
                float ur = halton_value(S, hd, idx, dim);
                dim++;
                if (ur < qq) cont = false;
                else beta2 = beta2 / (1 - qq);
            }
            sv.beta2 = beta2;
            if (cont) {
                sv.cont = true;
                sv.lp = specular || (typeC & T_TRANS);
                sv.no = spawn_origin(it.p, it.n, wiC);
                sv.wiC = wiC;
            }
        }
    }
    sv.dim = dim;
}

// The vertex as the light loop sees it (rebuilt from the vertex records by k_nee): the
// bitangent is n x dpdu.
__device__ __forceinline__ Inter nee_inter(V3 p, V3 n, V3 dpdu, V3 dir, int32_t entity, int32_t material,
                                           int32_t light, bool front) {
    Inter it{};
    it.p = p;
    it.n = n;
    it.dpdu = dpdu;
    it.dpdv = cross(n, dpdu);
    it.dir = dir;
    it.entity = entity;
    it.material = material;
    it.light = light;
    it.front = front;
    return it;
}

// Light sampling half of EsimateDirect for one light (:199-230). When the sample carries
// f * Li != 0, calls use(so, sd, C) with the shadow ray (origin, unnormalised direction to
// the sampled point) and the unshadowed contribution, and returns true. (A callback rather
// than out-parameters: the caller's stores stay inside the branch, which keeps k_nee at
// 168 instead of 177 VGPRs.)
template <class Use>
__device__ __forceinline__ bool light_term(const DevScene& S, const DevLight& Lt, bool delta, const Inter& it,
                                           const Bsdf& bs, V3 wo, float uL0, float uL1, Use&& use) {
    V3 lightPos;
    float pdfL;
    V3 Li = sample_li(S, Lt, it.p, it.n, uL0, uL1, lightPos, pdfL);
    V3 wi = normalize(lightPos - it.p);
    if (pdfL != 0.f && !is_zero(Li)) {
        V3 f = bsdf_f(bs, it, wo, wi) * fmax_ref(0.f, dot(it.n, wi));
        float pB = bsdf_pdf(bs, it, wo, wi, T_ALL & ~T_SPEC);
        if (!is_zero(f)) {
            V3 C = delta ? f * Li / pdfL : f * Li * power_heuristic(pdfL, pB) / pdfL;
            bool back = dot(lightPos - it.p, it.n) < 0;  // SpawnRayTo
            V3 so = it.p + (back ? -it.n : it.n) * 1e-4f;
            V3 sd = lightPos - so;
            use(so, sd, C);
            return true;
        }
    }
    return false;
}

// The shared MIS ray, Scene::Intersect(isec.SpawnRay(wi)) from the vertex (sp, sn): only
// the light owning the hit entity can receive f * Le * w / pdf_bsdf (:255-273). Returns
// that light (-1 none) and the term in Lb.
template <int ACC, class Lds>
__device__ __forceinline__ int mis_term(const DevScene& S, const View& V, V3 sp, V3 sn, V3 wi, V3 f, float pdfB,
                                        int own, int integrator, Lds& lds, V3& Lb) {
    V3 o = spawn_origin(sp, sn, wi);
    Hit h;
    int lh = -1;
    Lb = V3{0, 0, 0};
    bool found;
    found = accel_trace<ACC, false>(S, ACC == kAccBvh8 ? V.root8 : V.root, V.kd_lo, V.kd_hi, o, wi, 0.f, INFINITY, -1,
                                    h, lds);
    if (found) {
        int4 e = S.entities[h.entity];
        int li = e.w;
        if (li >= 0 && (integrator == C2_INTEGRATOR_DIRECT || li != own)) {
            const DevLight& Lt = S.lights[li];
            if (Lt.eligible && Lt.type == C2_LIGHT_AREA) {
                float pdfL = pdf_light(S, Lt, sp, sn, wi);
                if (pdfL != 0.f) {
                    float weight = power_heuristic(pdfB, pdfL);
                    // lightIsec.Le(-wi) with the hit's un-flipped normal
                    V3 hp, hn, hdu;
                    bool front;
                    if (e.x == C2_SHAPE_TRIANGLE) {
                        tri_frame(S, e.y, wi, h.u, h.v, hn, hdu, front);
                    } else {
                        analytic_frame(e.x, S.spheres + (size_t)e.y * 22, o, wi, h.t, hp, hn, hdu, front);
                    }
                    V3 nact = front ? hn : -hn;
                    V3 Li = dot(nact, -wi) < 0 ? V3{0, 0, 0} : V3{Lt.L[0], Lt.L[1], Lt.L[2]};
                    if (!is_zero(Li)) {
                        Lb = f * Li * weight / pdfB;
                        lh = li;
                    }
                }
            }
        }
    }
    return lh;
}

// FrameBuffer::AddSample, FrameBuffer.cpp:20-32 (weight 1); a NaN or Inf sample adds 0.
__device__ __forceinline__ void film_add(int mode, float4* film, float4* comp, double4* film64, int pixel, V3 c) {
    if (any_nan(c)) c = V3{0, 0, 0};
    if (any_inf(c)) c = V3{0, 0, 0};
    if (mode == C2_FILM_F64) {
        double4 f = film64[pixel];
        f.x += (double)c.x;
        f.y += (double)c.y;
        f.z += (double)c.z;
        f.w += 1.0;
        film64[pixel] = f;
    } else if (mode == C2_FILM_F32_KAHAN) {
        float4 f = film[pixel], cp = comp[pixel];
        float y, tt;
        y = c.x - cp.x; tt = f.x + y; cp.x = (tt - f.x) - y; f.x = tt;
        y = c.y - cp.y; tt = f.y + y; cp.y = (tt - f.y) - y; f.y = tt;
        y = c.z - cp.z; tt = f.z + y; cp.z = (tt - f.z) - y; f.z = tt;
        f.w += 1.f;
        film[pixel] = f;
        comp[pixel] = cp;
    } else {
        float4 f = film[pixel];
        V3 acc = V3{f.x, f.y, f.z} + c * 1.f;
        film[pixel] = make_float4(acc.x, acc.y, acc.z, f.w + 1.f);
    }
}

// Block-aggregated appends: one atomic per block of kBlock threads instead of one per
// wavefront, because every block of a launch appends to the same counter and the
// atomics on one address serialise in L2. Every thread of the block must make the
// call (block-uniform control flow); each call starts with a barrier, so one LDS
// scratch area can serve consecutive calls.
constexpr int kWaves = kBlock / 64;
constexpr int kAppendLights = 32;  // lights per light-major append group

// Predicate append: returns this thread's list position, or -1 when !pred.
// sm: kWaves + 1 ints of LDS.
__device__ __forceinline__ int32_t block_append(int32_t* counter, bool pred, int32_t* sm) {
    const int lane = __lane_id(), w = (int)threadIdx.x >> 6;
    const uint64_t m = __ballot(pred);
    __syncthreads();
    if (lane == 0) sm[w] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int k = 0; k < kWaves; k++) {
            int c = sm[k];
            sm[k] = tot;
            tot += c;
        }
        sm[kWaves] = tot ? atomicAdd(counter, tot) : 0;
    }
    __syncthreads();
    if (!pred) return -1;
    const uint64_t lower = lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
    return sm[kWaves] + sm[w] + __popcll(m & lower);
}

// Slot of a shadow ray in its wavefront's packed region: the region of the 64 vertices
// of queue positions i & ~63 starts at (i & ~63) * lmax; `off` is the wave's rays of the
// lights before this one, `m` the ballot of the lanes whose vertex emits this light's ray.
__device__ __forceinline__ int32_t wave_slot(int32_t i, int lmax, int32_t off, uint64_t m) {
    const int lane = __lane_id();
    const uint64_t lower = lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
    return (i & ~63) * lmax + off + __popcll(m & lower);
}

// Appends the shadow rays of a group of nb lights (bit j of `bits`: the group's light j
// emits one) to `list` in light-major order within the block (all of the block's rays of
// the group's first light, then the next light's, ...). The entries are the rays' slots
// (wave_slot, with `off` the wave's rays before the group). sm: kWaves * kAppendLights + 1
// ints of LDS.
__device__ __forceinline__ void block_append_lights(int32_t* counter, uint32_t bits, int nb, int32_t* sm,
                                                    int32_t* list, int32_t i, int lmax, int32_t off) {
    const int lane = __lane_id(), w = (int)threadIdx.x >> 6;
    __syncthreads();
    for (int j = 0; j < nb; j++) {
        uint64_t m = __ballot((bits >> j) & 1u);
        if (lane == 0) sm[w * kAppendLights + j] = __popcll(m);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int j = 0; j < nb; j++)
            for (int k = 0; k < kWaves; k++) {
                int c = sm[k * kAppendLights + j];
                sm[k * kAppendLights + j] = tot;
                tot += c;
            }
        sm[kWaves * kAppendLights] = tot ? atomicAdd(counter, tot) : 0;
    }
    __syncthreads();
    const uint64_t lower = lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
    const int32_t base = sm[kWaves * kAppendLights];
    for (int j = 0; j < nb; j++) {
        const uint64_t m = __ballot((bits >> j) & 1u);
        if ((bits >> j) & 1u) list[base + sm[w * kAppendLights + j] + __popcll(m & lower)] = wave_slot(i, lmax, off, m);
        off += __popcll(m);
    }
}

__global__ void __launch_bounds__(kBlock) k_generate(DevScene S, PathBuffers B, PassParams P) {
    __shared__ int32_t sm[kWaves + 1];
    int stride = gridDim.x * blockDim.x;
    const int npath = P.nsamp * P.pix_count;
    for (int base = blockIdx.x * blockDim.x; base < npath; base += stride) {
        int i = base + (int)threadIdx.x;
        bool live = i < npath;
        const int js = live ? i / P.pix_count : 0;
        int lin = P.pix_begin + (i - js * P.pix_count);
        // the view owning linear index lin: last entry with first index <= lin
        int lo = 0, hi = P.npv - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (P.pv[mid].y <= lin) lo = mid;
            else hi = mid - 1;
        }
        const int v = P.pv[lo].x;
        const View& V = P.views[v];
        lin -= P.pv[lo].y;
        int x, y;
        window_pixel(V, lin, x, y);
        int pixel = y * V.width + x;
        // training pixels: a phase per sample, so the samples of one launch train different pixels
        const int phase = P.cache_bounce >= 0
                              ? (int)(((uint32_t)(P.sample + js) * 2654435761u) % (uint32_t)P.train_ratio)
                              : 0;
        bool train = live && P.cache_bounce >= 0 && (pixel % P.train_ratio) == phase;
        int tp = P.cache_bounce >= 0 ? block_append(&B.counters[5], train, sm) : -1;  // uniform condition
        if (tp >= P.train_paths) train = false;  // beyond the record capacity: an ordinary path
        if (!live) continue;
        int32_t idx;
        const V3 d = camera_dir(S, V, x, y, P.sample + js, idx);
        B.ray_o[i] = make_float4(V.eye.x, V.eye.y, V.eye.z, 0.f);
        B.ray_d[i] = make_float4(d.x, d.y, d.z, 0.f);
        B.beta[i] = make_float4(1.f, 1.f, 1.f, 0.f);
        B.L[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        int fl = F_LIGHTPATH;
        if (train) {
            fl |= F_TRAIN | (tp << 8);
            B.tlist[tp] = i;
            for (int b = 0; b < P.max_depth; b++) B.trec_valid[(int64_t)tp * P.max_depth + b] = 0;
        }
        B.meta[i] = make_int4((int32_t)(V.film_off + pixel), idx, 2, fl);
        B.vid[i] = v;
        B.queue[0][i] = i;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) B.counters[0] = npath;
}

__global__ void k_chunk_reset(PathBuffers B) {
    B.counters[5] = 0;
    B.counters[7] = 0;
}


// Shading class of a hit: material type and its distribution (or, for glass, whether it
// is rough); the last two classes are "no material" and "no hit".
__device__ __forceinline__ int shade_class(const DevScene& S, int ent) {
    if (ent < 0) return kClasses - 1;
    const int m = S.entities[ent].z;
    if (m < 0) return kClasses - 2;
    const c2_material& M = S.materials[m];
    const int sub = M.type == C2_MAT_GLASS ? (M.roughness[0] != 0.f || M.roughness[1] != 0.f ? 1 : 0) : (M.ndf & 3);
    return min(M.type, 6) * 4 + sub;
}

// The same class on the host, for a material record.
inline int shade_class_host(const c2_material& M) {
    const int sub = M.type == C2_MAT_GLASS ? (M.roughness[0] != 0.f || M.roughness[1] != 0.f ? 1 : 0) : (M.ndf & 3);
    return std::min(M.type, 6) * 4 + sub;
}
// Grouping pays when the scene's surfaces span at least this many shading classes: the
// sphere field (diffuse, GGX, smooth and rough glass) renders 9-11% faster grouped at
// 960 x 540 and 1920 x 1080, the Cornell box (diffuse and mirror) 7-13% faster ungrouped
// at 256 x 256 to 1920 x 1080 (the two extra launches per bounce cost more than two
// classes diverge).
constexpr int kGroupMinClasses = 3;

// Groups a bounce's queue by shading class before the vertex kernels (k_shade, k_nee,
// k_resolve), so that a wavefront's paths mostly run one material's code instead of
// the code of every material its paths hit. Each path's result does not depend on the
// order. Counting pass: per-block histogram in LDS, one atomic per class and block.
__global__ void __launch_bounds__(kBlock) k_class_count(DevScene S, PathBuffers B, int q) {
    __shared__ int32_t h[kClasses];
    if (threadIdx.x < kClasses) h[threadIdx.x] = 0;
    __syncthreads();
    const int n = B.counters[q];
    const int32_t* queue = B.queue[q];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int c = shade_class(S, __float_as_int(B.hit[queue[i]].w));
        B.cls[i] = (uint8_t)c;
        atomicAdd(&h[c], 1);
    }
    __syncthreads();
    if (threadIdx.x < kClasses && h[threadIdx.x]) atomicAdd(&B.ccount[threadIdx.x], h[threadIdx.x]);
}

// Scatter pass: class c starts at the sum of the counts below it; a block reserves its
// slots of each class with one atomic per class. Within a block the order is kept (a
// lane's rank among the same-class lanes before it, from five ballots), so a class's
// paths stay in queue order, which is close to image order: the shadow and extension
// rays they spawn stay spatially coherent.
__global__ void __launch_bounds__(kBlock) k_class_scatter(PathBuffers B, int q) {
    static_assert(kClasses == 32, "five class bits");
    __shared__ int32_t base[kClasses], wcnt[kWaves][kClasses], off[kClasses];
    const int lane = __lane_id(), w = (int)threadIdx.x >> 6;
    if (threadIdx.x < kClasses) {
        int b = 0;
        for (int c = 0; c < (int)threadIdx.x; c++) b += B.ccount[c];
        base[threadIdx.x] = b;
    }
    const int n = B.counters[q];
    const int32_t* queue = B.queue[q];
    const uint64_t lower = lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
    for (int i0 = blockIdx.x * blockDim.x; i0 < n; i0 += gridDim.x * blockDim.x) {
        for (int k = (int)threadIdx.x; k < kWaves * kClasses; k += blockDim.x) (&wcnt[0][0])[k] = 0;
        __syncthreads();
        const int i = i0 + (int)threadIdx.x;
        const bool live = i < n;
        const int c = live ? B.cls[i] : 0;
        uint64_t same = __ballot(live);
        for (int bit = 0; bit < 5; bit++) {
            const uint64_t m = __ballot(live && ((c >> bit) & 1));
            same &= ((c >> bit) & 1) ? m : ~m;
        }
        const int rank = __popcll(same & lower);
        if (live && rank == 0) wcnt[w][c] = __popcll(same);
        __syncthreads();
        if (threadIdx.x < kClasses) {
            const int cc = (int)threadIdx.x;
            int t = 0;
            for (int k = 0; k < kWaves; k++) {
                const int v = wcnt[k][cc];
                wcnt[k][cc] = t;  // exclusive prefix over the block's waves
                t += v;
            }
            off[cc] = t ? atomicAdd(&B.ccount[kClasses + cc], t) : 0;
        }
        __syncthreads();
        if (live) B.queue_s[base[c] + off[c] + wcnt[w][c] + rank] = queue[i];
        __syncthreads();
    }
}

template <int ACC>
__global__ void __launch_bounds__(kBlock) k_extend(DevScene S, PathBuffers B, PassParams P, int q) {
    __shared__ TravLdsFor<ACC, false> lds;
    // The bounce's append counters (next queue, shadow rays, MIS rays, cache queries) and
    // the class histogram start at zero. Their readers in the previous bounce have finished:
    // they ran before this launch on the lane's stream, and shadow rays forked to the side
    // stream were joined before k_resolve. (A kernel of its own for this cost a launch per
    // bounce.)
    static_assert(2 * kClasses <= kBlock, "one thread per histogram entry");
    if (blockIdx.x == 0 && threadIdx.x < 2 * kClasses) {
        if (threadIdx.x == 0) {
            B.counters[4] = 0;
            B.counters[q ^ 1] = 0;
            B.counters[2] = 0;
            B.counters[3] = 0;
        }
        B.ccount[threadIdx.x] = 0;
    }
    int n = B.counters[q];
    const int32_t* queue = B.queue[q];
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&B.totals[0], (unsigned long long)n);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int p = queue[i];
        float4 o = B.ray_o[p], d = B.ray_d[p];
        const View& V = P.views[B.vid[p]];
        Hit h;
        bool hit;
        hit = accel_trace<ACC, false>(S, ACC == kAccBvh8 ? V.root8 : V.root, V.kd_lo, V.kd_hi, V3{o.x, o.y, o.z},
                                      V3{d.x, d.y, d.z}, 0.f, INFINITY, -1, h, lds);
        B.hit[p] = make_float4(h.t, h.u, h.v, __int_as_float(hit ? h.entity : -1));
    }
}

// Texture stage of scenes with images, between k_extend and k_shade (its own kernel so
// that the double-evaluated atan2/acos of sphere and disk uv cost the traversal and
// shading kernels no registers): the Kd texel of a textured hit, and for a ray that
// left the scene L += beta * skybox(dir) (PathTracingIntegrator.cpp:39-46; the
// direct-lighting integrator returns it unweighted, DirectLightingIntegrator.cpp:34-38,
// and beta is 1 there).
__global__ void __launch_bounds__(kBlock) k_surface(DevScene S, PathBuffers B, PassParams P, int q) {
    int n = B.counters[q];
    const int32_t* queue = B.queue[q];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int p = queue[i];
        float4 hh = B.hit[p];
        int ent = __float_as_int(hh.w);
        float4 rd = B.ray_d[p];
        V3 d{rd.x, rd.y, rd.z};
        if (ent < 0) {
            const View& V = P.views[B.vid[p]];
            if (V.sky[0] > 0) {
                float4 b4 = B.beta[p];
                V3 Ls = V3{b4.x, b4.y, b4.z} * sky_lookup(S, V.sky, d);
                float4 L = B.L[p];
                B.L[p] = make_float4(L.x + Ls.x, L.y + Ls.y, L.z + Ls.z, L.w);
            }
            continue;
        }
        const int4 e = S.entities[ent];
        if (e.z < 0) continue;
        const int kd = S.materials[e.z].kd_texture;
        if (kd <= 0) continue;
        float4 ro = B.ray_o[p];
        float u, v;
        surface_uv(S, e, V3{ro.x, ro.y, ro.z}, d, Hit{ent, hh.x, hh.y, hh.z}, u, v);
        B.texel[i] = texel_index(S, kd, u, v);
    }
}

// Vertex stage: emission, BSDF, the BSDF-sampled MIS direction (shared by all
// lights), continuation sample and Russian roulette; the light loop runs in k_nee.
template <int MAXT>
__global__ void __launch_bounds__(kBlock) k_shade(DevScene S, PathBuffers B, PassParams P, int q, int bounce) {
    __shared__ int32_t sm[kWaves + 1];
    int n = B.counters[q];
    const int32_t* queue = B.queue[q];
    int32_t* next = B.queue[q ^ 1];
    // grid-stride loop over block-uniform bases: every lane of a wavefront runs the
    // same iterations, as block_append's barriers and ballots require
    int stride = gridDim.x * blockDim.x;
    for (int base = blockIdx.x * blockDim.x; base < n; base += stride) {
        int i = base + (int)threadIdx.x;
        bool live = i < n;
        bool cont = false;
        int p = live ? queue[i] : 0;
        // a path that ends in the radiance cache here gets its query slot first, so that its
        // query is written where it is computed
        int cpos = -1;
        if (P.cache_bounce == bounce) {  // uniform over the launch
            bool query = false;
            if (live) {
                const int ent = __float_as_int(B.hit[p].w);
                query = ent >= 0 && S.entities[ent].z >= 0 && (B.meta[p].w & F_TRAIN) == 0;
            }
            cpos = block_append(&B.counters[4], query, sm);
        }
        if (live) {
            int4 meta = B.meta[p];
            const View& V = P.views[B.vid[p]];
            float4 hh = B.hit[p];
            int ent = __float_as_int(hh.w);
            int flags = 0;
            int32_t vent = -1, vmat = -1, vown = -1, vfl = 0;  // vertex record
            int32_t idx = meta.y, dim = meta.z;
            float4 b4 = B.beta[p];
            V3 beta{b4.x, b4.y, b4.z};
            V3 beta2 = beta;
            if (ent >= 0) {
                float4 ro = B.ray_o[p], rd = B.ray_d[p];
                V3 o{ro.x, ro.y, ro.z}, d{rd.x, rd.y, rd.z};
                Hit h{ent, hh.x, hh.y, hh.z};
                Inter it = make_inter(S, o, d, h);
                // emission when the previous bounce was specular / transmissive, :54-57
                float4 Lpre = B.L[p];
                if (meta.w & F_LIGHTPATH) {
                    V3 L = V3{Lpre.x, Lpre.y, Lpre.z} + beta * hit_emission(S, it, d);
                    Lpre = make_float4(L.x, L.y, L.z, 0.f);
                    B.L[p] = Lpre;
                }
                const bool training = (meta.w & F_TRAIN) != 0;
                if (cpos >= 0) {
                    // the path ends in the radiance cache: L += beta * cache(x, wo) after this kernel
                    // (cpos >= 0 exactly when it.material >= 0, the path is not a training path
                    // and this is the cache bounce)
                    const CacheQuery cq = cache_query(S, it, d, B.texel + i);
                    float4* rec = B.nrc_feat + (int64_t)cpos * kNrcRec4;
                    for (int k = 0; k < kNrcQ4; k++) rec[k] = cq.q[k];
                    rec[kNrcQ4] = make_float4(beta.x, beta.y, beta.z, 0.f);
                    dim = cpos;  // the Halton state is dead: k_accumulate finds the prediction here
                    flags = F_QUERY;
                } else if (it.material < 0) {
                    // transparent: SpawnRay(dir), no dimensions consumed, :63-68. The
                    // direct-lighting integrator ends the path at a transparent emitter
                    // (DirectLightingIntegrator.cpp:46-53).
                    if (P.integrator == C2_INTEGRATOR_PATH || it.light < 0) {
                        V3 no = spawn_origin(it.p, it.n, d);
                        B.ray_o[p] = make_float4(no.x, no.y, no.z, 0.f);
                        cont = true;
                    }
                    flags = meta.w & F_LIGHTPATH;
                } else {
                    if (training) {  // training row: features, throughput and radiance so far
                        int64_t r = ((int64_t)(meta.w >> 8) * P.max_depth + bounce);
                        float4* rec = B.trec + r * (kNrcQ4 + 2);
                        const CacheQuery tq = cache_query(S, it, d, B.texel + i);
                        for (int k = 0; k < kNrcQ4; k++) rec[k] = tq.q[k];
                        rec[kNrcQ4] = make_float4(beta.x, beta.y, beta.z, 0.f);
                        rec[kNrcQ4 + 1] = Lpre;
                        B.trec_valid[r] = 1;
                    }
                    // the Kd texel k_surface found, read only for a textured material
                    const int32_t texel = S.materials[it.material].kd_texture > 0 ? B.texel[i] : 0;
                    SurfVertex sv;
                    shade_surface<MAXT>(S, V, it, d, beta, idx, dim, bounce, texel, P.integrator, P.arg_order, sv);
                    B.vtx_p[i] = make_float4(it.p.x, it.p.y, it.p.z, sv.uL0);
                    B.vtx_n[i] = make_float4(it.n.x, it.n.y, it.n.z, sv.uL1);
                    B.vtx_du[i] = make_float4(it.dpdu.x, it.dpdu.y, it.dpdu.z, __int_as_float(sv.fBnz));
                    B.vtx_d[i] = make_float4(d.x, d.y, d.z, sv.pdfB);
                    B.vtx_h[i] = make_int2(idx, B.vid[p]);
                    B.mis_w[i] = make_float4(sv.wiB.x, sv.wiB.y, sv.wiB.z, 0.f);
                    B.mis_f[i] = make_float4(sv.misF.x, sv.misF.y, sv.misF.z, 0.f);
                    vent = ent;
                    vmat = it.material;
                    vown = it.light;
                    vfl = (it.front ? V_FRONT : 0) | V_NEE;
                    dim = sv.dim;
                    flags = F_NEE;
                    beta2 = sv.beta2;
                    if (sv.cont) {
                        cont = true;
                        B.ray_o[p] = make_float4(sv.no.x, sv.no.y, sv.no.z, 0.f);
                        B.ray_d[p] = make_float4(sv.wiC.x, sv.wiC.y, sv.wiC.z, 0.f);
                        if (sv.lp) flags |= F_LIGHTPATH;
                    }
                }
            }
            if (cont) flags |= F_CONTINUE;
            flags |= meta.w & ~0xF;  // keep the training tag
            B.meta[p] = make_int4(meta.x, idx, dim, flags);
            // the vertex record of every queued vertex (k_nee and k_resolve read its flags)
            if (cont) {
                vfl |= V_CONTINUE;
                B.beta2[i] = make_float4(beta2.x, beta2.y, beta2.z, 0.f);
            }
            B.vtx_i[i] = make_int4(vent, vmat, vown, vfl);
        }
        int qpos = block_append(&B.counters[q ^ 1], live && cont, sm);
        if (qpos >= 0) next[qpos] = p;
    }
}

// Training rows of this launch: target = (L_final - L after the vertex's emission) / beta.
// A thread takes kCollectPer (training path, bounce) records, kBlock apart, so every load is
// coalesced over the block's records; the block appends its valid rows with one atomic on the
// row counter (a wave scan of the per-thread counts, then the waves' sums in LDS), and each
// thread writes its rows in record order from its offset. One atomic per wavefront instead
// serialised ~21 k atomics on the counter per launch.
constexpr int kCollectPer = 8;
__global__ void __launch_bounds__(kBlock) k_cache_collect(PathBuffers B, PassParams P) {
    __shared__ int32_t wsum[kBlock / 64];
    __shared__ int32_t bbase;
    const int64_t nr = (int64_t)min(B.counters[5], P.train_paths) * P.max_depth;
    const int cap = P.train_cap * P.nsamp;
    const int lane = __lane_id(), wv = (int)threadIdx.x >> 6;
    const int64_t span = (int64_t)kBlock * kCollectPer;
    for (int64_t r0 = (int64_t)blockIdx.x * span; r0 < nr; r0 += (int64_t)gridDim.x * span) {  // block-uniform
        float4 tgt[kCollectPer];
        uint32_t okm = 0;
#pragma unroll
        for (int k = 0; k < kCollectPer; k++) {
            const int64_t r = r0 + (int64_t)k * kBlock + threadIdx.x;
            tgt[k] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (r < nr && B.trec_valid[r]) {
                const float4* rec = B.trec + r * (kNrcQ4 + 2);
                const float4 Lf = B.L[B.tlist[r / P.max_depth]];
                const float4 be = rec[kNrcQ4], lp = rec[kNrcQ4 + 1];
                const float t0v = be.x > 0.f ? (Lf.x - lp.x) / be.x : 0.f;
                const float t1v = be.y > 0.f ? (Lf.y - lp.y) / be.y : 0.f;
                const float t2v = be.z > 0.f ? (Lf.z - lp.z) / be.z : 0.f;
                if (isfinite(t0v) && isfinite(t1v) && isfinite(t2v)) {
                    okm |= 1u << k;
                    tgt[k] = make_float4(fmaxf(t0v, 0.f), fmaxf(t1v, 0.f), fmaxf(t2v, 0.f), 0.f);
                }
            }
        }
        const int cnt = __popc(okm);
        int incl = cnt;  // inclusive scan over the wavefront
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int v = __shfl_up(incl, d);
            if (lane >= d) incl += v;
        }
        if (lane == 63) wsum[wv] = incl;
        __syncthreads();
        if (threadIdx.x == 0) {
            int tot = 0;
            for (int w = 0; w < kBlock / 64; w++) {
                const int t = wsum[w];
                wsum[w] = tot;
                tot += t;
            }
            bbase = tot ? atomicAdd(&B.counters[7], tot) : 0;
        }
        __syncthreads();
        int slot = bbase + wsum[wv] + incl - cnt;
#pragma unroll
        for (int k = 0; k < kCollectPer; k++) {
            if (!((okm >> k) & 1u)) continue;
            if (slot < cap) {
                const float4* rec = B.trec + (r0 + (int64_t)k * kBlock + threadIdx.x) * (kNrcQ4 + 2);
                for (int j = 0; j < kNrcQ4; j++) B.bfeat[(int64_t)slot * kNrcQ4 + j] = rec[j];
                B.btarget[slot] = tgt[k];
            }
            slot++;
        }
        __syncthreads();  // wsum and bbase are rewritten by the next iteration
    }
}

// Light loop of UniformSampleAllLights / EsimateDirect (:181-278), one thread per
// vertex looping over the eligible lights in light order: the vertex record is loaded
// and the BSDF built once per light, then per light j the light sample (a shadow ray in
// the vertex's next free slot) and the test whether the shared BSDF-sampled direction
// needs the MIS ray (Pdf_Li != 0).
// (Four waves per SIMD asked for: 128 VGPRs and 168 B of scratch per lane instead of 165
// VGPRs at three waves. The kernel waits on memory more than it issues, so the fourth wave
// pays for the spills: +1.5-1.9% on the 1080p sphere field; five waves (96 VGPRs, 292 B)
// lose 2%.)
template <int MAXT>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_nee(DevScene S, PathBuffers B, PassParams P, int q) {
    __shared__ int32_t sm[kWaves * kAppendLights + 1];
    int n = B.counters[q];
    const int E = P.lmax;
    int stride = gridDim.x * blockDim.x;
    for (int base = blockIdx.x * blockDim.x; base < n; base += stride) {
        int i = base + (int)threadIdx.x;
        bool live = i < n;
        int4 vi = live ? B.vtx_i[i] : make_int4(-1, -1, -1, 0);
        const bool nee = (vi.w & V_NEE) != 0;
        int eb = 0, ne = 0, root = -1;
        Inter it{};
        int32_t texel = 0;
        uint32_t hidx = 0;  // Halton index: the BSDF's Phong redraw stream
        V3 wo{0, 0, 0}, misDir{0, 0, 0};
        float uL0 = 0.f, uL1 = 0.f;
        bool misCand = false;
        if (nee) {
            const int2 vh = B.vtx_h[i];
            const View& V = P.views[vh.y];
            eb = V.elig_begin;
            ne = V.num_elig;
            root = P.wide8 ? V.root8 : V.root;
            float4 vp = B.vtx_p[i], vn = B.vtx_n[i], vd = B.vtx_du[i], vdir = B.vtx_d[i];
            it = nee_inter(V3{vp.x, vp.y, vp.z}, V3{vn.x, vn.y, vn.z}, V3{vd.x, vd.y, vd.z},
                           V3{vdir.x, vdir.y, vdir.z}, vi.x, vi.y, vi.z, (vi.w & V_FRONT) != 0);
            // this bounce's texel (k_surface wrote it for this queue position)
            if (it.material >= 0 && S.materials[it.material].kd_texture > 0) texel = B.texel[i];
            hidx = (uint32_t)vh.x;
            wo = -it.dir;
            uL0 = vp.w;
            uL1 = vn.w;
            misCand = __float_as_int(vd.w) != 0;
            if (misCand) {
                float4 mw = B.mis_w[i];
                misDir = V3{mw.x, mw.y, mw.z};
            }
        }
        bool misHit = false;
        uint32_t shbits = 0;
        int32_t woff = 0, goff = 0;  // the wave's shadow rays so far; before this group of lights
        // block-uniform trip count (the appends below need every thread); lanes whose
        // view has fewer eligible lights idle through the tail
        for (int j = 0; j < E; j++) {
            const bool lj = j < ne;
            const int li = lj ? B.elig[eb + j] : -1;
            const bool delta = lj && S.lights[li].type == C2_LIGHT_POINT;
            bool shadow = false;
            float4 ro{}, rd{}, rc{};
            if (nee && lj && (P.integrator == C2_INTEGRATOR_DIRECT || it.light != li)) {
                const DevLight Lt = S.lights[li];
                // rebuilt per light from the (cached) material record rather than kept
                // live across the loop: fewer registers, more waves per SIMD
                const Bsdf bs = make_bsdf<MAXT>(S, it, hidx, 0u, texel);
                // (i) light sampling, :199-230
                light_term(S, Lt, delta, it, bs, wo, uL0, uL1, [&](V3 so, V3 sd, V3 C) {
                    ro = make_float4(so.x, so.y, so.z, __int_as_float(delta ? -1 : Lt.entity));
                    rd = make_float4(sd.x, sd.y, sd.z, __int_as_float(root));
                    rc = make_float4(C.x, C.y, C.z, 0.f);
                    shadow = true;
                });
                // (ii) BSDF sampling needs the MIS ray when Pdf_Li != 0, :235-257
                if (!delta && misCand && !misHit && pdf_light(S, Lt, it.p, it.n, misDir) != 0.f) misHit = true;
            }
            // the ray goes to the wave's next free slots, in lane order
            const uint64_t mj = __ballot(shadow);
            if (shadow) {
                const int32_t slot = wave_slot(i, E, woff, mj);
                B.sh_o[slot] = ro;
                B.sh_d[slot] = rd;
                B.sh_c[slot] = rc;
            }
            woff += __popcll(mj);
            shbits |= shadow ? (1u << (j % kAppendLights)) : 0u;
            if (j % kAppendLights == kAppendLights - 1 || j == E - 1) {
                block_append_lights(&B.counters[2], shbits, j % kAppendLights + 1, sm, B.sh_list, i, E, goff);
                if (live) B.sh_mask[(j / kAppendLights) * B.cap + i] = shbits;
                goff = woff;
                shbits = 0;
            }
        }
        if (live) B.nee_out[i] = misHit ? 1 : 0;
        int mpos = block_append(&B.counters[3], misHit, sm);
        if (mpos >= 0) B.mis_list[mpos] = i;
    }
}

// The shadow rays of blocks blk of nblk (grid-stride).
template <int ACC, class Lds>
__device__ __forceinline__ void occlude_blocks(const DevScene& S, const PathBuffers& B, Lds& lds, int blk, int nblk) {
    int n = B.counters[2];
    if (blk == 0 && threadIdx.x == 0) atomicAdd(&B.totals[1], (unsigned long long)n);
    for (int i = blk * blockDim.x + threadIdx.x; i < n; i += nblk * blockDim.x) {
        int s = B.sh_list[i];
        float4 o = B.sh_o[s], d = B.sh_d[s];
        Hit h;
        bool occluded;
        occluded = accel_trace<ACC, true>(S, __float_as_int(d.w), nullptr, nullptr, V3{o.x, o.y, o.z},
                                          V3{d.x, d.y, d.z}, 0.f, 1 - 1e-5f, __float_as_int(o.w), h, lds);
        B.sh_vis[s] = occluded ? 0 : 1;
    }
}

template <int ACC>
__global__ void __launch_bounds__(kBlock) k_occlude(DevScene S, PathBuffers B) {
    __shared__ TravLdsFor<ACC, true> lds;
    occlude_blocks<ACC>(S, B, lds, (int)blockIdx.x, (int)gridDim.x);
}

// The shared MIS ray: Scene::Intersect(isec.SpawnRay(wi)); only the light owning the
// hit entity can receive f * Le * w / pdf_bsdf, :255-273. Blocks blk of nblk (grid-stride).
template <int ACC, class Lds>
__device__ __forceinline__ void mis_blocks(const DevScene& S, const PathBuffers& B, const PassParams& P, Lds& lds,
                                           int blk, int nblk) {
    int n = B.counters[3];
    if (blk == 0 && threadIdx.x == 0) atomicAdd(&B.totals[2], (unsigned long long)n);
    for (int i = blk * blockDim.x + threadIdx.x; i < n; i += nblk * blockDim.x) {
        const int v = B.mis_list[i];  // the vertex's queue position
        float4 mp = B.vtx_p[v], mn = B.vtx_n[v], mw = B.mis_w[v], mf = B.mis_f[v];
        V3 sp{mp.x, mp.y, mp.z}, sn{mn.x, mn.y, mn.z}, wi{mw.x, mw.y, mw.z}, f{mf.x, mf.y, mf.z};
        float pdfB = B.vtx_d[v].w;
        int own = B.vtx_i[v].z;
        const View& V = P.views[B.vtx_h[v].y];
        V3 Lb;
        const int lh = mis_term<ACC>(S, V, sp, sn, wi, f, pdfB, own, P.integrator, lds, Lb);
        B.mis_L[v] = make_float4(Lb.x, Lb.y, Lb.z, 0.f);
        B.mis_hit[v] = lh;
    }
}

template <int ACC>
__global__ void __launch_bounds__(kBlock) k_mis(DevScene S, PathBuffers B, PassParams P) {
    __shared__ TravLdsFor<ACC, false> lds;
    mis_blocks<ACC>(S, B, P, lds, (int)blockIdx.x, (int)gridDim.x);
}

// Both in one launch, for chunks that trace them in line: blocks [0, gs) take the shadow
// rays, the rest the MIS rays (one launch less per bounce; its registers are those of the
// MIS rays, so large chunks keep the two kernels).
template <int ACC>
__global__ void __launch_bounds__(kBlock) k_occlude_mis(DevScene S, PathBuffers B, PassParams P, int gs) {
    __shared__ union {
        TravLdsFor<ACC, true> any;
        TravLdsFor<ACC, false> closest;
    } lds;
    if ((int)blockIdx.x < gs) occlude_blocks<ACC>(S, B, lds.any, (int)blockIdx.x, gs);
    else mis_blocks<ACC>(S, B, P, lds.closest, (int)blockIdx.x - gs, (int)gridDim.x - gs);
}

// L += beta * sum over eligible lights in light order of EsimateDirect's result:
// the unoccluded light sample (0 + C) plus, for the light owning the MIS hit, Lb. The
// slots of the wave's shadow rays are found again by replaying k_nee's ballots over the
// vertices' light masks (the grid-stride loop over block-uniform bases keeps vertex i in
// lane i % 64, as in k_nee).
__global__ void __launch_bounds__(kBlock) k_resolve(PathBuffers B, PassParams P, int q) {
    int n = B.counters[q];
    const int32_t* queue = B.queue[q];
    const int E = P.lmax;
    const int stride = gridDim.x * blockDim.x;
    for (int base = blockIdx.x * blockDim.x; base < n; base += stride) {  // wave-uniform trip count
        const int i = base + (int)threadIdx.x;
        const bool live = i < n;
        const int vfl = live ? B.vtx_i[i].w : 0;
        const bool nee = (vfl & V_NEE) != 0;
        int p = 0, lh = -1, eb = 0, ne = 0;
        V3 Lb{0, 0, 0};
        if (nee) {
            p = queue[i];
            if (B.nee_out[i]) lh = B.mis_hit[i];
            if (lh >= 0) {
                float4 m = B.mis_L[i];
                Lb = V3{m.x, m.y, m.z};
            }
            const View& V = P.views[B.vtx_h[i].y];
            eb = V.elig_begin;
            ne = V.num_elig;
        }
        V3 U{0, 0, 0};
        int32_t woff = 0;
        uint32_t bits = 0;
        for (int j = 0; j < E; j++) {
            if (j % kAppendLights == 0) bits = live ? B.sh_mask[(j / kAppendLights) * B.cap + i] : 0u;
            const bool ray = (bits >> (j % kAppendLights)) & 1u;
            const uint64_t mj = __ballot(ray);
            if (nee && j < ne) {
                const int li = B.elig[eb + j];
                V3 E3{0, 0, 0};
                bool any = false;
                if (ray) {
                    const int32_t slot = wave_slot(i, E, woff, mj);
                    if (B.sh_vis[slot] == 1) {
                        const float4 c = B.sh_c[slot];
                        E3 = E3 + V3{c.x, c.y, c.z};
                        any = true;
                    }
                }
                if (li == lh) {
                    E3 = E3 + Lb;
                    any = true;
                }
                if (any) U = U + E3;
            }
            woff += __popcll(mj);
        }
        if (nee) {
            float4 b4 = B.beta[p], L4 = B.L[p];
            V3 L = V3{L4.x, L4.y, L4.z} + V3{b4.x, b4.y, b4.z} * U;
            B.L[p] = make_float4(L.x, L.y, L.z, 0.f);
        }
        if (vfl & V_CONTINUE) B.beta[queue[i]] = B.beta2[i];  // (a flagged vertex is live)
    }
}

// FrameBuffer::AddSample, FrameBuffer.cpp:20-32 (weight 1, in sample order): the film
// adds of a chunk's samples of one pixel run in one thread in sample order, so
// every pixel's sum is formed in the reference's order (sample 0, 1, ...).
__global__ void __launch_bounds__(kBlock) k_accumulate(PathBuffers B, PassParams P, int mode, float4* film,
                                                       float4* comp, double4* film64) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < P.pix_count; t += gridDim.x * blockDim.x) {
        for (int js = 0; js < P.nsamp; js++) {
            const int i = js * P.pix_count + t;
            float4 L4 = B.L[i];
            const int4 m = B.meta[i];
            if (m.w & F_QUERY) {  // L + beta * cache prediction (k_nrc_infer wrote the product)
                const float4 c = B.nrc_feat[(int64_t)m.z * kNrcRec4 + kNrcQ4];
                L4 = make_float4(L4.x + c.x, L4.y + c.y, L4.z + c.z, L4.w);
            }
            film_add(mode, film, comp, film64, m.x, V3{L4.x, L4.y, L4.z});
        }
    }
}

// ============================================================================ persistent kernel

// One persistent launch: every pixel of the listed views is one work item, handed out in
// list order from a device counter; a lane that takes a pixel traces samples
// [sample_begin, min(spp, sample_end)) of it one after another, each path from the camera
// ray to its end in registers, and adds each sample to the film itself, so a pixel's
// samples are added in sample order, as k_accumulate adds them.
struct PersistParams {
    const View* views;
    const int2* pv;        // (view, first item), ascending by first item
    int32_t npv;
    int32_t nitems;        // pixels of the listed views' windows
    int32_t* counter;      // next unclaimed item; 0 when the launch starts
    int32_t sample_begin, sample_end;
    int32_t steps;         // path steps (direct lighting: the transparent chain limit)
    int32_t arg_order, integrator, film_mode;
    float4* film;
    float4* comp;
    double4* film64;
    const int32_t* elig;
    unsigned long long* totals;  // extension, shadow, MIS rays
};

// The closest-hit and any-hit stacks of a lane are never live at once, so they share
// the block's LDS (a lane's entries lie at offsets = lane mod BS in both).
template <int BS>
union PersistLds {
    TravLds<false, BS> closest;
    TravLds<true, BS> any;
};

// BS threads per block; W > 0 limits the registers so that at least W waves fit on a SIMD
// (the rest spill to scratch), W = 0 gives the kernel the registers it needs.
template <int BS, int W>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(W > 0 ? W : 1)))
k_persist(DevScene S, PersistParams Q) {
    __shared__ PersistLds<BS> lds;
    const int lane = __lane_id();
    const uint64_t lower = lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
    const bool direct = Q.integrator == C2_INTEGRATOR_DIRECT;
    // the lane's pixel
    int v = 0, x = 0, y = 0, pixel = 0, k = 0, kend = 0;
    bool done = false, live = false;
    // the lane's path
    V3 o{0, 0, 0}, d{0, 0, 0}, beta{1, 1, 1}, L{0, 0, 0};
    int32_t idx = 0, dim = 0, b = 0;
    bool lightpath = false;
    unsigned long long next = 0, nsh = 0, nmis = 0;
    while (true) {
        // lanes whose pixel has no sample left take the next items, one atomic per wave
        const bool want = !done && !live && k >= kend;
        const uint64_t m = __ballot(want);
        if (m) {
            const int first = __ffsll((unsigned long long)m) - 1;
            int base = 0;
            if (lane == first) base = atomicAdd(Q.counter, __popcll(m));
            base = __shfl(base, first);
            if (want) {
                const int item = base + __popcll(m & lower);
                if (item >= Q.nitems) {
                    done = true;
                } else {
                    int lo = 0, hi = Q.npv - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (Q.pv[mid].y <= item) lo = mid;
                        else hi = mid - 1;
                    }
                    v = Q.pv[lo].x;
                    const View& V = Q.views[v];
                    window_pixel(V, item - Q.pv[lo].y, x, y);
                    pixel = (int)V.film_off + y * V.width + x;
                    k = Q.sample_begin;
                    kend = min(V.spp, Q.sample_end);
                }
            }
        }
        if (!__ballot(!done)) break;  // the queue is drained and every path has ended
        if (!done && !live && k < kend) {  // next sample of the pixel: the camera ray
            const View& V = Q.views[v];
            d = camera_dir(S, V, x, y, k, idx);
            o = V.eye;
            beta = V3{1.f, 1.f, 1.f};
            L = V3{0.f, 0.f, 0.f};
            dim = 2;
            b = 0;
            lightpath = true;
            live = true;
        }
        if (!live) continue;
        const View& V = Q.views[v];
        bool cont = false;
        if (b < Q.steps) {
            // extension ray (k_extend)
            Hit h;
            const bool hit = traverse<false, BS>(S, V.root, o, d, 0.f, INFINITY, -1, h, lds.closest);
            next++;
            const int ent = hit ? h.entity : -1;
            h.entity = ent;
            if (ent < 0) {
                // a ray that left the scene: the skybox (k_surface)
                if (S.tex && V.sky[0] > 0) L = L + beta * sky_lookup(S, V.sky, d);
            } else {
                const int4 e = S.entities[ent];
                int32_t texel = 0;  // Kd texel of a textured material (k_surface)
                if (S.tex && e.z >= 0) {
                    const int kd = S.materials[e.z].kd_texture;
                    if (kd > 0) {
                        float u, vv;
                        surface_uv(S, e, o, d, h, u, vv);
                        texel = texel_index(S, kd, u, vv);
                    }
                }
                // vertex (k_shade)
                Inter it = make_inter(S, o, d, h);
                if (lightpath) L = L + beta * hit_emission(S, it, d);
                if (it.material < 0) {
                    // transparent: SpawnRay(dir), no dimensions consumed
                    if (!direct || it.light < 0) {
                        o = spawn_origin(it.p, it.n, d);
                        cont = true;
                    }
                } else {
                    SurfVertex sv;
                    shade_surface(S, V, it, d, beta, idx, dim, b, texel, Q.integrator, Q.arg_order, sv);
                    dim = sv.dim;
                    // light loop (k_nee, k_mis, k_occlude, k_resolve)
                    const Inter itn = nee_inter(it.p, it.n, it.dpdu, d, ent, it.material, it.light, it.front);
                    const V3 wo = -itn.dir;
                    const int ne = V.num_elig, eb = V.elig_begin;
                    bool misHit = false;  // some light has Pdf_Li != 0 along the MIS direction
                    if (sv.fBnz)
                        for (int j = 0; j < ne && !misHit; j++) {
                            const int li = Q.elig[eb + j];
                            const DevLight Lt = S.lights[li];
                            if ((direct || itn.light != li) && Lt.type != C2_LIGHT_POINT &&
                                pdf_light(S, Lt, itn.p, itn.n, sv.wiB) != 0.f)
                                misHit = true;
                        }
                    int lh = -1;
                    V3 Lb{0, 0, 0};
                    if (misHit) {
                        lh = mis_term<kAccBvh4>(S, V, itn.p, itn.n, sv.wiB, sv.misF, sv.pdfB, itn.light, Q.integrator,
                                             lds.closest, Lb);
                        nmis++;
                    }
                    // sum over the eligible lights in light order (k_resolve's order)
                    V3 U{0, 0, 0};
                    for (int j = 0; j < ne; j++) {
                        const int li = Q.elig[eb + j];
                        V3 E{0, 0, 0};
                        bool any = false;
                        if (direct || itn.light != li) {
                            const DevLight Lt = S.lights[li];
                            const bool delta = Lt.type == C2_LIGHT_POINT;
                            const Bsdf bs = make_bsdf(S, itn, (uint32_t)idx, 0u, texel);
                            light_term(S, Lt, delta, itn, bs, wo, sv.uL0, sv.uL1, [&](V3 so, V3 sd, V3 C) {
                                nsh++;
                                Hit hs;
                                if (!traverse<true, BS>(S, V.root, so, sd, 0.f, 1 - 1e-5f, delta ? -1 : Lt.entity,
                                                        hs, lds.any)) {
                                    E = E + C;
                                    any = true;
                                }
                            });
                        }
                        if (li == lh) {
                            E = E + Lb;
                            any = true;
                        }
                        if (any) U = U + E;
                    }
                    L = L + beta * U;
                    lightpath = false;
                    if (sv.cont) {
                        cont = true;
                        o = sv.no;
                        d = sv.wiC;
                        beta = sv.beta2;
                        lightpath = sv.lp;
                    }
                }
            }
        }
        b++;
        if (!cont || b >= Q.steps) {  // the path has ended: FrameBuffer::AddSample
            film_add(Q.film_mode, Q.film, Q.comp, Q.film64, pixel, L);
            k++;
            live = false;
        }
    }
    if (next) atomicAdd(&Q.totals[0], next);
    if (nsh) atomicAdd(&Q.totals[1], nsh);
    if (nmis) atomicAdd(&Q.totals[2], nmis);
}

__global__ void k_set_i32(int32_t* p, int32_t v) { *p = v; }

__global__ void k_zero4(float4* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = make_float4(0, 0, 0, 0);
}
__global__ void k_zero_d4(double4* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = make_double4(0, 0, 0, 0);
}

// ============================================================================ host side

struct DeviceArena {
    std::vector<void*> ptrs;
    int64_t bytes = 0;
    template <class T>
    T* alloc(size_t count) {
        void* p = nullptr;
        size_t b = std::max<size_t>(count * sizeof(T), 16);
        hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess) {
            (void)hipGetLastError();  // clear it: a later launch check must not report this failure
            fail(C2_ERR_NOMEM, "hipMalloc(" + std::to_string(b) + " bytes): " + hipGetErrorString(e));
        }
        ptrs.push_back(p);
        bytes += (int64_t)b;
        return static_cast<T*>(p);
    }
    template <class T>
    T* upload(const std::vector<T>& v) {
        T* p = alloc<T>(v.size());
        if (!v.empty()) HIP_CHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return p;
    }
    ~DeviceArena() {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

// Half extent along y of a disk's box, Disk::GetBoundingBox (Disk.cpp:34-43): r * sinTheta of the
// world normal. q is the shared sphere/disk record (centre, radius, local-to-world, world-to-local).
static float disk_offset_y(const float* q) {
    V3 wn = mat3_mul(q + 4, V3{0, 1, 0});
    float sinTheta = std::sqrt(1.f - wn.y * wn.y);
    return q[3] * sinTheta;
}

static void validate_desc(const c2_scene_desc* d) {
    if (!d) fail(C2_ERR_INVALID, "null scene description");
    if (d->num_entities < 0 || d->num_triangles < 0 || d->num_spheres < 0 || d->num_lights < 0 ||
        d->num_materials < 0)
        fail(C2_ERR_INVALID, "negative counts in scene description");
    if (d->num_entities > 0 && (!d->entity_shape_kind || !d->entity_shape_index || !d->entity_material ||
                                !d->entity_light))
        fail(C2_ERR_INVALID, "missing entity arrays");
    for (int e = 0; e < d->num_entities; e++) {
        int k = d->entity_shape_kind[e], s = d->entity_shape_index[e];
        if (k == C2_SHAPE_TRIANGLE) {
            if (s < 0 || s >= d->num_triangles) fail(C2_ERR_INVALID, "entity triangle index out of range");
        } else if (k == C2_SHAPE_SPHERE || k == C2_SHAPE_DISK) {
            if (s < 0 || s >= d->num_spheres) fail(C2_ERR_INVALID, "entity sphere index out of range");
        } else {
            fail(C2_ERR_INVALID, "unknown entity shape kind");
        }
        int m = d->entity_material[e], l = d->entity_light[e];
        if (m < -1 || m >= d->num_materials) fail(C2_ERR_INVALID, "entity material out of range");
        if (l < -1 || l >= d->num_lights) fail(C2_ERR_INVALID, "entity light out of range");
    }
    for (int l = 0; l < d->num_lights; l++) {
        const c2_light& L = d->lights[l];
        if (L.type == C2_LIGHT_AREA && (L.entity < 0 || L.entity >= d->num_entities))
            fail(C2_ERR_INVALID, "area light entity out of range");
        if (L.type != C2_LIGHT_AREA && L.type != C2_LIGHT_POINT) fail(C2_ERR_INVALID, "unknown light type");
    }
    for (int m = 0; m < d->num_materials; m++) {
        const c2_material& M = d->materials[m];
        if (M.type < C2_MAT_DIFFUSE || M.type > C2_MAT_ORENNAYAR) fail(C2_ERR_INVALID, "invalid material type");
        if ((M.type == C2_MAT_MICRO || M.type == C2_MAT_PLASTIC) && M.ndf != C2_NDF_GGX && M.ndf != C2_NDF_BECKMANN)
            fail(C2_ERR_INVALID, "invalid NDF");
        if (M.kd_texture < 0 || M.kd_texture > d->num_textures) fail(C2_ERR_INVALID, "material texture id out of range");
        if (M.kd_texture > 0 && M.type != C2_MAT_DIFFUSE && M.type != C2_MAT_PHONG)
            fail(C2_ERR_INVALID, "only Diffuse and Phong materials take a Kd texture");
    }
    if (d->num_textures < 0 || (d->num_textures > 0 && !d->textures)) fail(C2_ERR_INVALID, "bad texture array");
    for (int t = 0; t < d->num_textures; t++) {
        const c2_texture& T = d->textures[t];
        if (T.width <= 0 || T.height <= 0 || !T.rgb) fail(C2_ERR_INVALID, "bad texture");
    }
    int sky = 0;
    for (int f = 0; f < 6; f++) {
        if (d->skybox[f] < 0 || d->skybox[f] > d->num_textures) fail(C2_ERR_INVALID, "skybox face id out of range");
        sky += d->skybox[f] > 0;
    }
    if (sky != 0 && sky != 6) fail(C2_ERR_INVALID, "a skybox needs all six faces");
    if ((int64_t)d->num_entities >= (1ll << kSlotBits)) fail(C2_ERR_UNSUPPORTED, "too many entities");
}

// Host copy of the scene in kernel layout + its device mirror.
struct SceneGPU {
    DeviceArena arena;
    DevScene dev{};
    BuiltBvh bvh;
    std::vector<float4> prims;
    std::vector<float> tri_pos, tri_shade, spheres;
    std::vector<int4> ents;
    std::vector<DevLight> lights;
    std::vector<uchar4> texels;
    std::vector<int4> tex;
    std::vector<float> tri_uv;
    std::vector<std::array<int32_t, 6>> sky;  // per scene: skybox face ids in the merged table
    // KD-tree accelerator (built instead of being traversed by the BVH when kd is set)
    bool kd = false;
    // 8-wide tree (nodes8) built beside the 4-wide one
    bool wide8 = false;
    std::vector<Node8> nodes8;  // scene s is rooted at offs[s].node8
    std::vector<int4> kd_nodes{make_int4(-1, 0, 0, 0)};  // node 0: "none"
    std::vector<float4> kd_rec;  // per leaf entry: 5 float4 (see DevScene::kd_rec)
    int64_t kd_refs = 0;
    std::vector<KdMaster> kdm;  // per scene: root and union box
    double build_seconds = 0;

    // Start of each scene in the concatenated arrays (a single scene has offsets 0).
    struct Offsets {
        int32_t node, slot, ent, tri, sph, mat, light, tex, node8;
    };
    std::vector<Offsets> offs;
    std::vector<Node4> nodes;  // every scene's tree; scene s is rooted at offs[s].node
    std::vector<c2_material> mats;

    int leaf_size = kLeafSize;  // binary leaves before the 4-wide collapse

    void build(const c2_scene_desc* d) { build(&d, 1); }

    // Concatenates scenes: each keeps its own BVH (built exactly as when alone) and
    // its indices are shifted by the arrays before it, so traversal from its root
    // visits the same nodes and primitives in the same order as a lone build.
    void build(const c2_scene_desc* const* descs, int n) {
        for (int k = 0; k < n; k++) validate_desc(descs[k]);
        for (int k = 0; k < n; k++) {
            Offsets o{(int32_t)nodes.size(), (int32_t)(prims.size() / 3), (int32_t)ents.size(),
                      (int32_t)(tri_pos.size() / 9), (int32_t)(spheres.size() / 22), (int32_t)mats.size(),
                      (int32_t)lights.size(), (int32_t)tex.size(), (int32_t)nodes8.size()};
            offs.push_back(o);
            append_scene(descs[k], o, k == 0);
            if ((int64_t)ents.size() >= (1ll << kSlotBits) || (int64_t)prims.size() / 3 >= (1ll << kSlotBits))
                fail(C2_ERR_UNSUPPORTED, "too many entities in the scene set");
            if ((int64_t)nodes.size() * (int64_t)sizeof(Node4) >= (1ll << 32) ||
                (int64_t)nodes8.size() * (int64_t)sizeof(Node8) >= (1ll << 32))
                fail(C2_ERR_UNSUPPORTED, "too many BVH nodes (the node array must stay below 4 GiB)");
        }
        dev.nodes = arena.upload(nodes);
        if (wide8) dev.nodes8 = arena.upload(nodes8.empty() ? std::vector<Node8>(1) : nodes8);
        dev.prims = arena.upload(prims);
        dev.entities = arena.upload(ents);
        dev.tri_pos = arena.upload(tri_pos);
        dev.tri_shade = arena.upload(tri_shade);
        dev.spheres = arena.upload(spheres);
        dev.materials = arena.upload(mats);
        dev.lights = arena.upload(lights);
        dev.perm = arena.upload(halton_perm_data());
        dev.perm_off = arena.upload(halton_perm_offsets());
        if (kd) {
            dev.kd_nodes = arena.upload(kd_nodes);
            dev.kd_rec = arena.upload(kd_rec.empty() ? std::vector<float4>(5, make_float4(0, 0, 0, 0)) : kd_rec);
        }
        if (!tex.empty()) {
            dev.texels = arena.upload(texels);
            dev.tex = arena.upload(tex);
            dev.tri_uv = arena.upload(tri_uv);
            // (b / 255)^2.2 per byte value: powf evaluated in double and rounded, the
            // convention of every transcendental here
            std::vector<float> lin(256);
            for (int b = 0; b < 256; b++) lin[b] = (float)std::pow((double)((float)b / 255.f), (double)2.2f);
            dev.lin = arena.upload(lin);
        }
        dev.num_lights = (int32_t)lights.size();
        dev.num_entities = (int32_t)ents.size();
    }

    // One scene in kernel layout, appended at offsets o. keep_bvh keeps the host
    // trees of the first scene (ray-visit counting and accelerator statistics).
    void append_scene(const c2_scene_desc* d, const Offsets& o, bool keep_bvh) {
        int nt = d->num_triangles, ns = d->num_spheres, ne = d->num_entities;
        std::vector<float> tp((size_t)nt * 9), ts((size_t)nt * 12);
        for (int t = 0; t < nt; t++) {
            const float* v = d->triangles + (size_t)t * 24;
            for (int k = 0; k < 3; k++) {
                for (int c = 0; c < 3; c++) {
                    tp[(size_t)t * 9 + k * 3 + c] = v[k * 8 + c];
                    ts[(size_t)t * 12 + k * 3 + c] = v[k * 8 + 3 + c];
                }
            }
            // Triangle::calculateDerivative, Triangle.cpp:128-148
            float a00 = v[6] - v[22], a01 = v[7] - v[23], a10 = v[14] - v[22], a11 = v[15] - v[23];
            float ood = 1.f / (a00 * a11 - a10 * a01);
            float i00 = +a11 * ood, i01 = -a01 * ood, i10 = -a10 * ood, i11 = +a00 * ood;
            V3 du{0, 0, 0};
            if (!(std::isinf(i00) || std::isinf(i01) || std::isnan(i00) || std::isnan(i01) || std::isnan(i10) ||
                  std::isnan(i11))) {
                V3 p0{v[0], v[1], v[2]}, p1{v[8], v[9], v[10]}, p2{v[16], v[17], v[18]};
                V3 a = p0 - p2, b = p1 - p2;
                du = normalize(V3{i00 * a.x + i10 * b.x, i00 * a.y + i10 * b.y, i00 * a.z + i10 * b.z});
            }
            ts[(size_t)t * 12 + 9] = du.x;
            ts[(size_t)t * 12 + 10] = du.y;
            ts[(size_t)t * 12 + 11] = du.z;
        }
        std::vector<float> sp3(d->spheres, d->spheres + (size_t)ns * 22);
        std::vector<Box3> boxes(ne);
        std::vector<int4> le(ne);
        for (int e = 0; e < ne; e++) {
            int k = d->entity_shape_kind[e], sh = d->entity_shape_index[e];
            le[e] = make_int4(k, sh, d->entity_material[e], d->entity_light[e]);
            Box3& b = boxes[e];
            if (k == C2_SHAPE_TRIANGLE) {
                const float* p = &tp[(size_t)sh * 9];
                for (int c = 0; c < 3; c++) {
                    b.lo[c] = fmin_ref(fmin_ref(p[c], p[3 + c]), p[6 + c]);
                    b.hi[c] = fmax_ref(fmax_ref(p[c], p[3 + c]), p[6 + c]);
                }
            } else if (k == C2_SHAPE_DISK) {  // Disk::GetBoundingBox, Disk.cpp:34-43
                const float* q = &sp3[(size_t)sh * 22];
                float off[3] = {q[3], disk_offset_y(q), q[3]};
                for (int c = 0; c < 3; c++) {
                    b.lo[c] = q[c] - off[c];
                    b.hi[c] = q[c] + off[c];
                }
            } else {
                const float* q = &sp3[(size_t)sh * 22];
                for (int c = 0; c < 3; c++) {
                    b.lo[c] = q[c] - q[3];
                    b.hi[c] = q[c] + q[3];
                }
            }
        }
        if (kd) {
            KdTree kt;
            auto k0 = std::chrono::steady_clock::now();
            build_kdtree(boxes, kt);
            build_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - k0).count();
            if (kt.depth > 20) fail(C2_ERR_INVALID, "KD-tree deeper than its limit");
            const int32_t nb = (int32_t)kd_nodes.size() - 1, ob = (int32_t)kd_refs;  // node 0 is shared
            for (size_t i = 1; i < kt.nodes.size(); i++) {
                const KdNode& nd = kt.nodes[i];
                int32_t sb;
                std::memcpy(&sb, &nd.split, 4);
                if (nd.axis < 0) kd_nodes.push_back(make_int4(-1, 0, nd.a + ob, nd.b));
                else kd_nodes.push_back(make_int4(nd.axis, sb, nd.a ? nd.a + nb : 0, nd.b ? nd.b + nb : 0));
            }
            // leaf entries as self-contained records: the entity's box (the reference's
            // GetBoundingBox, as the tree was built on) with its id and kind, then the
            // triangle's vertices or the sphere / disk record index
            for (int32_t e : kt.objs) {
                const Box3& b = boxes[e];
                const int k = le[e].x, sh = le[e].y;
                kd_rec.push_back(make_float4(b.lo[0], b.lo[1], b.lo[2], __int_as_float_host(e + o.ent)));
                kd_rec.push_back(make_float4(b.hi[0], b.hi[1], b.hi[2], __int_as_float_host(k)));
                if (k == C2_SHAPE_TRIANGLE) {
                    const float* v = &tp[(size_t)sh * 9];
                    kd_rec.push_back(make_float4(v[0], v[1], v[2], v[3]));
                    kd_rec.push_back(make_float4(v[4], v[5], v[6], v[7]));
                    kd_rec.push_back(make_float4(v[8], 0.f, 0.f, 0.f));
                } else {
                    kd_rec.push_back(make_float4(__int_as_float_host(sh + o.sph), 0.f, 0.f, 0.f));
                    kd_rec.push_back(make_float4(0.f, 0.f, 0.f, 0.f));
                    kd_rec.push_back(make_float4(0.f, 0.f, 0.f, 0.f));
                }
            }
            kd_refs += (int64_t)kt.objs.size();
            KdMaster m;
            for (int c = 0; c < 3; c++) {
                m.lo[c] = kt.master.lo[c];
                m.hi[c] = kt.master.hi[c];
            }
            m.root = kt.root ? kt.root + nb : 0;
            kdm.push_back(m);
        }
        BuiltBvh local;
        BuiltBvh& bv = keep_bvh ? bvh : local;
        auto t0 = std::chrono::steady_clock::now();
        build_bvh(boxes, bv, leaf_size);
        build_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        // nodes: inner children and leaf slots shifted; empty slots (count 0) stay -1
        for (const Node4& nd0 : bv.nodes) {
            Node4 nd = nd0;
            for (int c = 0; c < 4; c++) {
                uint32_t c8 = (nd.count >> (8 * c)) & 0xffu;
                if (nd.child[c] >= 0) nd.child[c] += o.node;
                else if (c8 != 0) nd.child[c] = ~(~nd.child[c] + o.slot);
            }
            nodes.push_back(nd);
        }
        if (wide8) {  // the 8-wide tree, shifted the same way (empty slots: inverted boxes)
            collapse_bvh8(bv);
            for (const Node8& nd0 : bv.nodes8) {
                Node8 nd = nd0;
                for (int c = 0; c < 8; c++) {
                    if (nd.child[c] >= 0) nd.child[c] += o.node8;
                    else if (nd.lo_x[c] <= nd.hi_x[c]) nd.child[c] = ~(~nd.child[c] + o.slot);
                }
                nodes8.push_back(nd);
            }
        }
        for (size_t slot = 0; slot < bv.order.size(); slot++) {
            int e = bv.order[slot];
            int k = le[e].x, sh = le[e].y;
            if (k == C2_SHAPE_TRIANGLE) {
                const float* p = &tp[(size_t)sh * 9];
                prims.push_back(make_float4(p[0], p[1], p[2], p[3]));
                prims.push_back(make_float4(p[4], p[5], p[6], p[7]));
                prims.push_back(make_float4(p[8], __int_as_float_host(e + o.ent), __int_as_float_host(C2_SHAPE_TRIANGLE),
                                            0.f));
            } else {
                const float* q = &sp3[(size_t)sh * 22];
                prims.push_back(make_float4(q[0], q[1], q[2], q[3]));
                prims.push_back(make_float4(__int_as_float_host(sh + o.sph), k == C2_SHAPE_DISK ? disk_offset_y(q) : 0.f,
                                            0.f, 0.f));
                prims.push_back(make_float4(0.f, __int_as_float_host(e + o.ent), __int_as_float_host(k), 0.f));
            }
        }
        for (int e = 0; e < ne; e++) {
            int4 x = le[e];
            x.y += x.x == C2_SHAPE_TRIANGLE ? o.tri : o.sph;
            if (x.z >= 0) x.z += o.mat;
            if (x.w >= 0) x.w += o.light;
            ents.push_back(x);
        }
        for (int l = 0; l < d->num_lights; l++) {
            const c2_light& L = d->lights[l];
            DevLight D;
            std::memset(&D, 0, sizeof(D));
            D.type = L.type;
            D.entity = L.entity;
            for (int c = 0; c < 3; c++) {
                D.L[c] = L.radiance[c];
                D.pos[c] = L.position[c];
            }
            if (L.type == C2_LIGHT_AREA) {
                D.shape_kind = le[L.entity].x;
                D.shape = le[L.entity].y;
                if (D.shape_kind == C2_SHAPE_TRIANGLE) {  // Triangle::SurfaceArea
                    const float* p = &tp[(size_t)D.shape * 9];
                    V3 v0{p[0], p[1], p[2]}, v1{p[3], p[4], p[5]}, v2{p[6], p[7], p[8]};
                    D.area = length(cross(v1 - v0, v2 - v0)) / 2.f;
                } else if (D.shape_kind == C2_SHAPE_DISK) {  // Disk::SurfaceArea, Disk.cpp:83-86
                    float r = sp3[(size_t)D.shape * 22 + 3];
                    D.area = kPi * r * r;
                } else {  // Sphere::SurfaceArea
                    float r = sp3[(size_t)D.shape * 22 + 3];
                    D.area = 4 * kPi * r * r;
                }
                V3 flux = V3{D.L[0], D.L[1], D.L[2]} * kPi * D.area;
                D.eligible = !is_zero(flux);
                D.entity += o.ent;
                D.shape += D.shape_kind == C2_SHAPE_TRIANGLE ? o.tri : o.sph;
            } else {
                D.eligible = !(D.L[0] == 0.f && D.L[1] == 0.f && D.L[2] == 0.f);
            }
            lights.push_back(D);
        }
        tri_pos.insert(tri_pos.end(), tp.begin(), tp.end());
        tri_shade.insert(tri_shade.end(), ts.begin(), ts.end());
        spheres.insert(spheres.end(), sp3.begin(), sp3.end());
        for (int m = 0; m < d->num_materials; m++) {
            c2_material mm = d->materials[m];
            if (mm.kd_texture > 0) mm.kd_texture += o.tex;
            mats.push_back(mm);
        }
        // textures: ids shift by the textures before this scene; vertex uv per triangle
        for (int t = 0; t < d->num_textures; t++) {
            const c2_texture& T = d->textures[t];
            tex.push_back(make_int4(T.width, T.height, (int32_t)texels.size(), 0));
            for (int64_t i = 0; i < (int64_t)T.width * T.height; i++)
                texels.push_back(make_uchar4(T.rgb[i * 3], T.rgb[i * 3 + 1], T.rgb[i * 3 + 2], 255));
            if (texels.size() >= (1ull << 31)) fail(C2_ERR_UNSUPPORTED, "textures exceed 2^31 texels");
        }
        for (int t = 0; t < nt; t++) {
            const float* v = d->triangles + (size_t)t * 24;
            for (int k = 0; k < 3; k++) {
                tri_uv.push_back(v[k * 8 + 6]);
                tri_uv.push_back(v[k * 8 + 7]);
            }
        }
        std::array<int32_t, 6> sk{};
        for (int f = 0; f < 6; f++) sk[f] = d->skybox[f] > 0 ? d->skybox[f] + o.tex : 0;
        sky.push_back(sk);
    }

    static float __int_as_float_host(int32_t i) {
        float f;
        std::memcpy(&f, &i, 4);
        return f;
    }

    // Algorithmic traversal counts on the binary SAH tree (32-B nodes, leaf <= 4):
    // node box tests and primitive tests per ray, same semantics as the kernels.
    void count_visits(const float* rays, int64_t n, bool any, double& nodes, double& tris) const {
        int64_t cn = 0, ct = 0;
        for (int64_t i = 0; i < n; i++) {
            const float* r = rays + i * 8;
            V3 o{r[0], r[1], r[2]}, d{r[3], r[4], r[5]};
            RayPre rp = make_ray(o, d);
            float tmin = any ? r[6] : 0.f, tmax = r[7];
            float best = INFINITY;
            int stack[128], sp = 0, cur = 0;
            bool done = false;
            while (!done) {
                const Bvh2Node& nd = bvh.bvh2[cur];
                cn++;
                float t0 = tmin, t1 = any ? tmax : fmin_ref(tmax, best);
                bool hitBox = slab(rp, nd.box.lo[0], nd.box.lo[1], nd.box.lo[2], nd.box.hi[0], nd.box.hi[1],
                                   nd.box.hi[2], t0, t1);
                if (hitBox && nd.leaf()) {
                    for (int k = 0; k < -nd.right; k++) {
                        int e = bvh.order[nd.left + k];
                        ct++;
                        int kind = ents[e].x, s = ents[e].y;
                        if (kind != C2_SHAPE_TRIANGLE) continue;
                        const float* p = &tri_pos[(size_t)s * 9];
                        V3 v0{p[0], p[1], p[2]}, v1{p[3], p[4], p[5]}, v2{p[6], p[7], p[8]};
                        if (any) {
                            if (tri_test(o, d, v0, v1, v2, tmin, tmax)) { done = true; break; }
                        } else {
                            V3 res = tri_solve(o, d, v0, v1, v2);
                            if (tri_accept_closest(res) && res.z < best) best = res.z;
                        }
                    }
                } else if (hitBox) {
                    int a = nd.left, b = nd.right;
                    int axis = 0;
                    float ex[3];
                    for (int c = 0; c < 3; c++) ex[c] = nd.box.hi[c] - nd.box.lo[c];
                    axis = (ex[0] >= ex[1] && ex[0] >= ex[2]) ? 0 : (ex[1] >= ex[2] ? 1 : 2);
                    bool neg = axis == 0 ? rp.neg[0] : (axis == 1 ? rp.neg[1] : rp.neg[2]);
                    if (neg) std::swap(a, b);
                    stack[sp++] = b;
                    cur = a;
                    continue;
                }
                if (done || sp == 0) break;
                cur = stack[--sp];
            }
        }
        nodes = n ? (double)cn / n : 0;
        tris = n ? (double)ct / n : 0;
    }
};

}  // namespace c2

using namespace c2;

struct c2_accel {
    SceneGPU scene;
    c2_accel_stats stats{};
};

static int grid_for(int64_t n) {
    int64_t g = (n + kBlock - 1) / kBlock;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, 65535));
}

extern "C" {

int c2_accel_build(const c2_scene_desc* desc, c2_accel** out) { return c2_accel_create(desc, C2_ACCEL_BVH, out); }

int c2_accel_create(const c2_scene_desc* desc, int32_t accel, c2_accel** out) {
    return guard([&] {
        if (!out) fail(C2_ERR_INVALID, "null output");
        if (accel != C2_ACCEL_BVH && accel != C2_ACCEL_KDTREE && accel != C2_ACCEL_BVH8)
            fail(C2_ERR_INVALID, "unknown accelerator kind");
        std::unique_ptr<c2_accel> a(new c2_accel());
        a->scene.kd = accel == C2_ACCEL_KDTREE;
        a->scene.wide8 = accel == C2_ACCEL_BVH8;
        a->scene.build(desc);
        a->stats.num_nodes_bvh2 = (int64_t)a->scene.bvh.bvh2.size();
        a->stats.num_nodes_bvh4 = (int64_t)a->scene.nodes.size();
        a->stats.max_depth_bvh4 = a->scene.bvh.depth4;
        a->stats.leaf_size = kLeafSize;
        a->stats.build_seconds = a->scene.build_seconds;
        a->stats.device_bytes = a->scene.arena.bytes;
        *out = a.release();
    });
}

void c2_accel_destroy(c2_accel* a) { delete a; }

int c2_accel_get_stats(const c2_accel* a, c2_accel_stats* out) {
    return guard([&] {
        if (!a || !out) fail(C2_ERR_INVALID, "null argument");
        *out = a->stats;
    });
}

int c2_accel_intersect_device(c2_accel* a, const float* rays, int64_t n, int32_t* prim, float* t, void* stream) {
    return guard([&] {
        if (!a || n < 0 || (n > 0 && (!rays || !prim || !t))) fail(C2_ERR_INVALID, "bad arguments");
        if (n == 0) return;
        if (a->scene.ents.empty()) {
            hipLaunchKernelGGL(k_fill_miss, dim3(grid_for(n)), dim3(kBlock), 0, (hipStream_t)stream, prim, t, n);
            HIP_CHECK(hipGetLastError());
            return;
        }
        if (a->scene.kd)
            hipLaunchKernelGGL(k_intersect<kAccKd>, dim3(grid_for(n)), dim3(kBlock), 0, (hipStream_t)stream,
                               a->scene.dev, rays, n, prim, t, a->scene.kdm[0]);
        else if (a->scene.wide8)
            hipLaunchKernelGGL(k_intersect<kAccBvh8>, dim3(grid_for(n)), dim3(kBlock), 0, (hipStream_t)stream,
                               a->scene.dev, rays, n, prim, t, KdMaster{});
        else
            hipLaunchKernelGGL(k_intersect<kAccBvh4>, dim3(grid_for(n)), dim3(kBlock), 0, (hipStream_t)stream,
                               a->scene.dev, rays, n, prim, t, KdMaster{});
        HIP_CHECK(hipGetLastError());
    });
}

int c2_accel_intersect_test_device(c2_accel* a, const float* rays, const int32_t* ignore, int64_t n, uint8_t* hit,
                                   void* stream) {
    return guard([&] {
        if (!a || n < 0 || (n > 0 && (!rays || !hit))) fail(C2_ERR_INVALID, "bad arguments");
        if (n == 0) return;
        if (a->scene.ents.empty()) {
            hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(n)), dim3(kBlock), 0, (hipStream_t)stream, hit, n, (uint8_t)0);
            HIP_CHECK(hipGetLastError());
            return;
        }
        if (a->scene.kd)
            hipLaunchKernelGGL(k_intersect_test<kAccKd>, dim3(grid_for(n)), dim3(kBlock), 0, (hipStream_t)stream,
                               a->scene.dev, rays, ignore, n, hit, a->scene.kdm[0]);
        else if (a->scene.wide8)
            hipLaunchKernelGGL(k_intersect_test<kAccBvh8>, dim3(grid_for(n)), dim3(kBlock), 0, (hipStream_t)stream,
                               a->scene.dev, rays, ignore, n, hit, KdMaster{});
        else
            hipLaunchKernelGGL(k_intersect_test<kAccBvh4>, dim3(grid_for(n)), dim3(kBlock), 0, (hipStream_t)stream,
                               a->scene.dev, rays, ignore, n, hit, KdMaster{});
        HIP_CHECK(hipGetLastError());
    });
}

int c2_accel_intersect(c2_accel* a, const float* rays, int64_t n, int32_t* prim, float* t) {
    return guard([&] {
        if (!a || n < 0 || (n > 0 && (!rays || !prim || !t))) fail(C2_ERR_INVALID, "bad arguments");
        if (n == 0) return;
        DeviceArena tmp;
        float* dr = tmp.alloc<float>((size_t)n * 8);
        int32_t* dp = tmp.alloc<int32_t>(n);
        float* dt = tmp.alloc<float>(n);
        HIP_CHECK(hipMemcpy(dr, rays, (size_t)n * 8 * sizeof(float), hipMemcpyHostToDevice));
        int rc = c2_accel_intersect_device(a, dr, n, dp, dt, nullptr);
        if (rc != C2_OK) fail(rc, c2_last_error());
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipMemcpy(prim, dp, n * sizeof(int32_t), hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(t, dt, n * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int c2_accel_intersect_test(c2_accel* a, const float* rays, const int32_t* ignore, int64_t n, uint8_t* hit) {
    return guard([&] {
        if (!a || n < 0 || (n > 0 && (!rays || !hit))) fail(C2_ERR_INVALID, "bad arguments");
        if (n == 0) return;
        DeviceArena tmp;
        float* dr = tmp.alloc<float>((size_t)n * 8);
        int32_t* di = ignore ? tmp.alloc<int32_t>(n) : nullptr;
        uint8_t* dh = tmp.alloc<uint8_t>(n);
        HIP_CHECK(hipMemcpy(dr, rays, (size_t)n * 8 * sizeof(float), hipMemcpyHostToDevice));
        if (ignore) HIP_CHECK(hipMemcpy(di, ignore, n * sizeof(int32_t), hipMemcpyHostToDevice));
        int rc = c2_accel_intersect_test_device(a, dr, di, n, dh, nullptr);
        if (rc != C2_OK) fail(rc, c2_last_error());
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipMemcpy(hit, dh, n, hipMemcpyDeviceToHost));
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------- staging without allocation
// c2_accel_intersect_noalloc stages rays and results in this file-scope device array
// instead of allocating device buffers per call: C2_NOALLOC_MAX_RAYS rays of 32 B in,
// then 4 B ids and 4 B t out (40 MiB). Each GPU holds its own copy until the process
// ends. In host code `scratch` only names the symbol; its device address, per device,
// comes from hipGetSymbolAddress on the first call made with that device current.
namespace {
constexpr int64_t kNoallocRays = C2_NOALLOC_MAX_RAYS;
constexpr size_t kScratchBytes = (size_t)kNoallocRays * (8 * sizeof(float) + sizeof(int32_t) + sizeof(float));
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

unsigned char* scratch_for_current_device() {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) fail(C2_ERR_UNSUPPORTED, "device index beyond the staging table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) fail(C2_ERR_DEVICE, "hipGetSymbolAddress returned no address for the staging array");
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    return g_scratch_ptr[dev];
}
}  // namespace

extern "C" {

int c2_accel_intersect_noalloc(c2_accel* a, const float* rays, int64_t n, int32_t* prim, float* t) {
    return guard([&] {
        if (!a || n < 0 || (n > 0 && (!rays || !prim || !t))) fail(C2_ERR_INVALID, "bad arguments");
        if (n > kNoallocRays)
            fail(C2_ERR_INVALID, "c2_accel_intersect_noalloc: " + std::to_string(n) + " rays exceed the staging maximum of " +
                                     std::to_string(kNoallocRays));
        if (n == 0) return;
        if (a->scene.ents.empty()) {  // no device work: every ray misses
            for (int64_t i = 0; i < n; i++) {
                prim[i] = -1;
                t[i] = INFINITY;
            }
            return;
        }
        unsigned char* base = scratch_for_current_device();
        float* dr = reinterpret_cast<float*>(base);
        int32_t* dp = reinterpret_cast<int32_t*>(base + (size_t)n * 8 * sizeof(float));
        float* dt = reinterpret_cast<float*>(base + (size_t)n * (8 * sizeof(float) + sizeof(int32_t)));
        HIP_CHECK(hipMemcpy(dr, rays, (size_t)n * 8 * sizeof(float), hipMemcpyHostToDevice));
        int rc = c2_accel_intersect_device(a, dr, n, dp, dt, nullptr);
        if (rc != C2_OK) fail(rc, c2_last_error());
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipMemcpy(prim, dp, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(t, dt, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int c2_accel_count_visits(const c2_accel* a, const float* rays, int64_t n, int32_t any_hit, double* mean_nodes,
                          double* mean_tris) {
    return guard([&] {
        if (!a || n < 0 || (n > 0 && !rays) || !mean_nodes || !mean_tris) fail(C2_ERR_INVALID, "bad arguments");
        if (a->scene.ents.empty()) {
            *mean_nodes = *mean_tris = 0;
            return;
        }
        a->scene.count_visits(rays, n, any_hit != 0, *mean_nodes, *mean_tris);
    });
}

}  // extern "C"

// ============================================================================ path tracer handle

struct c2_pt {
    std::unique_ptr<c2_accel> accel;
    bool use8 = false;  // traverse the 8-wide tree (accelerator C2_ACCEL_BVH8, or a batch's choice)
    bool group = true;  // group the queue by shading class from bounce kSortFrom (k_class_*)
    // direct lighting: some entity is transparent without a light, so a path may pass
    // through it (otherwise every path ends at its first hit, and one step is all)
    bool dl_chain = true;
    // the largest material type k_shade and k_nee are instantiated for: C2_MAT_MIRROR
    // (Diffuse and Mirror only), C2_MAT_MICRO (also Glass and Microfacet) or kMatAll; the
    // smaller instantiations leave out the other materials' code (fewer registers)
    int mat_bound = kMatAll;
    bool profiling = false;
    uint32_t prof_mask = 0;  // kernel classes timed with events
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> events;
    // ray capture request (c2_pt_capture_rays)
    int cap_bounce = -1, cap_kind = 0;
    int64_t cap_n = 0, cap_written = 0;
    float* cap_rays = nullptr;
    int32_t* cap_ignore = nullptr;
    c2_render_opts opts{};
    PassParams pp{};
    DeviceArena arena;
    PathBuffers B{};  // lane 0 (the radiance cache and ray capture run on lane 0 only)
    PathBuffers BL[3]{};  // lanes 1..3: further sets of path buffers (see enqueue_passes)
    int lanes = 1;
    int32_t chunk = 0;      // paths in flight per launch (at most chunk_cap)
    int32_t chunk_cap = 0;  // paths the buffers hold
    // views (one per scene of a batch) and the views taking part in each sample pass
    std::vector<View> views;
    struct PassTab {
        int32_t off, npv;  // entries pv[off, off + npv)
        int64_t npix;      // linear pass indices
    };
    std::vector<PassTab> passes;  // a single renderer uses passes[0] for every sample
    int2* pv = nullptr;
    // persistent kernel: every view as (view, first item), views in order of falling spp
    // (the longest pixels are handed out first), the items and the item counter
    int2* ppv = nullptr;
    int32_t nppv = 0;
    int64_t pitems = 0;
    int32_t* pcounter = nullptr;
    int64_t film_pixels = 0;
    float4* film = nullptr;
    float4* comp = nullptr;
    double4* film64 = nullptr;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;  // MIS rays, concurrent with the shadow rays
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // lanes 1..3: stream, MIS stream, fork / join events; per lane: end of its work and
    // its last film accumulation
    hipStream_t lstream[3] = {}, lside[3] = {};
    hipEvent_t lfork[3] = {}, ljoin[3] = {}, lend[3] = {};
    hipEvent_t ev_start = nullptr;
    hipEvent_t ev_acc[4] = {};
    int64_t launches = 0;
    int64_t samples = 0;
    // radiance cache
    c2_nrc* nrc = nullptr;
    DeviceArena cache_arena;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> cache_events;  // 0 infer, 1 train, 2 collect
    int64_t cache_queries = 0, cache_rows = 0;

    ~c2_pt() {
        // drain every stream before the arenas release the buffers they use
        for (hipStream_t q : {side, stream, lside[0], lstream[0], lside[1], lstream[1], lside[2], lstream[2]})
            if (q) (void)hipStreamSynchronize(q);
        for (auto* v : {&events, &cache_events})
            for (auto& e : *v) {
                (void)hipEventDestroy(e.second.first);
                (void)hipEventDestroy(e.second.second);
            }
        for (hipEvent_t e : {ev_fork, ev_join, ev_start, ev_acc[0], ev_acc[1], ev_acc[2], ev_acc[3]})
            if (e) (void)hipEventDestroy(e);
        for (int l = 0; l < 3; l++)
            for (hipEvent_t e : {lfork[l], ljoin[l], lend[l]})
                if (e) (void)hipEventDestroy(e);
        for (hipStream_t q : {side, stream, lside[0], lstream[0], lside[1], lstream[1], lside[2], lstream[2]})
            if (q) (void)hipStreamDestroy(q);
    }
};

namespace c2 {
void nrc_launch_infer(c2_nrc* c, const int32_t* countp, int32_t maxn, float4* recs, hipStream_t s);
void nrc_launch_train(c2_nrc* c, const float4* feats, const float4* targets, const int32_t* countp, int32_t js,
                      int32_t ns, int32_t maxn, hipStream_t s, bool adam);
}  // namespace c2

extern "C" {

}  // extern "C"

namespace {
// One image of a renderer: film size, Halton sample count (pass participation) and
// the pixel window that receives samples (win_w = 0: the whole film).
struct ViewSpec {
    int32_t width, height, spp;
    int32_t win_x0, win_y0, win_w, win_h;
};

void check_opts(const c2_render_opts* opts) {
    if (opts->max_depth < 0) fail(C2_ERR_INVALID, "render options: negative depth");
    if (opts->arg_order != C2_ARGORDER_RTL && opts->arg_order != C2_ARGORDER_LTR)
        fail(C2_ERR_INVALID, "render options: bad arg_order");
    if (opts->film_mode < C2_FILM_F32 || opts->film_mode > C2_FILM_F64)
        fail(C2_ERR_INVALID, "render options: bad film_mode");
    if (opts->queue_blocks < 0 || opts->paths_in_flight < 0)
        fail(C2_ERR_INVALID, "render options: negative queue_blocks / paths_in_flight");
    if (opts->integrator != C2_INTEGRATOR_PATH && opts->integrator != C2_INTEGRATOR_DIRECT)
        fail(C2_ERR_INVALID, "render options: bad integrator");
    if (opts->accel != C2_ACCEL_BVH && opts->accel != C2_ACCEL_KDTREE && opts->accel != C2_ACCEL_BVH8)
        fail(C2_ERR_INVALID, "render options: bad accel");
}

// Builds a renderer over n scenes (one view each) sharing one concatenated scene,
// path state sized for the largest sample pass, and one film per view.
// Per-path buffers of one lane for `ch` paths in flight with `lmax` shadow slots each;
// the eligible-light list and the running ray totals are shared by the lanes.
void alloc_path_buffers(PathBuffers& B, DeviceArena& A, size_t ch, int lmax, const int32_t* elig,
                        unsigned long long* totals) {
    const size_t slots = ch * (size_t)lmax;
    B.ray_o = A.alloc<float4>(ch);
    B.ray_d = A.alloc<float4>(ch);
    B.beta = A.alloc<float4>(ch);
    B.beta2 = A.alloc<float4>(ch);
    B.L = A.alloc<float4>(ch);
    B.meta = A.alloc<int4>(ch);
    B.vid = A.alloc<int32_t>(ch);
    B.hit = A.alloc<float4>(ch);
    B.vtx_p = A.alloc<float4>(ch);
    B.vtx_n = A.alloc<float4>(ch);
    B.vtx_du = A.alloc<float4>(ch);
    B.vtx_d = A.alloc<float4>(ch);
    B.vtx_i = A.alloc<int4>(ch);
    B.vtx_h = A.alloc<int2>(ch);
    B.nee_out = A.alloc<int32_t>(ch);
    B.cap = (int32_t)ch;
    B.sh_mask = A.alloc<uint32_t>(ch * (size_t)((lmax + kAppendLights - 1) / kAppendLights));
    B.mis_hit = A.alloc<int32_t>(ch);
    B.elig = elig;
    B.texel = A.alloc<int32_t>(ch);
    B.mis_w = A.alloc<float4>(ch);
    B.mis_f = A.alloc<float4>(ch);
    B.mis_L = A.alloc<float4>(ch);
    B.sh_o = A.alloc<float4>(slots);
    B.sh_d = A.alloc<float4>(slots);
    B.sh_c = A.alloc<float4>(slots);
    B.sh_vis = A.alloc<uint8_t>(slots);
    B.queue[0] = A.alloc<int32_t>(ch);
    B.queue[1] = A.alloc<int32_t>(ch);
    B.sh_list = A.alloc<int32_t>(slots);
    B.mis_list = A.alloc<int32_t>(ch);
    B.queue_s = A.alloc<int32_t>(ch);
    B.cls = A.alloc<uint8_t>(ch);
    B.ccount = A.alloc<int32_t>(2 * kClasses);
    B.counters = A.alloc<int32_t>(8);
    B.totals = totals;
    HIP_CHECK(hipMemset(B.counters, 0, 8 * sizeof(int32_t)));
    HIP_CHECK(hipMemset(B.ccount, 0, 2 * kClasses * sizeof(int32_t)));
}

void pt_build(c2_pt* pt, const c2_scene_desc* const* descs, const ViewSpec* vs, int n, const c2_render_opts* opts,
              bool per_sample_passes) {
    check_opts(opts);
    for (int k = 0; k < n; k++) {
        const ViewSpec& v = vs[k];
        if (v.width <= 0 || v.height <= 0 || v.spp <= 0) fail(C2_ERR_INVALID, "render options: non-positive size");
        if (v.win_w > 0 && (v.win_x0 < 0 || v.win_y0 < 0 || v.win_h <= 0 || v.win_x0 + v.win_w > v.width ||
                            v.win_y0 + v.win_h > v.height))
            fail(C2_ERR_INVALID, "render options: pixel window outside the film");
    }
    pt->opts = *opts;
    std::unique_ptr<c2_accel> acc(new c2_accel());
    acc->scene.leaf_size = kLeafSizeRender;
    acc->scene.kd = opts->accel == C2_ACCEL_KDTREE;
    // a batch keeps both widths, so that its autotuner can choose (c2_batch_autotune)
    acc->scene.wide8 = opts->accel == C2_ACCEL_BVH8 || (per_sample_passes && opts->accel == C2_ACCEL_BVH);
    acc->scene.build(descs, n);
    pt->accel = std::move(acc);
    const SceneGPU& SG = pt->accel->scene;
    {
        uint64_t classes = 0;  // shading classes of the entities' materials (all scenes)
        for (const int4& e : SG.ents)
            if (e.z >= 0) classes |= 1ull << shade_class_host(SG.mats[e.z]);
        pt->group = __builtin_popcountll(classes) >= kGroupMinClasses;
        bool pass_through = false;
        int maxType = -1;
        for (const int4& e : SG.ents) {
            pass_through = pass_through || (e.z < 0 && e.w < 0);
            if (e.z >= 0) maxType = std::max(maxType, SG.mats[e.z].type);
        }
        pt->dl_chain = pass_through;
        pt->mat_bound = maxType <= C2_MAT_MIRROR ? C2_MAT_MIRROR : maxType <= C2_MAT_MICRO ? C2_MAT_MICRO : kMatAll;
    }
    PassParams& P = pt->pp;
    pt->use8 = opts->accel == C2_ACCEL_BVH8;
    P.wide8 = pt->use8 ? 1 : 0;
    P.max_depth = opts->max_depth;
    P.arg_order = opts->arg_order;
    P.integrator = opts->integrator;
    P.cache_bounce = -1;
    P.nsamp = 1;
    P.train_ratio = 1;
    P.train_cap = 0;
    P.train_paths = 0;
    std::vector<int32_t> elig;
    int64_t film_off = 0;
    int32_t maxElig = 0;
    for (int k = 0; k < n; k++) {
        const ViewSpec& v = vs[k];
        View V{};
        V.width = v.width;
        V.height = v.height;
        V.spp = v.spp;
        V.root = descs[k]->num_entities > 0 ? SG.offs[k].node : -1;
        V.root8 = descs[k]->num_entities > 0 && SG.wide8 ? SG.offs[k].node8 : -1;
        if (SG.kd) {
            V.root = SG.kdm[k].root;
            for (int c = 0; c < 3; c++) {
                V.kd_lo[c] = SG.kdm[k].lo[c];
                V.kd_hi[c] = SG.kdm[k].hi[c];
            }
        }
        // camera (PerspCamera ctor: tan evaluated on the host with the float libm call)
        const c2_camera& c = descs[k]->camera;
        float aspect = (float)v.width / v.height;
        V.camH = 1.0f * std::tan(c.fov / 2);
        V.camW = V.camH * aspect;
        V.eye = V3{c.position[0], c.position[1], c.position[2]};
        V3 at{c.look_at[0], c.look_at[1], c.look_at[2]}, up0{c.up[0], c.up[1], c.up[2]};
        V.look = normalize(at - V.eye);
        V3 pr = dot(up0, V.look) * V.look;
        V.up = normalize(up0 - pr);
        V.right = normalize(cross(V.look, up0));
        V.win_x0 = v.win_w > 0 ? v.win_x0 : 0;
        V.win_y0 = v.win_w > 0 ? v.win_y0 : 0;
        V.win_w = v.win_w > 0 ? v.win_w : v.width;
        V.win_h = v.win_w > 0 ? v.win_h : v.height;
        HaltonParams hp = halton_params(v.width, v.height);
        V.stride0 = hp.stride;
        V.exp0 = hp.base_exponents[0];
        V.bexp1 = hp.base_exponents[1];
        V.scale1 = hp.base_scales[1];
        V.bscale0 = hp.base_scales[0];
        V.mult0 = hp.mult_inverse[0];
        V.mult1 = hp.mult_inverse[1];
        V.elig_begin = (int32_t)elig.size();
        int32_t lb = SG.offs[k].light, le = lb + descs[k]->num_lights, nd = 0;
        for (int32_t l = lb; l < le; l++) {
            const DevLight& L = SG.lights[l];
            if (!L.eligible) continue;
            elig.push_back(l);
            if (L.type != C2_LIGHT_POINT) nd++;
        }
        V.num_elig = (int32_t)elig.size() - V.elig_begin;
        V.nd_elig = nd;
        maxElig = std::max(maxElig, V.num_elig);
        V.film_off = film_off;
        for (int f = 0; f < 6; f++) V.sky[f] = SG.sky[k][f];
        film_off += (int64_t)v.width * v.height;
        pt->views.push_back(V);
    }
    if (film_off >= (1ll << 31)) fail(C2_ERR_UNSUPPORTED, "films exceed 2^31 pixels");
    pt->film_pixels = film_off;
    P.lmax = std::max<int32_t>(maxElig, 1);
    // sample passes: which views take part (spp > k) and where their pixels start
    std::vector<int2> pv;
    auto view_pixels = [&](int k) {
        const ViewSpec& v = vs[k];
        return v.win_w > 0 ? (int64_t)v.win_w * v.win_h : (int64_t)v.width * v.height;
    };
    int32_t npasses = 1;
    if (per_sample_passes)
        for (int k = 0; k < n; k++) npasses = std::max(npasses, vs[k].spp);
    for (int32_t k = 0; k < npasses; k++) {
        c2_pt::PassTab t{(int32_t)pv.size(), 0, 0};
        for (int v = 0; v < n; v++) {
            if (per_sample_passes && vs[v].spp <= k) continue;
            pv.push_back(make_int2(v, (int32_t)t.npix));
            t.npix += view_pixels(v);
            t.npv++;
        }
        if (t.npix >= (1ll << 31)) fail(C2_ERR_UNSUPPORTED, "a sample pass exceeds 2^31 paths");
        pt->passes.push_back(t);
    }
    int64_t maxPass = 0;
    for (auto& t : pt->passes) maxPass = std::max(maxPass, t.npix);
    // default: the samples per launch of the largest pass (enqueue_passes), half as many again
    // for a single renderer, whose chunks grow that far when it ends their count on a
    // multiple of the lanes (plan_passes)
    const int64_t spl = per_sample_passes                          ? kBatchSamplesPerChunk
                        : opts->integrator == C2_INTEGRATOR_DIRECT ? samples_per_launch(kDirectSamplesPerChunk, maxPass)
                                                                   : samples_per_launch(kSamplesPerChunk, maxPass);
    const int64_t dflt = maxPass * (per_sample_passes ? spl : spl + spl / 2);
    int64_t chunk = opts->paths_in_flight > 0 ? opts->paths_in_flight : dflt;
    // keep the path buffers within ~16 GiB: 8 GiB in each of two lanes, or 16 GiB in a
    // batch's one lane
    int64_t perPath = (int64_t)P.lmax * (3 * 16 + 4 + 1) + 16 * 15 + 48 + 4 * ((P.lmax + 31) / 32);
    int64_t maxChunk = std::max<int64_t>(4096, ((int64_t)(per_sample_passes ? 16 : 8) << 30) / perPath);
    chunk = std::max<int64_t>(1, std::min(std::min(chunk, std::max(maxPass, dflt)), maxChunk));
    pt->chunk = pt->chunk_cap = (int32_t)chunk;
    PathBuffers& B = pt->B;
    DeviceArena& A = pt->arena;
    View* dviews = A.upload(pt->views);
    pt->pv = A.upload(pv);
    {
        std::vector<int> order(n);
        for (int k = 0; k < n; k++) order[k] = k;
        std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return vs[a].spp > vs[c].spp; });
        std::vector<int2> ppv;
        int64_t items = 0;
        for (int k : order) {
            ppv.push_back(make_int2(k, (int32_t)items));
            items += view_pixels(k);
        }
        pt->pitems = items;
        pt->nppv = n;
        pt->ppv = A.upload(ppv);
        pt->pcounter = A.alloc<int32_t>(1);
    }
    P.views = dviews;
    if (elig.empty()) elig.push_back(-1);
    const int32_t* delig = A.upload(elig);
    unsigned long long* totals = A.alloc<unsigned long long>(4);
    HIP_CHECK(hipMemset(totals, 0, 4 * sizeof(unsigned long long)));
    pt->lanes = per_sample_passes ? 1 : kLanes;  // batches keep one lane
    alloc_path_buffers(B, A, (size_t)chunk, P.lmax, delig, totals);
    for (int l = 1; l < pt->lanes; l++) alloc_path_buffers(pt->BL[l - 1], A, (size_t)chunk, P.lmax, delig, totals);
    // Every view's film covers its whole frame: k_accumulate, the reset and the film
    // read-back index it by full-frame pixel (y * width + x), also when a pixel
    // window restricts which pixels receive samples.
    if (opts->film_mode == C2_FILM_F64) {
        pt->film64 = A.alloc<double4>(film_off);
    } else {
        pt->film = A.alloc<float4>(film_off);
        if (opts->film_mode == C2_FILM_F32_KAHAN) pt->comp = A.alloc<float4>(film_off);
    }
    HIP_CHECK(hipStreamCreateWithFlags(&pt->stream, hipStreamNonBlocking));
    HIP_CHECK(hipStreamCreateWithFlags(&pt->side, hipStreamNonBlocking));
    HIP_CHECK(hipEventCreateWithFlags(&pt->ev_fork, hipEventDisableTiming));
    HIP_CHECK(hipEventCreateWithFlags(&pt->ev_join, hipEventDisableTiming));
    if (pt->lanes > 1) {
        HIP_CHECK(hipEventCreateWithFlags(&pt->ev_start, hipEventDisableTiming));
        for (int l = 0; l < pt->lanes; l++) HIP_CHECK(hipEventCreateWithFlags(&pt->ev_acc[l], hipEventDisableTiming));
        for (int l = 0; l < pt->lanes - 1; l++) {
            HIP_CHECK(hipStreamCreateWithFlags(&pt->lstream[l], hipStreamNonBlocking));
            if (kMisSide) HIP_CHECK(hipStreamCreateWithFlags(&pt->lside[l], hipStreamNonBlocking));
            for (hipEvent_t* e : {&pt->lfork[l], &pt->ljoin[l], &pt->lend[l]})
                HIP_CHECK(hipEventCreateWithFlags(e, hipEventDisableTiming));
        }
    }
}

// Zeroes every film with kernels (also valid inside a captured graph).
void enqueue_film_reset(c2_pt* pt, hipStream_t s) {
    int64_t npix = pt->film_pixels;
    if (pt->film64) {
        hipLaunchKernelGGL(k_zero_d4, dim3(grid_for(npix)), dim3(kBlock), 0, s, pt->film64, npix);
    } else {
        hipLaunchKernelGGL(k_zero4, dim3(grid_for(npix)), dim3(kBlock), 0, s, pt->film, npix);
        if (pt->comp) hipLaunchKernelGGL(k_zero4, dim3(grid_for(npix)), dim3(kBlock), 0, s, pt->comp, npix);
    }
    HIP_CHECK(hipGetLastError());
}
}  // namespace

extern "C" {

int c2_pt_create(const c2_scene_desc* desc, const c2_render_opts* opts, c2_pt** out) {
    return guard([&] {
        if (!opts || !out) fail(C2_ERR_INVALID, "null argument");
        std::unique_ptr<c2_pt> pt(new c2_pt());
        ViewSpec v{opts->width, opts->height, opts->spp, opts->win_x0, opts->win_y0, opts->win_w, opts->win_h};
        pt_build(pt.get(), &desc, &v, 1, opts, false);
        *out = pt.release();
        int rc2 = c2_pt_reset(*out);
        if (rc2 != C2_OK) fail(rc2, c2_last_error());
    });
}

void c2_pt_destroy(c2_pt* pt) { delete pt; }

int c2_pt_reset(c2_pt* pt) {
    return guard([&] {
        if (!pt) fail(C2_ERR_INVALID, "null renderer");
        enqueue_film_reset(pt, pt->stream);
        HIP_CHECK(hipStreamSynchronize(pt->stream));
        pt->samples = 0;
    });
}

}  // extern "C"

namespace {
// Events that only time launches on the GPU: without the system-scope release and acquire a
// default event record carries (the host never reads memory through them), so a record adds
// no cache write-back to the launch it brackets.
inline void timing_event(hipEvent_t* e) { HIP_CHECK(hipEventCreateWithFlags(e, hipEventDisableSystemFence)); }

struct KTimer {
    c2_pt* pt;
    int cls;
    hipStream_t s;
    hipEvent_t a = nullptr, b = nullptr;
    KTimer(c2_pt* p, int c, hipStream_t st) : pt(p), cls(c), s(st) {
        if (!(pt->prof_mask & (1u << c))) return;
        timing_event(&a);
        timing_event(&b);
        HIP_CHECK(hipEventRecord(a, s));
    }
    ~KTimer() {
        if (!a) return;
        (void)hipEventRecord(b, s);
        pt->events.push_back({cls, {a, b}});
    }
};

// Copies the first rays of the current extension / shadow queue to the capture buffer.
__global__ void k_capture(PathBuffers B, int kind, int q, int lmax, int64_t n, float* rays, int32_t* ignore,
                          int32_t* written) {
    int cnt = kind == 0 ? B.counters[q] : B.counters[2];
    int64_t m = cnt < n ? cnt : n;
    if (blockIdx.x == 0 && threadIdx.x == 0) *written = (int32_t)m;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        float4 o, d;
        float tmax;
        int ig = -1;
        if (kind == 0) {
            int p = B.queue[q][i];
            o = B.ray_o[p];
            d = B.ray_d[p];
            tmax = INFINITY;
        } else {
            int s = B.sh_list[i];
            o = B.sh_o[s];
            d = B.sh_d[s];
            tmax = 1 - 1e-5f;
            ig = __float_as_int(o.w);
        }
        float* r = rays + i * 8;
        r[0] = o.x; r[1] = o.y; r[2] = o.z; r[3] = d.x; r[4] = d.y; r[5] = d.z; r[6] = 0.f; r[7] = tmax;
        if (ignore) ignore[i] = ig;
    }
}
}  // namespace

extern "C" {

}  // extern "C"

namespace {
// Enqueues Halton sample passes [sample_begin, + sample_count) on stream s. A single
// renderer uses one pass table for every sample; a batch uses table k for sample k
// (the views with spp > k).
void enqueue_passes(c2_pt* pt, int32_t sample_begin, int32_t sample_count, hipStream_t s0, bool per_sample_passes) {
        PassParams P = pt->pp;
        DevScene S = pt->accel->scene.dev;
        const int acc = pt->accel->scene.kd ? kAccKd : pt->use8 ? kAccBvh8 : kAccBvh4;
        P.wide8 = acc == kAccBvh8 ? 1 : 0;
        // lanes: chunk c runs on lane c % L (lane 0 = the caller's stream s0); lane 1
        // starts after the work already queued on s0, and s0 waits for it at the end.
        // One lane with the radiance cache (training between chunks), ray capture and
        // batches (per-sample pass tables, also captured into a graph).
        // A stream under graph capture also keeps one lane (no second branch in the graph).
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        HIP_CHECK(hipStreamIsCapturing(s0, &cap));
        const int L = (pt->lanes > 1 && P.cache_bounce < 0 && pt->cap_bounce < 0 && !per_sample_passes &&
                       cap == hipStreamCaptureStatusNone)
                          ? pt->lanes
                          : 1;
        // the chunks and lanes of this call (host-only plan, c2_sched.h)
        std::vector<SchedChunk> plan;
        {
            std::vector<SchedPass> tabs(pt->passes.size());
            for (size_t t = 0; t < tabs.size(); t++)
                tabs[t] = SchedPass{pt->passes[t].off, pt->passes[t].npv, pt->passes[t].npix};
            SchedInput in{};
            in.passes = tabs.data();
            in.npasses = (int32_t)tabs.size();
            in.per_sample_passes = per_sample_passes;
            in.sample_begin = sample_begin;
            in.sample_count = sample_count;
            in.chunk = pt->chunk;
            // samples per chunk: several whole passes when the buffers hold them (not with ray
            // capture, which reads one pass)
            int64_t maxPass = 0;
            for (const auto& t : pt->passes) maxPass = std::max<int64_t>(maxPass, t.npix);
            in.samples_per_chunk = pt->cap_bounce >= 0                      ? 1
                                   : per_sample_passes                      ? kBatchSamplesPerChunk
                                   : P.cache_bounce >= 0                   ? kCacheSamplesPerChunk
                                   : P.integrator == C2_INTEGRATOR_DIRECT
                                       ? samples_per_launch(kDirectSamplesPerChunk, maxPass)
                                       : samples_per_launch(kSamplesPerChunk, maxPass);
            in.lanes = L;
            if (!per_sample_passes && pt->cap_bounce < 0 && P.cache_bounce < 0 && in.samples_per_chunk > 1)
                in.max_samples_per_chunk = in.samples_per_chunk + in.samples_per_chunk / 2;
            if (plan_passes(in, plan) != 0) fail(C2_ERR_INVALID, "render: sample range outside the pass tables");
        }
        if (L > 1) {
            HIP_CHECK(hipEventRecord(pt->ev_start, s0));
            for (int l = 0; l < L - 1; l++) HIP_CHECK(hipStreamWaitEvent(pt->lstream[l], pt->ev_start, 0));
        }
        for (const SchedChunk& ch : plan) {
            const c2_pt::PassTab& tab = pt->passes[ch.tab];
            const int ns = ch.nsamp;
            P.pv = pt->pv + tab.off;
            P.npv = tab.npv;
            P.sample = ch.sample;
            P.nsamp = ns;
            {
                P.pix_begin = ch.pix_begin;
                P.pix_count = ch.pix_count;
                const int64_t npath = (int64_t)ns * P.pix_count;
                const int lane = ch.lane;
                if (lane < 0 || lane >= L) fail(C2_ERR_INVALID, "render: lane outside the renderer's lanes");
                const hipStream_t s = lane ? pt->lstream[lane - 1] : s0;
                const hipStream_t side = !kMisSide ? s : lane ? pt->lside[lane - 1] : pt->side;
                const hipEvent_t evf = lane ? pt->lfork[lane - 1] : pt->ev_fork;
                const hipEvent_t evj = lane ? pt->ljoin[lane - 1] : pt->ev_join;
                PathBuffers& LB = lane ? pt->BL[lane - 1] : pt->B;
                const int qcap = pt->opts.queue_blocks > 0 ? pt->opts.queue_blocks : kQueueBlocksDefault;
                int g = std::min(grid_for(npath), qcap);
                if (P.cache_bounce >= 0) hipLaunchKernelGGL(k_chunk_reset, dim3(1), dim3(1), 0, s, LB);
                {
                    KTimer kt(pt, C2_K_GENERATE, s);
                    hipLaunchKernelGGL(k_generate, dim3(g), dim3(kBlock), 0, s, S, LB, P);
                }
                pt->launches++;
                int q = 0;
                if (S.num_entities > 0) {
                    // Direct lighting: one vertex, reached through at most kDlChainLimit - 1
                    // transparent surfaces (the reference recurses through them without
                    // consuming depth); depth 0 renders black, as level == maxDepth at
                    // level 0 does.
                    // (without transparent entities that carry no light, every direct-lighting
                    // path ends at its first hit: the later steps would run on empty queues)
                    const int steps = P.integrator == C2_INTEGRATOR_DIRECT
                                          ? (P.max_depth > 0 ? (pt->dl_chain ? kDlChainLimit : 1) : 0)
                                          : P.max_depth;
                    for (int b = 0; b < steps; b++) {
                        if (pt->cap_bounce == b && pt->cap_kind == 0) {
                            hipLaunchKernelGGL(k_capture, dim3(grid_for(pt->cap_n)), dim3(kBlock), 0, s, LB, 0, q,
                                               P.lmax, pt->cap_n, pt->cap_rays, pt->cap_ignore,
                                               LB.counters + 6);
                        }
                        {
                            KTimer kt(pt, C2_K_EXTEND, s);
                            if (acc == kAccKd)
                                hipLaunchKernelGGL(k_extend<kAccKd>, dim3(g), dim3(kBlock), 0, s, S, LB, P, q);
                            else if (acc == kAccBvh8)
                                hipLaunchKernelGGL(k_extend<kAccBvh8>, dim3(g), dim3(kBlock), 0, s, S, LB, P, q);
                            else
                                hipLaunchKernelGGL(k_extend<kAccBvh4>, dim3(g), dim3(kBlock), 0, s, S, LB, P, q);
                        }
                        // the vertex kernels read this bounce's queue grouped by shading class
                        PathBuffers SB = LB;
                        if (b >= kSortFrom && pt->group) {
                            KTimer kt(pt, C2_K_SHADE, s);
                            hipLaunchKernelGGL(k_class_count, dim3(g), dim3(kBlock), 0, s, S, LB, q);
                            hipLaunchKernelGGL(k_class_scatter, dim3(g), dim3(kBlock), 0, s, LB, q);
                            SB.queue[q] = LB.queue_s;
                            pt->launches += 2;
                        }
                        if (S.tex) {
                            KTimer kt(pt, C2_K_SHADE, s);
                            hipLaunchKernelGGL(k_surface, dim3(g), dim3(kBlock), 0, s, S, SB, P, q);
                            pt->launches++;
                        }
                        {
                            KTimer kt(pt, C2_K_SHADE, s);
                            if (pt->mat_bound == C2_MAT_MIRROR)
                                hipLaunchKernelGGL(k_shade<C2_MAT_MIRROR>, dim3(g), dim3(kBlock), 0, s, S, SB, P, q, b);
                            else if (pt->mat_bound == C2_MAT_MICRO)
                                hipLaunchKernelGGL(k_shade<C2_MAT_MICRO>, dim3(g), dim3(kBlock), 0, s, S, SB, P, q, b);
                            else
                                hipLaunchKernelGGL(k_shade<kMatAll>, dim3(g), dim3(kBlock), 0, s, S, SB, P, q, b);
                        }
                        if (P.cache_bounce == b) {
                            hipEvent_t e0, e1;
                            timing_event(&e0);
                            timing_event(&e1);
                            HIP_CHECK(hipEventRecord(e0, s));
                            nrc_launch_infer(pt->nrc, LB.counters + 4, (int32_t)npath, LB.nrc_feat, s);
                            HIP_CHECK(hipEventRecord(e1, s));
                            pt->cache_events.push_back({0, {e0, e1}});
                        }
                        {
                            KTimer kt(pt, C2_K_NEE, s);
                            if (pt->mat_bound == C2_MAT_MIRROR)
                                hipLaunchKernelGGL(k_nee<C2_MAT_MIRROR>, dim3(g), dim3(kBlock), 0, s, S, SB, P, q);
                            else if (pt->mat_bound == C2_MAT_MICRO)
                                hipLaunchKernelGGL(k_nee<C2_MAT_MICRO>, dim3(g), dim3(kBlock), 0, s, S, SB, P, q);
                            else
                                hipLaunchKernelGGL(k_nee<kMatAll>, dim3(g), dim3(kBlock), 0, s, S, SB, P, q);
                        }
                        if (pt->cap_bounce == b && pt->cap_kind == 1) {
                            hipLaunchKernelGGL(k_capture, dim3(grid_for(pt->cap_n)), dim3(kBlock), 0, s, LB, 1, q,
                                               P.lmax, pt->cap_n, pt->cap_rays, pt->cap_ignore,
                                               LB.counters + 6);
                        }
                        // The shadow rays run on the side stream and the rare MIS rays (closest
                        // hit, a small grid) on the lane's stream, concurrently; both only read
                        // what k_nee wrote, and k_resolve waits for both. The MIS launch directly
                        // follows k_nee in its queue, so its blocks are resident before the
                        // shadow-ray grid (behind the cross-stream wait) fills the device: on the
                        // side stream instead, its blocks waited for slots that k_occlude's
                        // blocks freed, and its launches spanned most of k_occlude's.
                        // Small chunks and direct lighting's short steps run both in line on
                        // the lane's stream: the fork and join cost more than the overlap gains.
                        const bool inl = P.integrator == C2_INTEGRATOR_DIRECT || npath < kInlineShadowPaths;
                        hipStream_t os = inl ? s : side;
                        if (!inl) {
                            HIP_CHECK(hipEventRecord(evf, s));
                            HIP_CHECK(hipStreamWaitEvent(side, evf, 0));
                        }
                        int gs = std::min(grid_for(npath * P.lmax), qcap);
                        // small chunks: shadow and MIS rays in one launch
                        const bool fused = inl && npath < kInlineShadowPaths;
                        if (fused) {
                            KTimer kt(pt, C2_K_OCCLUDE, s);
                            const int gm = std::min(g, kMisGrid);
                            if (acc == kAccKd)
                                hipLaunchKernelGGL(k_occlude_mis<kAccKd>, dim3(gs + gm), dim3(kBlock), 0, s, S, LB, P, gs);
                            else if (acc == kAccBvh8)
                                hipLaunchKernelGGL(k_occlude_mis<kAccBvh8>, dim3(gs + gm), dim3(kBlock), 0, s, S, LB, P,
                                                   gs);
                            else
                                hipLaunchKernelGGL(k_occlude_mis<kAccBvh4>, dim3(gs + gm), dim3(kBlock), 0, s, S, LB, P,
                                                   gs);
                        } else {
                            KTimer kt(pt, C2_K_OCCLUDE, os);
                            if (acc == kAccKd)
                                hipLaunchKernelGGL(k_occlude<kAccKd>, dim3(gs), dim3(kBlock), 0, os, S, LB);
                            else if (acc == kAccBvh8)
                                hipLaunchKernelGGL(k_occlude<kAccBvh8>, dim3(gs), dim3(kBlock), 0, os, S, LB);
                            else
                                hipLaunchKernelGGL(k_occlude<kAccBvh4>, dim3(gs), dim3(kBlock), 0, os, S, LB);
                        }
                        if (!inl) HIP_CHECK(hipEventRecord(evj, side));
                        if (!fused) {
                            KTimer kt(pt, C2_K_MIS, s);
                            if (acc == kAccKd)
                                hipLaunchKernelGGL(k_mis<kAccKd>, dim3(std::min(g, kMisGrid)), dim3(kBlock), 0, s, S,
                                                   LB, P);
                            else if (acc == kAccBvh8)
                                hipLaunchKernelGGL(k_mis<kAccBvh8>, dim3(std::min(g, kMisGrid)), dim3(kBlock), 0, s,
                                                   S, LB, P);
                            else
                                hipLaunchKernelGGL(k_mis<kAccBvh4>, dim3(std::min(g, kMisGrid)), dim3(kBlock), 0, s,
                                                   S, LB, P);
                        }
                        if (!inl) HIP_CHECK(hipStreamWaitEvent(s, evj, 0));
                        {
                            KTimer kt(pt, C2_K_RESOLVE, s);
                            hipLaunchKernelGGL(k_resolve, dim3(g), dim3(kBlock), 0, s, SB, P, q);
                        }
                        pt->launches += fused ? 5 : 6;  // extend, shade, nee, occlude (+ mis), resolve
                        q ^= 1;
                    }
                }
                if (P.cache_bounce >= 0) {
                    // each record owns its two events (c2_pt_get_cache_times and the destructor
                    // destroy both of every record)
                    hipEvent_t e0, e1, e2, e3;
                    for (hipEvent_t* e : {&e0, &e1, &e2, &e3}) timing_event(e);
                    HIP_CHECK(hipEventRecord(e0, s));
                    const int64_t pairs = (int64_t)P.train_paths * P.max_depth;  // (training path, bounce)
                    hipLaunchKernelGGL(k_cache_collect, dim3(grid_for((pairs + kCollectPer - 1) / kCollectPer)),
                                       dim3(kBlock), 0, s, LB, P);
                    HIP_CHECK(hipEventRecord(e1, s));
                    HIP_CHECK(hipEventRecord(e2, s));
                    // one Adam step per sample of the launch, each on an equal slice of the rows
                    for (int js = 0; js < ns; js++)
                        nrc_launch_train(pt->nrc, LB.bfeat, LB.btarget, LB.counters + 7, js, ns, P.train_cap, s,
                                         true);
                    HIP_CHECK(hipEventRecord(e3, s));
                    pt->cache_events.push_back({2, {e0, e1}});
                    pt->cache_events.push_back({1, {e2, e3}});
                }
                // films accumulate in chunk order: after the previous chunk's (other lane)
                if (ch.wait_acc >= 0) HIP_CHECK(hipStreamWaitEvent(s, pt->ev_acc[ch.wait_acc], 0));
                {
                    KTimer kt(pt, C2_K_ACCUMULATE, s);
                    hipLaunchKernelGGL(k_accumulate, dim3(g), dim3(kBlock), 0, s, LB, P, pt->opts.film_mode,
                                       pt->film, pt->comp, pt->film64);
                }
                if (L > 1) HIP_CHECK(hipEventRecord(pt->ev_acc[lane], s));
                pt->launches++;
                HIP_CHECK(hipGetLastError());
            }
            pt->samples += (int64_t)ns * ch.pix_count;
        }
        for (int l = 0; l < L - 1; l++) {
            HIP_CHECK(hipEventRecord(pt->lend[l], pt->lstream[l]));
            HIP_CHECK(hipStreamWaitEvent(s0, pt->lend[l], 0));
        }
}

// The persistent kernel serves the BVH accelerator with the path tracer or direct
// lighting; the radiance cache (training between launches) and ray capture (one bounce of
// a pass) need the wavefront.
bool persistent_ok(const c2_pt* pt) {
    return !pt->accel->scene.kd && pt->pp.cache_bounce < 0 && pt->cap_bounce < 0;
}

template <int BS, int W>
int persist_grid() {
    // one block per free slot of every CU (the kernel loops until the queue is drained),
    // per device
    static int grid[64] = {};
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) fail(C2_ERR_UNSUPPORTED, "device index beyond 63");
    if (!grid[dev]) {
        int cus = 0, nb = 0;
        HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(&k_persist<BS, W>), BS, 0));
        grid[dev] = std::max(1, nb) * std::max(1, cus);
    }
    return grid[dev];
}

// Enqueues samples [sample_begin, + sample_count) of every view as one persistent launch
// (after a one-thread kernel zeroes the item counter: no memset node in a captured graph).
template <int BS, int W>
void launch_persist(hipStream_t s, const DevScene& S, const PersistParams& Q) {
    hipLaunchKernelGGL((k_persist<BS, W>), dim3(persist_grid<BS, W>()), dim3(BS), 0, s, S, Q);
}

// Launch shapes of the persistent kernel: threads per block x minimum waves per SIMD.
struct PersistShape {
    int block, waves;
    void (*launch)(hipStream_t, const DevScene&, const PersistParams&);
};
const PersistShape kPersistShapes[] = {
    {128, 0, &launch_persist<128, 0>}, {128, 4, &launch_persist<128, 4>}, {128, 6, &launch_persist<128, 6>},
    {128, 8, &launch_persist<128, 8>}, {256, 0, &launch_persist<256, 0>}, {256, 4, &launch_persist<256, 4>},
    {256, 6, &launch_persist<256, 6>}, {256, 8, &launch_persist<256, 8>}, {512, 0, &launch_persist<512, 0>},
    {512, 4, &launch_persist<512, 4>}, {512, 6, &launch_persist<512, 6>}, {512, 8, &launch_persist<512, 8>},
};

const PersistShape* persist_shape(int block, int waves) {
    for (const PersistShape& p : kPersistShapes)
        if (p.block == block && p.waves == waves) return &p;
    return nullptr;
}

void enqueue_persistent(c2_pt* pt, int32_t sample_begin, int32_t sample_count, hipStream_t s, int block,
                        int waves) {
    if (!persistent_ok(pt)) fail(C2_ERR_UNSUPPORTED, "persistent kernel: needs the BVH, no radiance cache, no capture");
    if (sample_begin < 0 || sample_count < 0) fail(C2_ERR_INVALID, "render: negative sample range");
    if (pt->pitems >= (1ll << 30)) fail(C2_ERR_UNSUPPORTED, "persistent kernel: more than 2^30 pixels");
    const PassParams& P = pt->pp;
    const DevScene S = pt->accel->scene.dev;
    PersistParams Q{};
    Q.views = P.views;
    Q.pv = pt->ppv;
    Q.npv = pt->nppv;
    Q.nitems = (int32_t)pt->pitems;
    Q.counter = pt->pcounter;
    Q.sample_begin = sample_begin;
    Q.sample_end = (int32_t)std::min<int64_t>((int64_t)sample_begin + sample_count, INT32_MAX);
    Q.steps = S.num_entities == 0 ? 0
              : P.integrator == C2_INTEGRATOR_DIRECT ? (P.max_depth > 0 ? kDlChainLimit : 0)
                                                     : P.max_depth;
    Q.arg_order = P.arg_order;
    Q.integrator = P.integrator;
    Q.film_mode = pt->opts.film_mode;
    Q.film = pt->film;
    Q.comp = pt->comp;
    Q.film64 = pt->film64;
    Q.elig = pt->B.elig;
    Q.totals = pt->B.totals;
    const PersistShape* shape = persist_shape(block, waves);
    if (!shape) fail(C2_ERR_INVALID, "persistent kernel: no such launch shape (threads per block, waves per SIMD)");
    hipLaunchKernelGGL(k_set_i32, dim3(1), dim3(1), 0, s, pt->pcounter, 0);
    shape->launch(s, S, Q);
    HIP_CHECK(hipGetLastError());
    pt->launches += 2;
    for (const View& V : pt->views) {
        const int64_t ns = std::max<int64_t>(0, std::min<int64_t>(V.spp, Q.sample_end) - sample_begin);
        pt->samples += ns * V.win_w * V.win_h;
    }
}

// Film of view v as FrameBuffer::GetImageDataRGB32F returns it: color / weight,
// rows flipped (y = 0 at the top).
void read_film(c2_pt* pt, int v, float* rgb) {
    HIP_CHECK(hipDeviceSynchronize());
    const View& V = pt->views[v];
    int W = V.width, H = V.height;
    int64_t npix = (int64_t)W * H;
    if (pt->film64) {
        std::vector<double4> f(npix);
        HIP_CHECK(hipMemcpy(f.data(), pt->film64 + V.film_off, npix * sizeof(double4), hipMemcpyDeviceToHost));
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                const double4& c = f[(size_t)y * W + x];
                float* o = rgb + ((size_t)(H - y - 1) * W + x) * 3;
                o[0] = (float)(c.x / c.w);
                o[1] = (float)(c.y / c.w);
                o[2] = (float)(c.z / c.w);
            }
    } else {
        std::vector<float4> f(npix);
        HIP_CHECK(hipMemcpy(f.data(), pt->film + V.film_off, npix * sizeof(float4), hipMemcpyDeviceToHost));
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                const float4& c = f[(size_t)y * W + x];
                V3 q = V3{c.x, c.y, c.z} / c.w;
                float* o = rgb + ((size_t)(H - y - 1) * W + x) * 3;
                o[0] = q.x;
                o[1] = q.y;
                o[2] = q.z;
            }
    }
}
}  // namespace

extern "C" {

int c2_pt_render(c2_pt* pt, int32_t sample_begin, int32_t sample_count, void* stream) {
    return guard([&] {
        if (!pt) fail(C2_ERR_INVALID, "null renderer");
        if (sample_begin < 0 || sample_count < 0) fail(C2_ERR_INVALID, "negative sample range");
        enqueue_passes(pt, sample_begin, sample_count, stream ? (hipStream_t)stream : pt->stream, false);
    });
}

int c2_pt_sync(c2_pt* pt) {
    return guard([&] {
        if (!pt) fail(C2_ERR_INVALID, "null renderer");
        HIP_CHECK(hipStreamSynchronize(pt->stream));
        HIP_CHECK(hipDeviceSynchronize());
    });
}

int c2_pt_get_film(c2_pt* pt, float* rgb) {
    return guard([&] {
        if (!pt || !rgb) fail(C2_ERR_INVALID, "null argument");
        read_film(pt, 0, rgb);
    });
}

int c2_pt_set_profiling(c2_pt* pt, int32_t enable) { return c2_pt_profile_kernels(pt, enable ? 0xffu : 0u); }

int c2_pt_profile_kernels(c2_pt* pt, uint32_t mask) {
    return guard([&] {
        if (!pt) fail(C2_ERR_INVALID, "null renderer");
        HIP_CHECK(hipDeviceSynchronize());
        for (auto& e : pt->events) {
            (void)hipEventDestroy(e.second.first);
            (void)hipEventDestroy(e.second.second);
        }
        pt->events.clear();
        pt->prof_mask = mask & ((1u << C2_K_COUNT) - 1u);
        pt->profiling = pt->prof_mask != 0;
    });
}

int c2_pt_get_kernel_times(c2_pt* pt, c2_kernel_times* out) {
    return guard([&] {
        if (!pt || !out) fail(C2_ERR_INVALID, "null argument");
        HIP_CHECK(hipDeviceSynchronize());
        c2_kernel_times r{};
        for (auto& e : pt->events) {
            float ms = 0.f;
            HIP_CHECK(hipEventElapsedTime(&ms, e.second.first, e.second.second));
            r.ms[e.first] += ms;
            r.launches[e.first] += 1;
        }
        *out = r;
    });
}

int c2_pt_capture_rays(c2_pt* pt, int32_t sample, int32_t bounce, int32_t kind, float* rays_out,
                       int32_t* ignore_out, int64_t n, int64_t* written) {
    return guard([&] {
        if (!pt || !rays_out || !written || n <= 0 || (kind != 0 && kind != 1) || bounce < 0)
            fail(C2_ERR_INVALID, "bad arguments");
        DeviceArena tmp;
        n = std::min<int64_t>(n, (int64_t)pt->chunk * (kind == 0 ? 1 : pt->pp.lmax));
        pt->cap_rays = tmp.alloc<float>((size_t)n * 8);
        pt->cap_ignore = ignore_out ? tmp.alloc<int32_t>(n) : nullptr;
        pt->cap_n = n;
        pt->cap_kind = kind;
        pt->cap_bounce = bounce;
        int64_t save = pt->chunk;
        int rc = c2_pt_render(pt, sample, 1, nullptr);
        pt->cap_bounce = -1;
        (void)save;
        if (rc != C2_OK) fail(rc, c2_last_error());
        HIP_CHECK(hipStreamSynchronize(pt->stream));
        int32_t w = 0;
        HIP_CHECK(hipMemcpy(&w, pt->B.counters + 6, sizeof(int32_t), hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(rays_out, pt->cap_rays, (size_t)w * 8 * sizeof(float), hipMemcpyDeviceToHost));
        if (ignore_out) HIP_CHECK(hipMemcpy(ignore_out, pt->cap_ignore, (size_t)w * sizeof(int32_t), hipMemcpyDeviceToHost));
        pt->cap_rays = nullptr;
        pt->cap_ignore = nullptr;
        *written = w;
    });
}

int c2_pt_set_cache(c2_pt* pt, c2_nrc* nrc, int32_t cache_bounce, int32_t train_ratio) {
    return guard([&] {
        if (!pt) fail(C2_ERR_INVALID, "null renderer");
        HIP_CHECK(hipStreamSynchronize(pt->stream));
        if (!nrc) {
            pt->nrc = nullptr;
            pt->pp.cache_bounce = -1;
            return;
        }
        if (pt->pp.integrator != C2_INTEGRATOR_PATH) fail(C2_ERR_INVALID, "cache: path tracer only");
        if (cache_bounce < 0 || cache_bounce >= pt->pp.max_depth || train_ratio < 1)
            fail(C2_ERR_INVALID, "cache: bounce must be in [0, depth) and train_ratio >= 1");
        c2_nrc_stats st{};
        int32_t cap = (int32_t)std::min<int64_t>(1 << 16, (int64_t)pt->chunk * pt->pp.max_depth);
        int64_t tpaths = (int64_t)pt->chunk / train_ratio + pt->views[0].height + 64;
        PathBuffers& B = pt->B;
        DeviceArena& A = pt->cache_arena;
        B.nrc_feat = A.alloc<float4>((size_t)pt->chunk * kNrcRec4);
        B.tlist = A.alloc<int32_t>(tpaths);
        B.trec = A.alloc<float4>((size_t)tpaths * pt->pp.max_depth * (kNrcQ4 + 2));
        B.trec_valid = A.alloc<uint8_t>((size_t)tpaths * pt->pp.max_depth);
        B.bfeat = A.alloc<float4>((size_t)cap * kCacheSamplesPerChunk * kNrcQ4);
        B.btarget = A.alloc<float4>((size_t)cap * kCacheSamplesPerChunk);
        (void)st;
        pt->nrc = nrc;
        pt->pp.cache_bounce = cache_bounce;
        pt->pp.train_ratio = train_ratio;
        pt->pp.train_cap = cap;
        pt->pp.train_paths = (int32_t)tpaths;
    });
}

int c2_pt_get_cache_times(c2_pt* pt, c2_cache_times* out) {
    return guard([&] {
        if (!pt || !out) fail(C2_ERR_INVALID, "null argument");
        HIP_CHECK(hipDeviceSynchronize());
        c2_cache_times r{};
        int passes[3] = {0, 0, 0};
        double ms[3] = {0, 0, 0};
        for (auto& e : pt->cache_events) {
            float t = 0;
            HIP_CHECK(hipEventElapsedTime(&t, e.second.first, e.second.second));
            ms[e.first] += t;
            passes[e.first]++;
            (void)hipEventDestroy(e.second.first);
            (void)hipEventDestroy(e.second.second);
        }
        pt->cache_events.clear();
        r.infer_ms = passes[0] ? ms[0] : 0;
        r.train_ms = passes[1] ? ms[1] : 0;
        r.collect_ms = passes[2] ? ms[2] : 0;
        *out = r;
    });
}

int c2_pt_get_stats(const c2_pt* pt, c2_pt_stats* out) {
    return guard([&] {
        if (!pt || !out) fail(C2_ERR_INVALID, "null argument");
        unsigned long long t[4] = {0, 0, 0, 0};
        HIP_CHECK(hipMemcpy(t, pt->B.totals, sizeof(t), hipMemcpyDeviceToHost));
        out->samples = pt->samples;
        out->extension_rays = (int64_t)t[0];
        out->shadow_rays = (int64_t)t[1];
        out->mis_rays = (int64_t)t[2];
        out->kernel_launches = pt->launches;
    });
}

}  // extern "C"

// ============================================================================ HBM calibration copy

namespace c2 {

// Reads n4 float4 and writes n4 float4, each element exactly once.
__global__ void __launch_bounds__(kBlock) k_stream_copy(const float4* __restrict__ src, float4* __restrict__ dst,
                                                        int64_t n4) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace c2

extern "C" int c2_util_stream_copy(int64_t bytes, int32_t reps, double* avg_ms) {
    using namespace c2;
    return guard([&] {
        if (bytes <= 0 || bytes % 16 || bytes > ((int64_t)8 << 30) || reps <= 0 || !avg_ms)
            fail(C2_ERR_INVALID, "bytes must be a positive multiple of 16 up to 8 GiB, reps > 0");
        float4 *src = nullptr, *dst = nullptr;
        HIP_CHECK(hipMalloc(&src, bytes));
        if (hipMalloc(&dst, bytes) != hipSuccess) {
            (void)hipFree(src);
            fail(C2_ERR_DEVICE, "hipMalloc failed");
        }
        const int64_t n4 = bytes / 16;
        const int grid = (int)std::min<int64_t>((n4 + kBlock - 1) / kBlock, 1 << 16);
        hipEvent_t e0, e1;
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        hipError_t err = hipMemset(src, 0, bytes);
        float ms = 0;
        if (err == hipSuccess) {
            hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(kBlock), 0, nullptr, src, dst, n4);  // warm
            (void)hipEventRecord(e0, nullptr);
            for (int r = 0; r < reps; ++r)
                hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(kBlock), 0, nullptr, src, dst, n4);
            (void)hipEventRecord(e1, nullptr);
            err = hipEventSynchronize(e1);
            if (err == hipSuccess) err = hipGetLastError();
            if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipFree(src);
        (void)hipFree(dst);
        if (err != hipSuccess) fail(C2_ERR_DEVICE, std::string("stream copy: ") + hipGetErrorString(err));
        *avg_ms = ms / reps;
    });
}

// ============================================================================ batched scenes

// Many scenes rendered together: every scene is one view of a single renderer over
// the concatenated scene set, and sample pass k runs the wavefront over the pixels
// of every view with spp > k, so small scenes share launches instead of each
// paying for its own. The full render (film reset + every pass) can be captured
// once into a hipGraph and replayed with one launch.
struct c2_batch {
    std::unique_ptr<c2_pt> pt;
    int32_t schedule = C2_SCHED_WAVEFRONT;
    int32_t block = 256;  // persistent kernel: threads per block
    int32_t waves = 0;    // persistent kernel: minimum waves per SIMD (0: as the registers allow)
    int32_t max_spp = 0;
    int64_t samples_per_render = 0;  // sum over scenes of width * height * spp
    int64_t renders = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~c2_batch() {
        if (pt && pt->stream) (void)hipStreamSynchronize(pt->stream);
        drop_graph();
    }
};

namespace {
// The whole render of a batch on stream s: the film reset, then every sample pass of the
// wavefront or the one persistent launch.
void enqueue_batch(c2_batch* b, hipStream_t s) {
    c2_pt* pt = b->pt.get();
    enqueue_film_reset(pt, s);
    if (b->schedule == C2_SCHED_PERSISTENT) enqueue_persistent(pt, 0, b->max_spp, s, b->block, b->waves);
    else enqueue_passes(pt, 0, b->max_spp, s, true);
}
}  // namespace

extern "C" {

int c2_batch_create(const c2_batch_item* items, int32_t n, const c2_render_opts* opts, c2_batch** out) {
    return guard([&] {
        if (!items || n <= 0 || !opts || !out) fail(C2_ERR_INVALID, "bad arguments");
        std::vector<const c2_scene_desc*> descs(n);
        std::vector<ViewSpec> vs(n);
        std::unique_ptr<c2_batch> b(new c2_batch());
        for (int k = 0; k < n; k++) {
            if (!items[k].scene) fail(C2_ERR_INVALID, "null scene in batch");
            descs[k] = items[k].scene;
            vs[k] = ViewSpec{items[k].width, items[k].height, items[k].spp, 0, 0, 0, 0};
            b->max_spp = std::max(b->max_spp, items[k].spp);
            b->samples_per_render += (int64_t)items[k].width * items[k].height * std::max(0, items[k].spp);
        }
        b->pt.reset(new c2_pt());
        pt_build(b->pt.get(), descs.data(), vs.data(), n, opts, true);
        enqueue_film_reset(b->pt.get(), b->pt->stream);
        HIP_CHECK(hipStreamSynchronize(b->pt->stream));
        *out = b.release();
    });
}

void c2_batch_destroy(c2_batch* b) { delete b; }

int c2_batch_autotune(c2_batch* b, c2_batch_tuning* out) {
    return guard([&] {
        if (!b) fail(C2_ERR_INVALID, "null batch");
        c2_pt* pt = b->pt.get();
        hipStream_t s = pt->stream;
        hipEvent_t e0, e1;
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        const uint32_t saveProf = pt->prof_mask;
        pt->prof_mask = 0;
        auto timed = [&](auto&& enqueue) {
            HIP_CHECK(hipEventRecord(e0, s));
            enqueue();
            HIP_CHECK(hipEventRecord(e1, s));
            HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0;
            HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            return ms;
        };
        int tried = 0;
        // 1. the wavefront: the first sample passes (up to 16) over paths in flight x
        //    queue-kernel grid caps. More than pass 0: larger buffers let the later passes,
        //    which hold fewer views, share launches (up to kBatchSamplesPerChunk samples), which
        //    pass 0 alone does not show (on the 64-scene batch pass 0 ranked 12 M and 24 M
        //    paths in flight within noise, whole renders differ by 15%).
        const int32_t win = std::max(1, std::min(b->max_spp, 16));
        std::vector<int32_t> chunks;
        for (int32_t c = pt->chunk_cap; c >= (1 << 16) && chunks.size() < 3; c /= 2) chunks.push_back(c);
        if (chunks.empty()) chunks.push_back(pt->chunk_cap);
        const int32_t qbs[] = {4096, 8192, 16384};
        float bestW = INFINITY;
        int32_t bestC = pt->chunk, bestQ = pt->opts.queue_blocks;
        enqueue_passes(pt, 0, win, s, true);  // warm-up: code objects, caches
        for (int32_t c : chunks)
            for (int32_t q : qbs) {
                pt->chunk = c;
                pt->opts.queue_blocks = q;
                const float ms = timed([&] { enqueue_passes(pt, 0, win, s, true); });
                tried++;
                if (ms < bestW) {
                    bestW = ms;
                    bestC = c;
                    bestQ = q;
                }
            }
        pt->chunk = bestC;
        pt->opts.queue_blocks = bestQ;
        // ... and the BVH width at those values (the 8-wide tree exists for batches of BVH scenes)
        const bool have8 = pt->accel->scene.wide8 && !pt->accel->scene.kd;
        float ms8 = -1.f;
        if (have8) {
            const bool save8 = pt->use8;
            pt->use8 = !save8;
            enqueue_passes(pt, 0, 1, s, true);  // warm-up of the other width's kernels
            const float other = timed([&] { enqueue_passes(pt, 0, win, s, true); });
            tried++;
            ms8 = save8 ? bestW : other;
            if (other < bestW) bestW = other;
            else pt->use8 = save8;
        }
        // 2. the persistent kernel: sample 0 of every pixel over the block sizes
        const bool pok = persistent_ok(pt);
        float bestP = INFINITY;
        int32_t bestB = 256, bestWv = 0;
        if (pok) {
            for (const PersistShape& sh : kPersistShapes) {
                enqueue_persistent(pt, 0, 1, s, sh.block, sh.waves);  // warm-up
                const float ms = timed([&] { enqueue_persistent(pt, 0, 1, s, sh.block, sh.waves); });
                tried++;
                if (ms < bestP) {
                    bestP = ms;
                    bestB = sh.block;
                    bestWv = sh.waves;
                }
            }
        }
        // 3. the two winners over the whole render, each as the captured graph that
        //    c2_batch_render(b, 1) replays (captured, replayed once to warm, then timed)
        auto graph_ms = [&](int32_t sched) {
            b->schedule = sched;
            b->block = bestB;
            b->waves = bestWv;
            b->drop_graph();
            int rc = c2_batch_render(b, 1);
            if (rc != C2_OK) fail(rc, c2_last_error());
            HIP_CHECK(hipStreamSynchronize(s));
            return timed([&] { HIP_CHECK(hipGraphLaunch(b->exec, s)); });
        };
        b->renders = 0;
        const float wMs = graph_ms(C2_SCHED_WAVEFRONT);
        const float pMs = pok ? graph_ms(C2_SCHED_PERSISTENT) : INFINITY;
        b->schedule = pMs < wMs ? C2_SCHED_PERSISTENT : C2_SCHED_WAVEFRONT;
        b->block = bestB;
        b->waves = bestWv;
        b->renders = 0;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        pt->prof_mask = saveProf;
        pt->samples = 0;
        b->drop_graph();  // a captured graph holds the old launch shapes
        enqueue_film_reset(pt, s);
        HIP_CHECK(hipStreamSynchronize(s));
        if (out) {
            out->paths_in_flight = bestC;
            out->queue_blocks = bestQ;
            out->pass_ms = bestW;
            out->candidates = tried;
            out->schedule = b->schedule;
            out->block_threads = bestB;
            out->waves_per_simd = bestWv;
            out->sample0_ms_persistent = pok ? bestP : -1.0;
            out->render_ms_wavefront = wMs;
            out->render_ms_persistent = pok ? pMs : -1.0;
            out->bvh_width = pt->use8 ? 8 : 4;
            out->pass_ms_width8 = ms8;
        }
    });
}

int c2_batch_configure(c2_batch* b, const c2_batch_tuning* cfg) {
    return guard([&] {
        if (!b || !cfg) fail(C2_ERR_INVALID, "null argument");
        c2_pt* pt = b->pt.get();
        if (cfg->schedule != C2_SCHED_WAVEFRONT && cfg->schedule != C2_SCHED_PERSISTENT)
            fail(C2_ERR_INVALID, "schedule: C2_SCHED_WAVEFRONT or C2_SCHED_PERSISTENT");
        if (cfg->schedule == C2_SCHED_PERSISTENT) {
            if (!persist_shape(cfg->block_threads, cfg->waves_per_simd))
                fail(C2_ERR_INVALID, "persistent launch shape: 128, 256 or 512 threads x 0, 4, 6 or 8 waves per SIMD");
            if (!persistent_ok(pt))
                fail(C2_ERR_UNSUPPORTED, "persistent kernel: needs the BVH, no radiance cache");
        }
        if (cfg->paths_in_flight < 0 || cfg->queue_blocks < 0) fail(C2_ERR_INVALID, "negative launch shape");
        if (cfg->bvh_width != 0 && cfg->bvh_width != 4 && cfg->bvh_width != 8)
            fail(C2_ERR_INVALID, "BVH width: 4 or 8 (0 keeps the current one)");
        if (cfg->bvh_width == 8 && (!pt->accel->scene.wide8 || pt->accel->scene.kd))
            fail(C2_ERR_UNSUPPORTED, "BVH width 8: the batch has no 8-wide tree");
        b->schedule = cfg->schedule;
        if (cfg->bvh_width != 0) pt->use8 = cfg->bvh_width == 8;
        if (cfg->schedule == C2_SCHED_PERSISTENT) {
            b->block = cfg->block_threads;
            b->waves = cfg->waves_per_simd;
        }
        if (cfg->paths_in_flight > 0) pt->chunk = std::min(cfg->paths_in_flight, pt->chunk_cap);
        if (cfg->queue_blocks > 0) pt->opts.queue_blocks = cfg->queue_blocks;
        b->drop_graph();
    });
}

int c2_batch_render(c2_batch* b, int32_t use_graph) {
    return guard([&] {
        if (!b) fail(C2_ERR_INVALID, "null batch");
        c2_pt* pt = b->pt.get();
        hipStream_t s = pt->stream;
        if (!use_graph) {
            enqueue_batch(b, s);
        } else {
            if (!b->exec) {
                const uint32_t saveProf = pt->prof_mask;
                pt->prof_mask = 0;  // no timing events inside the graph
                HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
                try {
                    enqueue_batch(b, s);
                } catch (...) {
                    hipGraph_t g = nullptr;
                    (void)hipStreamEndCapture(s, &g);
                    if (g) (void)hipGraphDestroy(g);
                    pt->prof_mask = saveProf;
                    throw;
                }
                pt->prof_mask = saveProf;
                HIP_CHECK(hipStreamEndCapture(s, &b->graph));
                HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
            }
            HIP_CHECK(hipGraphLaunch(b->exec, s));
        }
        b->renders++;
    });
}

int c2_batch_sync(c2_batch* b) {
    return guard([&] {
        if (!b) fail(C2_ERR_INVALID, "null batch");
        HIP_CHECK(hipStreamSynchronize(b->pt->stream));
    });
}

int c2_batch_get_film(c2_batch* b, int32_t item, float* rgb) {
    return guard([&] {
        if (!b || !rgb) fail(C2_ERR_INVALID, "null argument");
        if (item < 0 || item >= (int32_t)b->pt->views.size()) fail(C2_ERR_INVALID, "item out of range");
        read_film(b->pt.get(), item, rgb);
    });
}

int c2_batch_get_stats(const c2_batch* b, c2_batch_stats* out) {
    return guard([&] {
        if (!b || !out) fail(C2_ERR_INVALID, "null argument");
        unsigned long long t[4] = {0, 0, 0, 0};
        HIP_CHECK(hipMemcpy(t, b->pt->B.totals, sizeof(t), hipMemcpyDeviceToHost));
        out->samples = b->renders * b->samples_per_render;
        out->samples_per_render = b->samples_per_render;
        out->extension_rays = (int64_t)t[0];
        out->shadow_rays = (int64_t)t[1];
        out->mis_rays = (int64_t)t[2];
        out->passes = b->max_spp;
        out->paths_in_flight = b->pt->chunk;
        out->queue_blocks = b->pt->opts.queue_blocks;
        out->schedule = b->schedule;
        out->block_threads = b->block;
        out->waves_per_simd = b->waves;
        out->graph_nodes = 0;
        if (b->graph) {
            size_t nn = 0;
            HIP_CHECK(hipGraphGetNodes(b->graph, nullptr, &nn));
            out->graph_nodes = (int64_t)nn;
        }
    });
}

}  // extern "C"
