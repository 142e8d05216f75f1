/*
 * crystal2.h -- C-ABI of the MI355X (gfx950) wavefront path tracer.
 *
 * Drop-in boundary for Crystal2's PathTracer + MIS hot path. Every entry point
 * is `extern "C"`, takes plain pointers and sizes, returns an int status
 * (C2_OK = 0, negative on error) and never lets a C++ exception cross the
 * boundary; c2_last_error() returns a thread-local message for the last
 * failure on the calling thread.
 *
 * Reference interfaces replaced (paths relative to the Crystal2 repository):
 *   Scene/config loaders   src/Core/Scene.cpp:23-199, src/Loaders/{JsonLoader,ConfigLoader,ObjLoader}.cpp,
 *                          src/Cameras/Camera.cpp:7-20, src/Integrators/Integrator.cpp:14-47
 *   Accelerator            src/Accelerators/Accelerator.h:8-22  (Build / Intersect / IntersectTest)
 *   Integrator             src/Integrators/Integrator.h:8-18     (Render(Scene*, Camera*, FrameBuffer*)),
 *                          src/Integrators/SamplerIntegrator.cpp:17-72,
 *                          src/Integrators/PathTracingIntegrator.cpp:28-278
 *   Sampler                src/Samplers/HaltonSampler.cpp:118-291 (index + value per dimension)
 *   FrameBuffer            src/Core/FrameBuffer.cpp:20-75         (AddSample / GetImageDataRGB32F)
 *
 * Ownership: descriptors passed in are read during the call only; handles own
 * their device memory. A handle is not internally synchronised: use one per
 * host thread (the reference calls Intersect concurrently from its tile pool,
 * src/Integrators/SamplerIntegrator.cpp:67; here one handle per thread plays
 * that role).
 */
#ifndef CRYSTAL2_H
#define CRYSTAL2_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status codes */
enum {
    C2_OK = 0,
    C2_ERR_INVALID = -1,     /* bad argument / inconsistent descriptor            */
    C2_ERR_PARSE = -2,       /* JSON/OBJ syntax or schema error (SJson semantics)  */
    C2_ERR_IO = -3,          /* file could not be opened                          */
    C2_ERR_DEVICE = -4,      /* HIP runtime error (message names the call)        */
    C2_ERR_NOMEM = -5,       /* host or device allocation failed                  */
    C2_ERR_UNSUPPORTED = -6  /* valid reference input this build does not handle  */
};

/* Thread-local message describing the last error on this thread ("" if none). */
const char* c2_last_error(void);

/* ---------------------------------------------------------------- scene description
 * A flattened, already-assembled scene: what Scene::CreateScene builds
 * (src/Core/Scene.cpp:94-199). Entity order is JSON object order; the entity
 * index is the primitive ID everywhere (hits, ignore lists, light attachment).
 */
enum { C2_SHAPE_TRIANGLE = 0, C2_SHAPE_SPHERE = 1, C2_SHAPE_DISK = 2 };

enum {
    C2_MAT_DIFFUSE = 0,  /* Lambertain                     src/Materials/DiffuseMaterial.cpp:16-20 */
    C2_MAT_MIRROR = 1,   /* SpecularReflection + NoOp       src/Materials/MirrorMaterial.cpp:11-13  */
    C2_MAT_GLASS = 2,    /* SpecularFresnel | MicrofacetBlend src/Materials/Glass.cpp:16-34         */
    C2_MAT_MICRO = 3,    /* MicrofacetReflection + Schlick  src/Materials/MicrofacetMaterial.cpp:18-42 */
    C2_MAT_PHONG = 4,    /* BlinnPhongReflection            src/Materials/PhongMaterial.cpp:13-22    */
    C2_MAT_PLASTIC = 5,  /* FresnelBlend                    src/Materials/Plastic.cpp:13-27          */
    C2_MAT_ORENNAYAR = 6 /* OrenNayarReflection             src/Materials/DiffuseMaterial.cpp:29-34  */
};
enum { C2_NDF_GGX = 0, C2_NDF_BECKMANN = 1 };
enum { C2_LIGHT_AREA = 0, C2_LIGHT_POINT = 1 };
/* Accelerator::GetAccelerator names (src/Accelerators/Accelerator.cpp:8-23): "BVH" and
 * "GPUBVH" run the GPU BVH; "KDTree" runs the reference's KD-tree (the same tree, traversed
 * on the GPU with the reference's windows, so results match it rather than the BVH). */
/* C2_ACCEL_BVH8: the same SAH tree collapsed 8 wide (256-byte nodes) instead of 4 wide; the
 * results are those of C2_ACCEL_BVH (closest hits tie to the lowest entity index). */
enum { C2_ACCEL_BVH = 0, C2_ACCEL_KDTREE = 1, C2_ACCEL_BVH8 = 2 };

typedef struct c2_material {
    int32_t type;          /* C2_MAT_*                                      */
    int32_t ndf;           /* C2_NDF_* (Micro, Plastic)                     */
    int32_t phong_k;       /* Blinn-Phong exponent K                        */
    int32_t kd_texture;    /* Diffuse, Phong: 0 = the constant colour, k >= 1 = the
                              texture c2_scene_add_texture returned as k ("Kd",
                              Material.cpp:25-29,91-95)                      */
    float color[3];        /* albedo / R (constant texture); Plastic: Kd    */
    float roughness[2];    /* NDF alpha_x, alpha_y (Glass: 0,0 = specular)  */
    float ior;             /* IOR (Glass, Micro)                            */
    float ks[3];           /* Plastic: Ks                                   */
    float sigma;           /* OrenNayar: Sigma                              */
} c2_material;

typedef struct c2_light {
    int32_t type;          /* C2_LIGHT_*                                         */
    int32_t entity;        /* AREA: entity whose shape emits (one light per tri) */
    float radiance[3];     /* AREA: Lemit; POINT: Flux                          */
    float position[3];     /* POINT only                                        */
} c2_light;

typedef struct c2_camera {
    float position[3];
    float look_at[3];
    float up[3];
    float fov;             /* radians, FOV/180*pi as Camera::CreateCamera computes it */
} c2_camera;

/* An 8-bit RGB image as stbi_load(..., 3) returns it: rows top-down, 3 bytes per texel.
 * Texels are read nearest-neighbour with the reference's uv clamp and row flip, and
 * converted to linear as (byte/255)^2.2 (src/Textures/Texel.cpp:15-38). */
typedef struct c2_texture {
    int32_t width, height;
    const uint8_t* rgb;
} c2_texture;

typedef struct c2_scene_desc {
    int32_t num_entities;
    const int32_t* entity_shape_kind;   /* C2_SHAPE_*                                */
    const int32_t* entity_shape_index;  /* index into triangles[] or spheres[]       */
    const int32_t* entity_material;     /* index into materials[], -1 = transparent  */
    const int32_t* entity_light;        /* index into lights[] (attached area light), -1 = none */

    int32_t num_triangles;
    const float* triangles;   /* [num_triangles][3 vertices][8]: pos.xyz, normal.xyz, uv.xy,
                                 world space, normals already transformed + normalised
                                 (src/Shapes/TriangleMesh.cpp:29-51)                     */
    int32_t num_spheres;
    const float* spheres;     /* [num_spheres][22]: analytic shapes (spheres and disks):
                                 center.xyz, radius, local2world[9], world2local[9] (glm
                                 column-major, src/Shapes/Sphere.cpp:19-30, Disk.cpp:19-29) */
    int32_t num_materials;
    const c2_material* materials;
    int32_t num_lights;
    const c2_light* lights;   /* the reference's light order (Scene::_lights)       */
    c2_camera camera;
    int32_t num_textures;
    const c2_texture* textures;  /* texture id k is textures[k - 1]                    */
    int32_t skybox[6];        /* texture ids of the cubemap faces posx, negx, posy, negy,
                                 posz, negz; all 0 = no skybox (Scene.cpp:128-143)     */
} c2_scene_desc;

/* ---------------------------------------------------------------- host scene builder
 * Owns the arrays a c2_scene_desc points into. c2_scene_load reads the
 * reference's scene JSON (+ OBJ files, relative to base_dir or the CWD when
 * base_dir is NULL) with the reference's number parsing and object order.
 */
typedef struct c2_scene c2_scene;

typedef struct c2_config {
    int32_t width, height, spp, threads;
    int32_t max_depth;
    char accel[16];        /* "Acceleration": "BVH" | "Brute" | "KDTree" | "GPUBVH" */
    char integrator[32];   /* "PathTracer" | "WavefrontPathTracer" | ...           */
    char sampler[16];      /* "Halton" (the only sampler on this path)             */
} c2_config;

int  c2_config_load(const char* config_json_path, c2_config* out);
int  c2_scene_load(const char* scene_json_path, const char* base_dir, c2_scene** out);
int  c2_scene_create(c2_scene** out);
void c2_scene_destroy(c2_scene* scene);
/* returns material index in *index */
int  c2_scene_add_material(c2_scene* scene, const c2_material* material, int32_t* index);
/* Triangles with unshared per-corner vertex data [ntri][3][8] (pos, normal, uv),
 * transformed by the optional glm-style column-major 4x4 `transform` exactly as
 * TriangleMesh does (NULL = identity). material -1 = transparent; lemit != NULL
 * attaches one Diffused area light per triangle. */
int  c2_scene_add_triangles(c2_scene* scene, const float* vertex_data, int32_t ntri,
                            const float* transform, int32_t material, const float* lemit);
/* Sphere at center with radius and XYZ rotation in degrees (src/Shapes/Sphere.cpp:19-30). */
int  c2_scene_add_sphere(c2_scene* scene, const float center[3], float radius,
                         const float rotation_deg[3], int32_t material, const float* lemit);
/* Disk::CreateDisk (src/Shapes/Disk.cpp:7-29): an entity of kind C2_SHAPE_DISK whose
 * record sits in the spheres[] array. */
int  c2_scene_add_disk(c2_scene* scene, const float center[3], float radius,
                       const float rotation_deg[3], int32_t material, const float* lemit);
int  c2_scene_add_point_light(c2_scene* scene, const float position[3], const float flux[3]);
/* ImageTexels (src/Textures/Texel.cpp:4-38): copies width*height*3 bytes of RGB, rows
 * top-down; *id (>= 1) is what c2_material.kd_texture and c2_scene_set_skybox take. */
int  c2_scene_add_texture(c2_scene* scene, int32_t width, int32_t height, const uint8_t* rgb,
                          int32_t* id);
/* ImageCubemapTexture (src/Textures/CubemapTexture.cpp:3-14): texture ids of the faces
 * posx, negx, posy, negy, posz, negz (all > 0), or all 0 to remove the skybox. Rays that
 * leave the scene then add beta * skybox(direction) (PathTracingIntegrator.cpp:39-46). */
int  c2_scene_set_skybox(c2_scene* scene, const int32_t face_ids[6]);
/* stbi_load(path, &w, &h, &c, 3) as the reference calls it (Texel.cpp:5): PNG or baseline
 * JPEG to 8-bit RGB with stb_image's arithmetic. *rgb is malloc'd; free it with
 * c2_image_free. Progressive JPEG gives C2_ERR_UNSUPPORTED. */
int  c2_image_load(const char* path, int32_t* width, int32_t* height, uint8_t** rgb);
void c2_image_free(uint8_t* rgb);
int  c2_scene_set_camera(c2_scene* scene, const c2_camera* camera);
/* Fills *out with pointers into the scene's arrays (valid until the scene changes). */
int  c2_scene_get_desc(const c2_scene* scene, c2_scene_desc* out);
/* Binary scene cache (no counterpart in the reference, whose OBJ parsing is sscanf-bound,
 * src/Loaders/ObjLoader.cpp:90-149): every array of the scene in one file, reloaded at
 * file-read speed into a scene equal to the saved one. */
int  c2_scene_save(const c2_scene* scene, const char* path);
int  c2_scene_load_binary(const char* path, c2_scene** out);

/* ---------------------------------------------------------------- film output
 * rgb: width*height*3 floats, rows top-down, as c2_pt_get_film returns them.
 * c2_film_write_hdr writes the file stbi_write_hdr("result.hdr", w, h, 3, rgb) writes at
 * the end of the reference's render (src/Core/RayTracer.cpp:40), byte for byte.
 * c2_film_write_pfm writes the film losslessly as a little-endian PFM. */
int  c2_film_write_hdr(const char* path, int32_t width, int32_t height, const float* rgb);
int  c2_film_write_pfm(const char* path, int32_t width, int32_t height, const float* rgb);

/* ---------------------------------------------------------------- Accelerator (GPU BVH)
 * Build: SAH binary BVH (leaf <= 4) collapsed to a 4-wide BVH with 128-byte,
 * cache-line aligned nodes in HBM. Intersect / IntersectTest keep the
 * reference's semantics (src/Accelerators/BVH.cpp:56-154): per-primitive box
 * pre-clip, closest hit by strict t (ties: lowest entity index), any-hit over
 * the box-clipped [tmin, tmax], `ignore` entity skipped.
 *
 * Rays are [n][8] float32: o.xyz, d.xyz, tmin, tmax. Closest hit ignores tmin
 * and clips with tmax (the reference passes tMin=0, tMax=inf).
 * `*_device` variants take device pointers and a hipStream_t (NULL = default
 * stream) and return without synchronising; host variants copy in and out.
 */
typedef struct c2_accel c2_accel;

typedef struct c2_accel_stats {
    int64_t num_nodes_bvh2;       /* binary SAH BVH nodes                       */
    int64_t num_nodes_bvh4;       /* 4-wide nodes after collapse                */
    int32_t max_depth_bvh4;
    int32_t leaf_size;
    double  build_seconds;
    int64_t device_bytes;         /* node + primitive bytes resident in HBM     */
} c2_accel_stats;

int  c2_accel_build(const c2_scene_desc* scene, c2_accel** out);  /* c2_accel_create(scene, C2_ACCEL_BVH) */
/* KDTree::Build / Intersect / IntersectTest (src/Accelerators/KDTree.cpp:59-266) when accel
 * is C2_ACCEL_KDTREE. */
int  c2_accel_create(const c2_scene_desc* scene, int32_t accel, c2_accel** out);
void c2_accel_destroy(c2_accel* accel);
int  c2_accel_get_stats(const c2_accel* accel, c2_accel_stats* out);
int  c2_accel_intersect(c2_accel* accel, const float* rays, int64_t n,
                        int32_t* prim, float* t);
int  c2_accel_intersect_test(c2_accel* accel, const float* rays, const int32_t* ignore,
                             int64_t n, uint8_t* hit);
int  c2_accel_intersect_device(c2_accel* accel, const float* rays, int64_t n,
                               int32_t* prim, float* t, void* stream);
int  c2_accel_intersect_test_device(c2_accel* accel, const float* rays, const int32_t* ignore,
                                    int64_t n, uint8_t* hit, void* stream);
/* c2_accel_intersect without any device allocation. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (serving loops that cast host
 * rays repeatedly, code that forbids hipMalloc). Same arguments and results as
 * c2_accel_intersect; the rays and results are staged in one fixed device array of
 * the library (40 MiB on each GPU it runs on, held until the process ends), so
 * n <= C2_NOALLOC_MAX_RAYS (larger n: C2_ERR_INVALID). Synchronous, on the device
 * current when called (the accelerator's). Calls that overlap on one GPU (several
 * threads) share that array and race: serialise them. */
#define C2_NOALLOC_MAX_RAYS (1 << 20)
int  c2_accel_intersect_noalloc(c2_accel* accel, const float* rays, int64_t n,
                                int32_t* prim, float* t);
/* Algorithmic traversal counts on the host binary SAH BVH (32-byte nodes,
 * leaf <= 4), the figure behind "algorithmic bytes per ray":
 *   bytes/ray = 24 + 8 + nodes*32 + tris*36. any_hit selects IntersectTest. */
int  c2_accel_count_visits(const c2_accel* accel, const float* rays, int64_t n, int32_t any_hit,
                           double* mean_nodes, double* mean_tris);

/* ---------------------------------------------------------------- Integrator (wavefront PT + MIS)
 * One renderer per scene. c2_pt_render adds Halton samples
 * [sample_begin, sample_begin + sample_count) of every pixel to the device
 * film in sample order (bit-for-bit the reference's per-pixel sequential
 * FrameBuffer::AddSample). One call with sample_begin=0, sample_count=spp is
 * the reference's Render(); smaller calls are the bench's "steps".
 */
typedef struct c2_pt c2_pt;

enum { C2_ARGORDER_RTL = 0, C2_ARGORDER_LTR = 1 };
enum { C2_FILM_F32 = 0, C2_FILM_F32_KAHAN = 1, C2_FILM_F64 = 2 };
/* PathTracingIntegrator (PathTracingIntegrator.cpp:28-278) or DirectLightingIntegrator
 * (DirectLightingIntegrator.cpp:22-156, transparent chains limited to 7 surfaces). */
enum { C2_INTEGRATOR_PATH = 0, C2_INTEGRATOR_DIRECT = 1 };

typedef struct c2_render_opts {
    int32_t width, height;
    int32_t spp;              /* Halton sampler sample count (index stride)           */
    int32_t max_depth;        /* PathTracer "Depth"                                    */
    int32_t arg_order;        /* C2_ARGORDER_*: order of Get1D/Get2D at
                                 PathTracingIntegrator.cpp:79 (RTL = GCC/MSVC)       */
    int32_t film_mode;        /* C2_FILM_*                                             */
    int32_t paths_in_flight;  /* paths per launch (path buffers of one lane); 0 = six
                                 samples of every pixel: a launch then carries four
                                 consecutive samples (frames under 2^18 pixels up to
                                 16, about 2^20 paths), up to six when that ends the
                                 call's launches on a multiple of the lanes; below one
                                 pass, a pass is split into chunks                   */
    int32_t queue_blocks;     /* grid cap (blocks) of the per-bounce queue kernels,
                                 which loop grid-stride; 0 = the library default     */
    int32_t win_x0, win_y0;   /* pixel window (film coordinates, y = 0 at the bottom); */
    int32_t win_w, win_h;     /* win_w = 0: the whole film                             */
    int32_t integrator;       /* C2_INTEGRATOR_*: "PathTracer" or "DirectLight"
                                 (src/Integrators/Integrator.cpp:27-40)               */
    int32_t accel;            /* C2_ACCEL_*: config "Acceleration" (0 = the BVH)         */
} c2_render_opts;

typedef struct c2_pt_stats {
    int64_t samples;          /* camera paths evaluated                               */
    int64_t extension_rays;   /* closest-hit path rays                                 */
    int64_t shadow_rays;      /* any-hit NEE rays                                      */
    int64_t mis_rays;         /* closest-hit BSDF-sampled MIS rays                     */
    int64_t kernel_launches;
} c2_pt_stats;

int  c2_pt_create(const c2_scene_desc* scene, const c2_render_opts* opts, c2_pt** out);
void c2_pt_destroy(c2_pt* pt);
int  c2_pt_reset(c2_pt* pt);
/* Enqueues samples [sample_begin, + sample_count) on `stream` (NULL: the renderer's own).
 * Consecutive chunks of one call alternate between two sets of path buffers on two
 * internal streams; the call's work starts after what `stream` already holds, and
 * `stream` waits for all of it at the end. Films are the same for any chunking. */
int  c2_pt_render(c2_pt* pt, int32_t sample_begin, int32_t sample_count, void* stream);
/* Film as FrameBuffer::GetImageDataRGB32F: W*H*3 float, rows top-down (host memory). */
int  c2_pt_get_film(c2_pt* pt, float* rgb);
int  c2_pt_get_stats(const c2_pt* pt, c2_pt_stats* out);
/* Synchronise the renderer's stream; returns the first asynchronous error. */
int  c2_pt_sync(c2_pt* pt);

/* Per-kernel-class timing with HIP events recorded on the launching stream around
 * every launch while profiling is enabled (adds no synchronisation). With two lanes the
 * times are wall times while the other lane's kernels share the GPU. C2_K_SHADE also
 * covers the shading-class grouping kernels and k_surface. */
enum { C2_K_GENERATE = 0, C2_K_EXTEND = 1, C2_K_SHADE = 2, C2_K_OCCLUDE = 3, C2_K_MIS = 4,
       C2_K_RESOLVE = 5, C2_K_ACCUMULATE = 6, C2_K_NEE = 7, C2_K_COUNT = 8 };
typedef struct c2_kernel_times {
    double  ms[C2_K_COUNT];        /* summed kernel time per class            */
    int64_t launches[C2_K_COUNT];  /* launches per class                      */
} c2_kernel_times;
int  c2_pt_set_profiling(c2_pt* pt, int32_t enable);   /* clears the record  */
/* Times only the kernel classes whose bit (1 << C2_K_*) is set (0 = none); clears the
 * record. No counterpart in the reference: it keeps the event overhead of measuring one
 * kernel inside a timed region small. */
int  c2_pt_profile_kernels(c2_pt* pt, uint32_t class_mask);
int  c2_pt_get_kernel_times(c2_pt* pt, c2_kernel_times* out);  /* synchronises */

/* Renders Halton sample `sample` (adding it to the film) and copies up to n rays
 * of one queue at bounce `bounce`: kind 0 = extension (path) rays, kind 1 =
 * shadow rays (with the entity they ignore). rays_out: [n][8] (o, d, tmin, tmax);
 * ignore_out may be NULL. *written = rays copied. */
int  c2_pt_capture_rays(c2_pt* pt, int32_t sample, int32_t bounce, int32_t kind, float* rays_out,
                        int32_t* ignore_out, int64_t n, int64_t* written);

/* ---------------------------------------------------------------- Radiance cache (new, no reference counterpart)
 * An online-trained MLP (64-dim input encoding, 4 hidden layers x 64 ReLU, 3
 * outputs) with FP16 MFMA forward / backward and FP32 Adam. A query is 16 floats:
 * position, outgoing direction, facing normal, roughness, albedo, then 3 slots the
 * library fills itself (the azimuths of the direction and the normal; values passed
 * there are ignored); the prediction is the radiance scattered from that point along
 * the direction.
 * Parameters: [5 layers][64 in][64 out] weights then [5][64] biases (the output
 * layer uses 3 of its 64 columns). */
typedef struct c2_nrc c2_nrc;

typedef struct c2_nrc_opts {
    float learning_rate;      /* Adam step size (1e-2)                    */
    float beta1, beta2, epsilon;
    int32_t seed;             /* weight initialisation                    */
    int32_t batch_capacity;   /* training rows per step (2^16)            */
    float bbox_min[3], bbox_max[3];  /* position normalisation            */
} c2_nrc_opts;

typedef struct c2_nrc_stats {
    int64_t steps;
    double  flop_forward_per_query;   /* 33,152 */
    int64_t queries;                  /* inference queries run in the renderer */
    int64_t train_rows;               /* training rows consumed by Adam steps  */
} c2_nrc_stats;

#define C2_NRC_PARAMS (5 * 64 * 64 + 5 * 64)

int  c2_nrc_create(const c2_nrc_opts* opts, c2_nrc** out);
void c2_nrc_destroy(c2_nrc* nrc);
int  c2_nrc_get_params(c2_nrc* nrc, float* params);        /* C2_NRC_PARAMS floats */
int  c2_nrc_set_params(c2_nrc* nrc, const float* params);
/* Host convenience: features [n][16] -> radiance [n][3]. */
int  c2_nrc_infer(c2_nrc* nrc, const float* features, int64_t n, float* out);
/* One training step on host data: features [n][16], targets [n][3]. Writes the
 * mean relative-L2 loss gradient (C2_NRC_PARAMS floats) and loss; apply != 0
 * also takes the Adam step. */
int  c2_nrc_train_step(c2_nrc* nrc, const float* features, const float* targets, int64_t n, int32_t apply,
                       float* grad_out, float* loss_out);
int  c2_nrc_get_stats(const c2_nrc* nrc, c2_nrc_stats* out);

/* Cache mode for a renderer: camera paths end at bounce `cache_bounce` with the
 * cache's prediction; one pixel in `train_ratio` per pass traces a full training
 * path whose vertices become training rows (target = radiance after the vertex /
 * throughput), and one Adam step runs per sample pass. nrc == NULL disables it. */
int  c2_pt_set_cache(c2_pt* pt, c2_nrc* nrc, int32_t cache_bounce, int32_t train_ratio);
/* Time of the cache's kernels since the last call, in ms (inference; training steps;
 * gathering the training rows, which runs no MFMA work), and the rows / queries they
 * processed (for MFMA utilisation). */
typedef struct c2_cache_times {
    double infer_ms, train_ms;
    int64_t queries, train_rows;
    double collect_ms;
} c2_cache_times;
int  c2_pt_get_cache_times(c2_pt* pt, c2_cache_times* out);

/* ---------------------------------------------------------------- Batched scenes
 * No counterpart in the reference, which renders one scene per process
 * (src/Core/RayTracer.cpp:11-22). Many scenes, each with its own film size and
 * sample count, render together: sample pass k traces every pixel of every scene
 * with spp > k in shared launches. Each scene's film is bit-identical to its own
 * c2_pt_render(0, spp) with the same options. The whole render can be captured
 * once into a hipGraph and replayed with one launch (use_graph = 1).
 * opts: max_depth, arg_order, film_mode, paths_in_flight, queue_blocks apply to
 * every scene; width / height / spp / the pixel window come from the items. */
typedef struct c2_batch c2_batch;
typedef struct c2_batch_item {
    const c2_scene_desc* scene;
    int32_t width, height, spp;
} c2_batch_item;
/* Schedules of a batch render. WAVEFRONT: every sample pass as the queue kernels
 * (generate, extend, shade, nee, occlude, mis, resolve, accumulate per bounce).
 * PERSISTENT: one kernel launch whose lanes take pixels from a device counter and
 * trace each sample's path from the camera to its end in registers (BVH accelerator,
 * path tracer or direct lighting). Both give the same films bit for bit. */
#define C2_SCHED_WAVEFRONT  0
#define C2_SCHED_PERSISTENT 1
typedef struct c2_batch_tuning {
    int32_t paths_in_flight;  /* wavefront: paths per launch                     */
    int32_t queue_blocks;     /* wavefront: grid cap of the queue kernels        */
    double pass_ms;           /* wavefront: sample passes 0 .. min(spp, 16) - 1  */
                              /* with the chosen values                          */
    int32_t candidates;       /* configurations timed                            */
    int32_t schedule;         /* C2_SCHED_*                                      */
    int32_t block_threads;    /* persistent: threads per block (128, 256, 512)   */
    int32_t waves_per_simd;   /* persistent: registers limited so that at least  */
                              /* this many waves fit a SIMD (4, 6, 8), or 0: none */
    double sample0_ms_persistent;  /* persistent: sample 0 of every pixel (-1: n/a) */
    double render_ms_wavefront;    /* whole render, captured graph, each schedule's */
    double render_ms_persistent;   /* best values (-1: persistent not applicable)   */
    int32_t bvh_width;        /* wavefront: BVH width traversed, 4 or 8 (the     */
                              /* persistent kernel traverses the 4-wide tree);   */
                              /* configure: 0 keeps the current width            */
    double pass_ms_width8;    /* wavefront: the same passes with the 8-wide tree  */
                              /* at the chosen values (-1: not timed)            */
} c2_batch_tuning;
typedef struct c2_batch_stats {
    int64_t samples, samples_per_render;
    int64_t extension_rays, shadow_rays, mis_rays;
    int32_t passes, paths_in_flight, queue_blocks;
    int64_t graph_nodes;      /* nodes of the captured graph (0 before capture)  */
    int32_t schedule, block_threads, waves_per_simd;
} c2_batch_stats;
int  c2_batch_create(const c2_batch_item* items, int32_t n, const c2_render_opts* opts, c2_batch** out);
void c2_batch_destroy(c2_batch* batch);
/* Tunes each schedule (wavefront: sample passes 0 .. min(spp, 16) - 1 over
 * paths-in-flight x queue_blocks, then the BVH width 4 or 8 at the best of those;
 * persistent: sample 0 over 128 / 256 / 512 threads per block x 0 / 4 / 6 / 8 waves per
 * SIMD), then times both winners over the whole render as captured graphs and keeps the
 * faster. Drops a captured graph and clears the films. */
int  c2_batch_autotune(c2_batch* batch, c2_batch_tuning* out);
/* Sets the schedule and launch shape explicitly (fields of c2_batch_tuning: schedule,
 * block_threads and waves_per_simd, paths_in_flight / queue_blocks when > 0, and
 * bvh_width when > 0; the timing fields are ignored). Drops a captured graph. */
int  c2_batch_configure(c2_batch* batch, const c2_batch_tuning* cfg);
/* Clears the films and renders every sample of every scene; asynchronous. */
int  c2_batch_render(c2_batch* batch, int32_t use_graph);
int  c2_batch_sync(c2_batch* batch);
int  c2_batch_get_film(c2_batch* batch, int32_t item, float* rgb);  /* item's W*H*3, rows top-down */
int  c2_batch_get_stats(const c2_batch* batch, c2_batch_stats* out);

/* ---------------------------------------------------------------- Sampler (Halton)
 * Reference HaltonSampler for a W x H film: vector index of (pixel, sample) and
 * sample value per dimension (dims >= 100 use the documented deterministic
 * substitute instead of the reference's shared mt19937). Host computation,
 * identical to the device code's tables. */
int  c2_halton_index(int32_t width, int32_t height, int32_t x, int32_t y, int32_t sample, int32_t* index);
int  c2_halton_value(int32_t width, int32_t height, int32_t index, int32_t dim, float* value);

/* ---------------------------------------------------------------- Utility
 * No counterpart in the reference. A plain streaming copy over `bytes` bytes
 * (read once, written once per launch; multiple of 16, at most 8 GiB) launched
 * `reps` times after one warm-up launch, on the default stream; *avg_ms = mean
 * launch time. It is the known-size kernel that calibrates HBM byte counters
 * (FETCH_SIZE / WRITE_SIZE) and the copy bandwidth the renderer's kernels are
 * compared against. Allocates and frees its own buffers. */
int  c2_util_stream_copy(int64_t bytes, int32_t reps, double* avg_ms);

#ifdef __cplusplus
}
#endif
#endif /* CRYSTAL2_H */
