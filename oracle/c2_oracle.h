/*
 * c2_oracle.h -- CPU restatement of Crystal2's PathTracer + MIS path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the GPU path: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. The
 * product library never links, loads or calls it.
 *
 * It consumes the same flattened scene description as the product
 * (include/crystal2.h, c2_scene_desc) and restates, function by function, the
 * reference's arithmetic (citations at each function in c2_oracle.cpp).
 *
 * Build variants (oracle/Makefile):
 *   libc2oracle.so       transcendentals evaluated in double and rounded to
 *                        float (the convention of the GPU path);
 *   libc2oracle_libm.so  transcendentals from the float libm functions the
 *                        reference calls (std::sin(float), ...), for parity
 *                        against the reference binary oracle/_ref/crystal2_ref.
 */
#ifndef C2_ORACLE_H
#define C2_ORACLE_H
#include <stdint.h>
#include "crystal2.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene orc_scene;

enum { ORC_ACCEL_BVH = 0, ORC_ACCEL_BRUTE = 1, ORC_ACCEL_KDTREE = 2 };
/* Closest-hit tie rule: ORC_TIE_FIRST = the reference's "first found in
 * traversal order wins" (strict <); ORC_TIE_LOWEST = among equal t, the lowest
 * entity index wins (Brute's rule, src/Accelerators/Brute.cpp:19-40). */
enum { ORC_TIE_FIRST = 0, ORC_TIE_LOWEST = 1 };

int  orc_scene_create(const c2_scene_desc* desc, int accel, int tie_rule, orc_scene** out);
void orc_scene_destroy(orc_scene* s);
/* Integrator of orc_render: ORC_INTEGRATOR_PATH (PathTracingIntegrator, the default)
 * or ORC_INTEGRATOR_DIRECT (DirectLightingIntegrator). */
enum { ORC_INTEGRATOR_PATH = 0, ORC_INTEGRATOR_DIRECT = 1 };
int  orc_scene_set_integrator(orc_scene* s, int integrator);
/* enable != 0: Halton dims >= 100 and Blinn-Phong redraws come from the
 * reference's own two default-seeded std::mt19937 generators, consumed in the
 * order of a one-thread reference render (orc_render with threads = 1 only;
 * other thread counts fail). 0 (the default): the hash streams the GPU uses.
 * skip > 0 discards that many draws of each generator first: the render then
 * stands for a reference render from another point of the same streams. */
int  orc_scene_set_reference_draws(orc_scene* s, int enable, uint64_t skip);
const char* orc_last_error(void);

/* Render film rows [y_begin, y_end) (film coordinates, y = 0 at the bottom)
 * with W x H / spp / depth. film: W*H*3 floats laid out like
 * FrameBuffer::GetImageDataRGB32F (rows top-down); rows outside the range are
 * left untouched. arg_order: 0 = Get2D before Get1D at
 * PathTracingIntegrator.cpp:79 (GCC/MSVC), 1 = Get1D first (clang). */
int orc_render(orc_scene* s, int width, int height, int spp, int depth, int arg_order,
               int threads, int y_begin, int y_end, float* film, double* seconds);
/* Same, but returns the raw per-pixel radiance of every sample:
 * samples: [H][W][spp][3] in film coordinates (y = 0 bottom), for small images. */
int orc_render_samples(orc_scene* s, int width, int height, int spp, int depth, int arg_order,
                       int threads, float* samples);

/* rays [n][8] (o, d, tmin, tmax). Closest hit as Scene::Intersect (tmin=0,
 * tmax=inf): prim = entity index or -1, t. Any hit as
 * Scene::IntersectTest(ray, tmin, tmax, shape of entity ignore[i]). */
int orc_intersect(orc_scene* s, const float* rays, int64_t n, int32_t* prim, float* t);
int orc_intersect_test(orc_scene* s, const float* rays, const int32_t* ignore, int64_t n, uint8_t* hit);

/* HaltonSampler for a W x H film: per query (x, y, k): vector index and the
 * values of dims [0, ndims). */
int orc_halton(int width, int height, const int32_t* queries, int64_t n, int ndims,
               int32_t* index, float* values);

#ifdef __cplusplus
}
#endif
#endif
