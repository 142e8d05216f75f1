// Driver for the oracle under the compiler's sanitizers (tests/test_oracle_sanitizers.py):
// loads a scene with the host scene loader, renders it with the CPU restatement on several
// threads and writes the film as raw float32 [H][W][3] for the test to compare with the
// uninstrumented oracle's film.
//
//   oracle_san scene.json base_dir width height spp depth accel integrator threads out.raw
//
// accel: 0 BVH, 2 KD-tree; integrator: 0 path tracer, 1 direct lighting.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../oracle/c2_oracle.h"

int main(int argc, char** argv) {
    if (argc != 11) {
        std::fprintf(stderr, "usage: %s scene.json base_dir W H spp depth accel integrator threads out.raw\n", argv[0]);
        return 2;
    }
    const int W = std::atoi(argv[3]), H = std::atoi(argv[4]), spp = std::atoi(argv[5]), depth = std::atoi(argv[6]);
    const int accel = std::atoi(argv[7]), integ = std::atoi(argv[8]), threads = std::atoi(argv[9]);
    c2_scene* scene = nullptr;
    if (c2_scene_load(argv[1], argv[2], &scene) != C2_OK) {
        std::fprintf(stderr, "scene load failed: %s\n", c2_last_error());
        return 1;
    }
    c2_scene_desc desc;
    if (c2_scene_get_desc(scene, &desc) != C2_OK) return 1;
    orc_scene* o = nullptr;
    if (orc_scene_create(&desc, accel, ORC_TIE_FIRST, &o) != 0 || orc_scene_set_integrator(o, integ) != 0) {
        std::fprintf(stderr, "oracle scene failed: %s\n", orc_last_error());
        return 1;
    }
    std::vector<float> film((size_t)W * H * 3, 0.f);
    double secs = 0;
    if (orc_render(o, W, H, spp, depth, 0, threads, 0, H, film.data(), &secs) != 0) {
        std::fprintf(stderr, "render failed: %s\n", orc_last_error());
        return 1;
    }
    FILE* f = std::fopen(argv[10], "wb");
    if (!f || std::fwrite(film.data(), sizeof(float), film.size(), f) != film.size()) return 1;
    std::fclose(f);
    orc_scene_destroy(o);
    c2_scene_destroy(scene);
    std::printf("rendered %dx%d spp %d depth %d on %d threads in %.2fs\n", W, H, spp, depth, threads, secs);
    return 0;
}
