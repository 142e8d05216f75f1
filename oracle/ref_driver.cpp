// Headless driver for the reference renderer compiled from /root/reference.
//
// Test infrastructure only (oracle/_ref). Built with -DC2_GPU_PLUGIN (make refgpu), the
// render and hits commands hand the scene file to the GPU plugins first, so a config may
// name "GPUBVH" and "WavefrontPathTracer". It replaces the reference's GUI
// entry point (Crystal2.cpp:8-37 + src/CrystalEngine.cpp) with a command line
// that calls the reference's own loaders, accelerator, sampler and integrator
// and dumps raw results, so the CPU restatement in oracle/ and the GPU path can
// be checked against the reference itself:
//
//   crystal2_ref render <config.json> <scene.json> <out.f32>
//       Scene::CreateScene + Camera::CreateCamera + Integrator::LoadIntegrator
//       (src/Core/RayTracer.cpp:11-22), Integrator::Render timed with
//       steady_clock, film dumped as FrameBuffer::GetImageDataRGB32F
//       (src/Core/FrameBuffer.cpp:60-75): W*H*3 float32, rows top-down.
//       Prints one JSON line {"render_s":..,"samples":..,"threads":..}.
//   crystal2_ref hits <config.json> <scene.json> <rays.bin> <out.bin>
//       rays.bin: int64 n, then n x 8 float32 (o.xyz, d.xyz, tmin, tmax).
//       out.bin : n int32 entity index (-1 = miss) from Scene::Intersect
//                 (src/Core/Scene.cpp:50-53), n float32 t, n float32 hit
//                 position xyz x3 ... see write order below, and n uint8 from
//                 Scene::IntersectTest(ray, tmin, tmax, nullptr).
//   crystal2_ref raycast <config.json> <scene.json> <rays.bin> <threads> <out.bin>
//       The ray-cast timing: Scene::Intersect over all rays, then
//       Scene::IntersectTest(ray, tmin, tmax, nullptr), each split into equal
//       contiguous ranges over <threads> std::threads (the reference calls both
//       concurrently from its tile pool) and timed with steady_clock.
//       out.bin: n int32 entity index (-1 = miss), n float32 t, n uint8 any hit.
//       Prints {"closest_s":..,"any_s":..,"rays":..,"threads":..,"hw_threads":..}.
//   crystal2_ref halton <W> <H> <spp> <queries.bin> <ndims> <out.bin>
//       queries.bin: int64 n, then n x 3 int32 (x, y, k). For each query:
//       StartPixel((x,y)); idx = GetVectorIndexForSample(k)
//       (src/Samplers/HaltonSampler.cpp:251-270); SampleValue(idx, d) for
//       d < ndims (:272-291). out.bin: n int32 idx, then n x ndims float32.
//   crystal2_ref writehdr <film.f32> <W> <H> <out.hdr>
//       stbi_write_hdr(out, W, H, 3, film) as RayTracer::TraceAsync calls it
//       (src/Core/RayTracer.cpp:40); film.f32 is W*H*3 float32, rows top-down.
//   crystal2_ref decode <image> <out.bin>
//       stbi_load(image, 3 components), the call ImageTexels::CreateImageTexels
//       makes (src/Textures/Texel.cpp:4-13). out.bin: int32 width, int32 height,
//       then width*height*3 bytes, rows top-down.
#include <Core/Scene.h>
#include <Core/FrameBuffer.h>
#include <Core/SurfaceInteraction.h>
#include <Core/Entities.h>
#include <Cameras/Camera.h>
#include <Integrators/Integrator.h>
#include <Loaders/JsonLoader.h>
#include <Loaders/ConfigLoader.h>
#include <Samplers/HaltonSampler.h>
#include <SJson/SJson.h>
#include <stbi/stb_image.h>
#include <stbi/stb_image_write.h>
#ifdef C2_GPU_PLUGIN
#include "../crystal2_amd/plugin/crystal2_gpu.h"
#endif

#include <chrono>
#include <stdexcept>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

struct Loaded {
    std::shared_ptr<SJson::SJsonNode> cfgNode, sceneNode;
    config::ConfigInfo info;
    std::unique_ptr<Scene> scene;
};

Loaded load(const char* cfgPath, const char* scenePath) {
    Loaded L;
    L.cfgNode = loader::JsonLoader::LoadJsonFile(cfgPath);
    L.sceneNode = loader::JsonLoader::LoadJsonFile(scenePath);
    L.info = config::ConfigLoader::LoadConfigInfo(L.cfgNode);
    L.scene = Scene::CreateScene(L.sceneNode->GetMember("Scene"), L.info);
    return L;
}

int cmdRender(int argc, char** argv) {
    if (argc < 5) return 2;
#ifdef C2_GPU_PLUGIN
    crystal2_gpu::SetSceneFile(argv[3]);  // before Scene::CreateScene builds a "GPUBVH"
#endif
    Loaded L = load(argv[2], argv[3]);
    auto camera = Camera::CreateCamera(L.sceneNode->GetMember("Camera"), L.info);
    auto integrator = crystal::Integrator::LoadIntegrator(L.cfgNode.get(), L.info);
    FrameBuffer fb(L.info.Width, L.info.Height);
    fb.Clear();
    auto t0 = std::chrono::steady_clock::now();
    integrator->Render(L.scene.get(), camera.get(), &fb);
    auto t1 = std::chrono::steady_clock::now();
    double secs = std::chrono::duration<double>(t1 - t0).count();
    auto data = fb.GetImageDataRGB32F();
    FILE* f = std::fopen(argv[4], "wb");
    if (!f) return 3;
    std::fwrite(data.get(), sizeof(float), (size_t)L.info.Width * L.info.Height * 3, f);
    std::fclose(f);
    long long samples = (long long)L.info.Width * L.info.Height * L.info.SamplesPerPixel;
    std::printf("{\"render_s\": %.6f, \"samples\": %lld, \"threads\": %d, \"hw_threads\": %u}\n",
                secs, samples, L.info.NumOfThreads, std::thread::hardware_concurrency());
    return 0;
}

int cmdHits(int argc, char** argv) {
    if (argc < 6) return 2;
#ifdef C2_GPU_PLUGIN
    crystal2_gpu::SetSceneFile(argv[3]);
#endif
    Loaded L = load(argv[2], argv[3]);
    auto objects = L.scene->GetObjects();
    std::unordered_map<const crystal::Entity*, int> index;
    for (size_t i = 0; i < objects.size(); i++) index[objects[i]] = (int)i;

    FILE* f = std::fopen(argv[4], "rb");
    if (!f) return 3;
    long long n = 0;
    if (std::fread(&n, sizeof(n), 1, f) != 1) return 3;
    std::vector<float> rays((size_t)n * 8);
    if (std::fread(rays.data(), sizeof(float), rays.size(), f) != rays.size()) return 3;
    std::fclose(f);

    std::vector<int> id((size_t)n, -1);
    std::vector<float> t((size_t)n, 0.f), pos((size_t)n * 3, 0.f), nrm((size_t)n * 3, 0.f);
    std::vector<unsigned char> any((size_t)n, 0);
    auto t0 = std::chrono::steady_clock::now();
    for (long long i = 0; i < n; i++) {
        const float* r = &rays[(size_t)i * 8];
        Ray3f ray(glm::vec3(r[0], r[1], r[2]), glm::vec3(r[3], r[4], r[5]));
        SurfaceInteraction isec;
        if (L.scene->Intersect(ray, &isec) && isec.GetHitEntity()) {
            id[i] = index[isec.GetHitEntity()];
            t[i] = isec.GetDistance();
            glm::vec3 p = isec.GetHitPos(), nn = isec.GetNormal();
            pos[i * 3] = p.x; pos[i * 3 + 1] = p.y; pos[i * 3 + 2] = p.z;
            nrm[i * 3] = nn.x; nrm[i * 3 + 1] = nn.y; nrm[i * 3 + 2] = nn.z;
        }
        any[i] = L.scene->IntersectTest(ray, r[6], r[7], nullptr) ? 1 : 0;
    }
    auto t1 = std::chrono::steady_clock::now();
    f = std::fopen(argv[5], "wb");
    if (!f) return 3;
    std::fwrite(id.data(), sizeof(int), id.size(), f);
    std::fwrite(t.data(), sizeof(float), t.size(), f);
    std::fwrite(pos.data(), sizeof(float), pos.size(), f);
    std::fwrite(nrm.data(), sizeof(float), nrm.size(), f);
    std::fwrite(any.data(), 1, any.size(), f);
    std::fclose(f);
    std::printf("{\"trace_s\": %.6f, \"rays\": %lld}\n",
                std::chrono::duration<double>(t1 - t0).count(), n);
    return 0;
}

int cmdRaycast(int argc, char** argv) {
    if (argc < 7) return 2;
    Loaded L = load(argv[2], argv[3]);
    auto objects = L.scene->GetObjects();
    std::unordered_map<const crystal::Entity*, int> index;
    for (size_t i = 0; i < objects.size(); i++) index[objects[i]] = (int)i;
    FILE* f = std::fopen(argv[4], "rb");
    if (!f) return 3;
    long long n = 0;
    if (std::fread(&n, sizeof(n), 1, f) != 1) return 3;
    std::vector<float> rays((size_t)n * 8);
    if (std::fread(rays.data(), sizeof(float), rays.size(), f) != rays.size()) return 3;
    std::fclose(f);
    const int nt = std::max(1, std::atoi(argv[5]));
    std::vector<int> id((size_t)n, -1);
    std::vector<float> t((size_t)n, 0.f);
    std::vector<unsigned char> any((size_t)n, 0);
    auto run = [&](bool anyHit) {
        std::vector<std::thread> pool;
        for (int w = 0; w < nt; w++)
            pool.emplace_back([&, w] {
                long long a = n * w / nt, b = n * (w + 1) / nt;
                for (long long i = a; i < b; i++) {
                    const float* r = &rays[(size_t)i * 8];
                    Ray3f ray(glm::vec3(r[0], r[1], r[2]), glm::vec3(r[3], r[4], r[5]));
                    if (anyHit) {
                        any[i] = L.scene->IntersectTest(ray, r[6], r[7], nullptr) ? 1 : 0;
                    } else {
                        SurfaceInteraction isec;
                        if (L.scene->Intersect(ray, &isec) && isec.GetHitEntity()) {
                            id[i] = index.at(isec.GetHitEntity());
                            t[i] = isec.GetDistance();
                        }
                    }
                }
            });
        for (auto& th : pool) th.join();
    };
    auto t0 = std::chrono::steady_clock::now();
    run(false);
    auto t1 = std::chrono::steady_clock::now();
    run(true);
    auto t2 = std::chrono::steady_clock::now();
    f = std::fopen(argv[6], "wb");
    if (!f) return 3;
    std::fwrite(id.data(), sizeof(int), id.size(), f);
    std::fwrite(t.data(), sizeof(float), t.size(), f);
    std::fwrite(any.data(), 1, any.size(), f);
    std::fclose(f);
    std::printf("{\"closest_s\": %.6f, \"any_s\": %.6f, \"rays\": %lld, \"threads\": %d, \"hw_threads\": %u}\n",
                std::chrono::duration<double>(t1 - t0).count(), std::chrono::duration<double>(t2 - t1).count(), n, nt,
                std::thread::hardware_concurrency());
    return 0;
}

int cmdHalton(int argc, char** argv) {
    if (argc < 8) return 2;
    int W = std::atoi(argv[2]), H = std::atoi(argv[3]), spp = std::atoi(argv[4]);
    int ndims = std::atoi(argv[6]);
    if (ndims > 100) return 4;  // dims >= 100 draw from a shared mt19937 (not reproducible)
    crystal::HaltonSampler sampler(spp, Bound2i(Point2i(0, 0), Point2i(W, H)));
    FILE* f = std::fopen(argv[5], "rb");
    if (!f) return 3;
    long long n = 0;
    if (std::fread(&n, sizeof(n), 1, f) != 1) return 3;
    std::vector<int> q((size_t)n * 3);
    if (std::fread(q.data(), sizeof(int), q.size(), f) != q.size()) return 3;
    std::fclose(f);
    std::vector<int> idx((size_t)n);
    std::vector<float> vals((size_t)n * ndims);
    for (long long i = 0; i < n; i++) {
        sampler.StartPixel(Point2i(q[i * 3], q[i * 3 + 1]));
        idx[i] = sampler.GetVectorIndexForSample(q[i * 3 + 2]);
        for (int d = 0; d < ndims; d++) vals[(size_t)i * ndims + d] = sampler.SampleValue(idx[i], d);
    }
    f = std::fopen(argv[7], "wb");
    if (!f) return 3;
    std::fwrite(idx.data(), sizeof(int), idx.size(), f);
    std::fwrite(vals.data(), sizeof(float), vals.size(), f);
    std::fclose(f);
    return 0;
}

int cmdWriteHdr(int argc, char** argv) {
    if (argc < 6) return 2;
    int W = std::atoi(argv[3]), H = std::atoi(argv[4]);
    std::vector<float> film((size_t)W * H * 3);
    FILE* f = std::fopen(argv[2], "rb");
    if (!f) return 3;
    size_t got = std::fread(film.data(), sizeof(float), film.size(), f);
    std::fclose(f);
    if (got != film.size()) return 3;
    return stbi_write_hdr(argv[5], W, H, 3, film.data()) ? 0 : 3;
}

int cmdDecode(int argc, char** argv) {
    if (argc < 4) return 2;
    int w = 0, h = 0, c = 0;
    unsigned char* px = stbi_load(argv[2], &w, &h, &c, 3);
    if (!px) return 3;
    FILE* f = std::fopen(argv[3], "wb");
    if (!f) return 3;
    int32_t wh[2] = {w, h};
    std::fwrite(wh, sizeof(int32_t), 2, f);
    std::fwrite(px, 1, (size_t)w * h * 3, f);
    std::fclose(f);
    stbi_image_free(px);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: crystal2_ref render|hits|raycast|halton|decode|writehdr ...\n");
        return 2;
    }
    std::string cmd = argv[1];
    int rc = 2;
    try {
        if (cmd == "render") rc = cmdRender(argc, argv);
        else if (cmd == "hits") rc = cmdHits(argc, argv);
        else if (cmd == "raycast") rc = cmdRaycast(argc, argv);
        else if (cmd == "halton") rc = cmdHalton(argc, argv);
        else if (cmd == "decode") rc = cmdDecode(argc, argv);
        else if (cmd == "writehdr") rc = cmdWriteHdr(argc, argv);
    } catch (const std::logic_error& e) {  // the plugins' std::invalid_argument
        std::fprintf(stderr, "error: %s\n", e.what());
        std::fflush(stderr);
        std::_Exit(4);
    } catch (const std::runtime_error& e) {
        // the reference's `std::exception("msg")` (ref_prelude.h spells `exception` as
        // runtime_error in this translation unit too, so it is named as such here)
        std::fprintf(stderr, "error: %s\n", e.what());
        std::fflush(stderr);
        std::_Exit(4);
    }
    if (rc == 2) std::fprintf(stderr, "bad arguments for %s\n", cmd.c_str());
    std::fflush(stdout);
    // The reference's FixedThreadPool workers are detached and still blocked on
    // their condition variable; leave without running static destructors.
    std::_Exit(rc);
}
