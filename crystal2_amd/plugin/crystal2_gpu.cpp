// Crystal2 plugins over the MI355X library: "GPUBVH" (Accelerator) and "WavefrontPathTracer"
// (Integrator). See crystal2_gpu.h for how a maintainer registers them.
#include "crystal2_gpu.h"

#include <Core/Entities.h>
#include <Core/FrameBuffer.h>
#include <Core/Scene.h>
#include <Core/SurfaceInteraction.h>
#include <Lights/AreaLight.h>
#include <SJson/SJson.h>

#include <crystal2.h>

#include <limits>
#include <mutex>
#include <stdexcept>
#include <unordered_map>
#include <vector>

namespace crystal2_gpu {
namespace {

std::mutex g_mu;
std::string g_sceneFile;

// A library status as the exception the reference throws for bad input.
void check(int rc) {
    if (rc != C2_OK) throw std::invalid_argument(std::string("crystal2 GPU: ") + c2_last_error());
}

// Get1D / Get2D as arguments of one call (PathTracingIntegrator.cpp:79) are evaluated
// right to left by MSVC and GCC and left to right by clang; the plugin is compiled by the
// compiler that compiles the reference, so it takes that compiler's order.
#if defined(__clang__)
constexpr int32_t kArgOrder = C2_ARGORDER_LTR;
#else
constexpr int32_t kArgOrder = C2_ARGORDER_RTL;
#endif

// The device scene of the scene file (owns the arrays its descriptor points into).
struct DeviceScene {
    c2_scene* scene = nullptr;
    c2_scene_desc desc{};
    DeviceScene() {
        std::string path;
        {
            std::lock_guard<std::mutex> lk(g_mu);
            path = g_sceneFile;
        }
        if (path.empty()) throw std::invalid_argument("crystal2 GPU: no scene file (crystal2_gpu::SetSceneFile)");
        check(c2_scene_load(path.c_str(), nullptr, &scene));
        check(c2_scene_get_desc(scene, &desc));
    }
    ~DeviceScene() { c2_scene_destroy(scene); }
    DeviceScene(const DeviceScene&) = delete;
    DeviceScene& operator=(const DeviceScene&) = delete;
};

class GpuBVH final : public Accelerator {
public:
    ~GpuBVH() override {
        if (_accel) c2_accel_destroy(_accel);
    }

    void Build(const std::vector<const crystal::IIntersectable*>& objects) override {
        _scene = std::make_unique<DeviceScene>();
        if (_scene->desc.num_entities != (int32_t)objects.size())
            throw std::invalid_argument("crystal2 GPU: the scene file's entities differ from the scene's objects");
        _objects = objects;
        _index.clear();
        for (size_t i = 0; i < objects.size(); i++) {
            _index[objects[i]] = (int32_t)i;
            // the integrators pass a light's attached shape as the shape to ignore
            // (PathTracingIntegrator.cpp:212, GeometricEntity::IntersectTest compares it
            // with its own shape), so the entity of that shape is the one skipped
            const auto* entity = dynamic_cast<const crystal::Entity*>(objects[i]);
            const crystal::AreaLight* light = entity ? entity->GetAreaLight() : nullptr;
            if (light && light->GetAttachedObject()) _index[light->GetAttachedObject()] = (int32_t)i;
        }
        if (_accel) c2_accel_destroy(_accel);
        _accel = nullptr;
        check(c2_accel_build(&_scene->desc, &_accel));
    }

    // BVH::Intersect (BVH.cpp:56-105): the closest entity comes from the device; its own
    // Intersect fills the interaction, which replaces *info when closer.
    bool Intersect(const Ray3f& ray, SurfaceInteraction* info, float tMin, float tMax) const override {
        const float r[8] = {ray.start.x, ray.start.y, ray.start.z, ray.dir.x, ray.dir.y, ray.dir.z, tMin, tMax};
        int32_t prim = -1;
        float t = 0.f;
        check(c2_accel_intersect(_accel, r, 1, &prim, &t));
        if (prim < 0) return false;
        SurfaceInteraction isec;
        if (!_objects[prim]->Intersect(ray, &isec)) return false;
        const float dis = isec.GetDistance();
        if (dis < info->GetDistance()) {
            isec.SetHitEntity(dynamic_cast<const crystal::Entity*>(_objects[prim]));
            *info = std::move(isec);
        }
        return dis < std::numeric_limits<float>::infinity();
    }

    bool IntersectTest(const Ray3f& ray, const crystal::IIntersectable* ignoreShape, float tMin,
                       float tMax) const override {
        const float r[8] = {ray.start.x, ray.start.y, ray.start.z, ray.dir.x, ray.dir.y, ray.dir.z, tMin, tMax};
        int32_t ignore = -1;
        if (ignoreShape) {
            auto it = _index.find(ignoreShape);
            if (it != _index.end()) ignore = it->second;
        }
        uint8_t hit = 0;
        check(c2_accel_intersect_test(_accel, r, &ignore, 1, &hit));
        return hit != 0;
    }

private:
    std::unique_ptr<DeviceScene> _scene;
    c2_accel* _accel = nullptr;
    std::vector<const crystal::IIntersectable*> _objects;
    std::unordered_map<const crystal::IIntersectable*, int32_t> _index;
};

class WavefrontPathTracer final : public crystal::Integrator {
public:
    WavefrontPathTracer(int maxDepth, const config::ConfigInfo& info) : _depth(maxDepth), _info(info) {}

    void Render(Scene*, Camera*, FrameBuffer* frameBuffer) override {
        DeviceScene scene;
        c2_render_opts o{};
        o.width = _info.Width;
        o.height = _info.Height;
        o.spp = _info.SamplesPerPixel;
        o.max_depth = _depth;
        o.arg_order = kArgOrder;
        o.film_mode = C2_FILM_F32;  // FrameBuffer's fp32 sums
        o.integrator = C2_INTEGRATOR_PATH;
        o.accel = _info.AccelType == "KDTree" ? C2_ACCEL_KDTREE : C2_ACCEL_BVH;
        c2_pt* pt = nullptr;
        check(c2_pt_create(&scene.desc, &o, &pt));
        std::unique_ptr<c2_pt, void (*)(c2_pt*)> owned(pt, c2_pt_destroy);
        check(c2_pt_render(pt, 0, o.spp, nullptr));  // every sample, in sample order
        std::vector<float> rgb((size_t)o.width * o.height * 3);
        check(c2_pt_get_film(pt, rgb.data()));  // GetImageDataRGB32F's layout: rows top-down
        // One sample of weight 1 per pixel: FrameBuffer::GetImageDataRGB32F then returns
        // (0 + c * 1) / 1 = c for every channel, the device film bit for bit.
        for (int row = 0; row < o.height; row++)
            for (int x = 0; x < o.width; x++) {
                const float* c = &rgb[((size_t)row * o.width + x) * 3];
                frameBuffer->AddSample(x, o.height - 1 - row, glm::vec3(c[0], c[1], c[2]), 1.f);
            }
    }

private:
    int _depth;
    config::ConfigInfo _info;
};

}  // namespace

void SetSceneFile(const std::string& sceneJsonPath) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_sceneFile = sceneJsonPath;
}

std::unique_ptr<Accelerator> MakeAccelerator() { return std::make_unique<GpuBVH>(); }

std::unique_ptr<crystal::Integrator> MakeIntegrator(JsonNode_CPTR integratorNode, const config::ConfigInfo& info) {
    const auto sampler = integratorNode->GetMember("Sampler")->GetMember("Type")->GetString();
    if (sampler != "Halton")
        throw std::invalid_argument("crystal2 GPU: WavefrontPathTracer samples with \"Halton\", not \"" + sampler + "\"");
    return std::make_unique<WavefrontPathTracer>((int)integratorNode->GetMember("Depth")->GetInt(), info);
}

}  // namespace crystal2_gpu
