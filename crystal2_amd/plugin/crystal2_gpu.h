// Crystal2 plugins over the MI355X library (include/crystal2.h): the "GPUBVH" accelerator
// and the "WavefrontPathTracer" integrator. A maintainer compiles crystal2_gpu.cpp inside
// the reference's source tree (it includes src/Accelerators/Accelerator.h and
// src/Integrators/Integrator.h), links libcrystal2.so, and adds one branch to each
// string-keyed factory:
//
//   Accelerator::GetAccelerator (src/Accelerators/Accelerator.cpp:8-23):
//       else if (name == "GPUBVH") return crystal2_gpu::MakeAccelerator();
//   Integrator::LoadIntegrator (src/Integrators/Integrator.cpp:27-40):
//       if (type == "WavefrontPathTracer") return crystal2_gpu::MakeIntegrator(integratorNode, configInfo);
//
// and, where the host loads the scene file (Crystal2.cpp:18 reads it from argv[2]):
//       crystal2_gpu::SetSceneFile(argv[2]);
//
// The reference's shapes keep their geometry private, so the plugins build the device
// scene from that same file (c2_scene_load: the reference's number parsing, JSON object
// order = entity order, OBJ paths relative to the working directory, as Scene::CreateScene
// reads them). Device entity i is then Scene::GetObjects()[i].
#pragma once
#include <Accelerators/Accelerator.h>
#include <Core/ConfigInfo.h>
#include <Integrators/Integrator.h>

#include <memory>
#include <string>

namespace crystal2_gpu {

// The scene file both plugins build their device scene from; call before
// Scene::CreateScene (which builds the accelerator) and Integrator::Render.
void SetSceneFile(const std::string& sceneJsonPath);

// "GPUBVH": Accelerator::Build / Intersect / IntersectTest (Accelerator.h:8-22) on the GPU
// BVH. Closest and any hit run on the device; the SurfaceInteraction of the closest hit is
// filled by that entity's own Intersect, as BVH::Intersect fills it (BVH.cpp:56-105).
// The ignored shape of IntersectTest is passed as its entity index.
std::unique_ptr<Accelerator> MakeAccelerator();

// "WavefrontPathTracer": Integrator::Render (Integrator.h:8-18) as the GPU wavefront path
// tracer, PathTracingIntegrator (PathTracingIntegrator.cpp:28-278) with the Halton sampler
// and the config's Depth, Width, Height and SamplesPerPixel. The film lands in the
// FrameBuffer so that GetImageDataRGB32F returns the GPU film bit for bit. The config's
// Acceleration "KDTree" selects the device KD-tree (the reference's tree and hits).
// Throws std::invalid_argument for a sampler other than "Halton".
std::unique_ptr<crystal::Integrator> MakeIntegrator(JsonNode_CPTR integratorNode, const config::ConfigInfo& info);

}  // namespace crystal2_gpu
