// Host-side internals shared by the C-ABI implementation files.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/crystal2.h"

namespace c2 {

// Error carried from the inside of the library to the C-ABI boundary, where it
// becomes a status code plus c2_last_error() text.
struct Error {
    int code;
    std::string msg;
};
[[noreturn]] void fail(int code, const std::string& msg);
void set_last_error(const std::string& msg);

// Runs f() and converts any c2::Error / std::exception into a status code.
template <class F>
int guard(F&& f) {
    try {
        f();
        set_last_error("");
        return C2_OK;
    } catch (const Error& e) {
        set_last_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return C2_ERR_NOMEM;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return C2_ERR_INVALID;
    }
}

// Halton sampler constants for a W x H film (src/Samplers/HaltonSampler.cpp:204-229).
struct HaltonParams {
    int32_t base_scales[2];
    int32_t base_exponents[2];
    int32_t stride;
    int32_t mult_inverse[2];
};
HaltonParams halton_params(int width, int height);
// Vector index of sample k of pixel (x, y) (HaltonSampler.cpp:251-270).
int32_t halton_index(const HaltonParams& p, int x, int y, int k);
// Value of dimension d at vector index idx (HaltonSampler.cpp:272-291, dims >= 100:
// deterministic hash substitute).
float halton_value(const HaltonParams& p, int32_t idx, int32_t dim);
// Permutation tables for dims 0..99 concatenated (std::shuffle over a default-seeded
// std::mt19937, HaltonSampler.cpp:164-178), their offsets and bases (first 100 primes).
const std::vector<uint16_t>& halton_perm_data();
const std::vector<int32_t>& halton_perm_offsets();
const std::vector<int32_t>& halton_primes();

// Deterministic substitute for the reference's shared mt19937 draws (dims >= 100,
// HaltonSampler.cpp:285-288, and Blinn-Phong rejection redraws,
// BlinnPhongReflection.cpp:31-53): a 32-bit integer hash of (index, stream)
// mapped exactly like uniformRandomFloat (src/Core/Utils.h:12-15).
inline uint32_t hash_u32(uint32_t a, uint32_t b) {
    uint32_t h = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u) * 0x85EBCA77u;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return h;
}

struct TextureData {
    int32_t width = 0, height = 0;
    std::vector<uint8_t> rgb;  // rows top-down
};

struct SceneData {
    std::vector<int32_t> kind, shape, material, light;
    std::vector<float> tris;     // [n][3][8]
    std::vector<float> spheres;  // [n][22]
    std::vector<c2_material> materials;
    std::vector<c2_light> lights;
    std::vector<TextureData> textures;     // id k = textures[k - 1]
    mutable std::vector<c2_texture> texture_desc;  // views of `textures` for c2_scene_desc
    int32_t skybox[6] = {0, 0, 0, 0, 0, 0};
    c2_camera camera{};
    bool has_camera = false;
};

}  // namespace c2

struct c2_scene {
    c2::SceneData d;
};
