// Image decoding for textures and the skybox: PNG and baseline JPEG to 8-bit RGB.
//
// The reference loads every image with stbi_load(path, ..., 3)
// (src/Textures/Texel.cpp:4-13, stb_image v2.23 vendored under deps/TREngine/deps/stbi).
// These decoders return the same bytes for the same file: PNG is lossless, so any
// correct decoder agrees; for JPEG the integer IDCT, the chroma upsampling filters and
// the fixed-point YCbCr conversion follow stb_image's arithmetic step for step.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace c2 {

struct ImageRGB {
    int width = 0, height = 0;
    std::vector<uint8_t> rgb;  // width * height * 3, rows top-down
};

// Decode a PNG or JPEG file (format from the signature). Throws c2::Error: C2_ERR_IO when
// the file cannot be read, C2_ERR_PARSE for malformed data, C2_ERR_UNSUPPORTED for
// progressive or arithmetic-coded JPEG and other formats.
ImageRGB load_image(const std::string& path);
ImageRGB decode_image(const uint8_t* data, size_t size);

}  // namespace c2
