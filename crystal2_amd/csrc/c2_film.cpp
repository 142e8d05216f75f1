// Film output and the binary scene cache (host side).
//
//  * c2_film_write_hdr: the Radiance RGBE file the reference writes at the end of a render
//    (stbi_write_hdr("result.hdr", w, h, 3, film), src/Core/RayTracer.cpp:40), byte for byte:
//    the same header, the frexp-based RGBE conversion and stb_image_write v1.13's per-channel
//    run-length scheme (runs of 3 or more equal bytes, dumps of at most 128).
//  * c2_film_write_pfm: a lossless float film (PFM, bottom-up rows, little-endian).
//  * c2_scene_save / c2_scene_load_binary: every array of a c2_scene in one file, so a scene
//    parsed once from JSON + OBJ (sscanf-bound) reloads at file-read speed.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "c2_host.h"

namespace c2 {
namespace {

struct File {
    FILE* f = nullptr;
    explicit File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

void put(File& f, const void* p, size_t n) {
    if (n && std::fwrite(p, 1, n, f.f) != n) fail(C2_ERR_IO, "write failed");
}

// stbiw__linear_to_rgbe
void to_rgbe(uint8_t rgbe[4], const float c[3]) {
    float m = c[1] > c[2] ? c[1] : c[2];
    m = c[0] > m ? c[0] : m;
    if (m < 1e-32f) {
        rgbe[0] = rgbe[1] = rgbe[2] = rgbe[3] = 0;
        return;
    }
    int e;
    float norm = (float)std::frexp(m, &e) * 256.0f / m;
    rgbe[0] = (uint8_t)(c[0] * norm);
    rgbe[1] = (uint8_t)(c[1] * norm);
    rgbe[2] = (uint8_t)(c[2] * norm);
    rgbe[3] = (uint8_t)(e + 128);
}

// one scanline: flat RGBE below 8 or from 32768 pixels, else the per-channel RLE
void hdr_scanline(std::vector<uint8_t>& out, int w, const float* row) {
    std::vector<uint8_t> px((size_t)w * 4);
    for (int x = 0; x < w; x++) to_rgbe(&px[(size_t)x * 4], row + (size_t)x * 3);
    if (w < 8 || w >= 32768) {
        out.insert(out.end(), px.begin(), px.end());
        return;
    }
    out.push_back(2);
    out.push_back(2);
    out.push_back((uint8_t)((w & 0xff00) >> 8));
    out.push_back((uint8_t)(w & 0xff));
    std::vector<uint8_t> ch(w);
    for (int c = 0; c < 4; c++) {
        for (int x = 0; x < w; x++) ch[x] = px[(size_t)x * 4 + c];
        int x = 0;
        while (x < w) {
            int r = x;  // start of the next run of three equal bytes
            while (r + 2 < w && !(ch[r] == ch[r + 1] && ch[r] == ch[r + 2])) r++;
            bool run = r + 2 < w;
            if (!run) r = w;
            while (x < r) {  // literal bytes before it, 128 at most per packet
                int len = std::min(r - x, 128);
                out.push_back((uint8_t)len);
                out.insert(out.end(), ch.begin() + x, ch.begin() + x + len);
                x += len;
            }
            if (run) {
                while (r < w && ch[r] == ch[x]) r++;
                while (x < r) {  // the run, 127 at most per packet
                    int len = std::min(r - x, 127);
                    out.push_back((uint8_t)(len + 128));
                    out.push_back(ch[x]);
                    x += len;
                }
            }
        }
    }
}

const char kMagic[8] = {'C', '2', 'S', 'C', 'E', 'N', 'E', '1'};

template <class T>
void put_vec(File& f, const std::vector<T>& v) {
    put(f, v.data(), v.size() * sizeof(T));
}

struct Reader {
    FILE* f;
    void get(void* p, size_t n) {
        if (n && std::fread(p, 1, n, f) != n) fail(C2_ERR_PARSE, "scene cache: truncated file");
    }
    template <class T>
    void vec(std::vector<T>& v, int64_t n) {
        if (n < 0 || n > (int64_t)1 << 34) fail(C2_ERR_PARSE, "scene cache: bad array size");
        v.resize((size_t)n);
        get(v.data(), (size_t)n * sizeof(T));
    }
};

}  // namespace
}  // namespace c2

using namespace c2;

extern "C" {

int c2_film_write_hdr(const char* path, int32_t w, int32_t h, const float* rgb) {
    return guard([&] {
        if (!path || !rgb || w <= 0 || h <= 0) fail(C2_ERR_INVALID, "bad arguments");
        std::vector<uint8_t> out;
        const char header[] = "#?RADIANCE\n# Written by stb_image_write.h\nFORMAT=32-bit_rle_rgbe\n";
        out.insert(out.end(), header, header + sizeof(header) - 1);
        char buf[128];
        int n = std::snprintf(buf, sizeof(buf), "EXPOSURE=          1.0000000000000\n\n-Y %d +X %d\n", h, w);
        out.insert(out.end(), buf, buf + n);
        for (int y = 0; y < h; y++) hdr_scanline(out, w, rgb + (size_t)y * w * 3);
        File f(path, "wb");
        if (!f.f) fail(C2_ERR_IO, std::string("cannot create ") + path);
        put(f, out.data(), out.size());
    });
}

int c2_film_write_pfm(const char* path, int32_t w, int32_t h, const float* rgb) {
    return guard([&] {
        if (!path || !rgb || w <= 0 || h <= 0) fail(C2_ERR_INVALID, "bad arguments");
        File f(path, "wb");
        if (!f.f) fail(C2_ERR_IO, std::string("cannot create ") + path);
        char hdr[64];
        int n = std::snprintf(hdr, sizeof(hdr), "PF\n%d %d\n-1.0\n", w, h);  // negative scale: little-endian
        put(f, hdr, (size_t)n);
        for (int y = h - 1; y >= 0; y--) put(f, rgb + (size_t)y * w * 3, (size_t)w * 3 * sizeof(float));
    });
}

int c2_scene_save(const c2_scene* s, const char* path) {
    return guard([&] {
        if (!s || !path) fail(C2_ERR_INVALID, "null argument");
        const SceneData& d = s->d;
        File f(path, "wb");
        if (!f.f) fail(C2_ERR_IO, std::string("cannot create ") + path);
        put(f, kMagic, 8);
        int64_t counts[8] = {(int64_t)d.kind.size(), (int64_t)d.tris.size(),    (int64_t)d.spheres.size(),
                             (int64_t)d.materials.size(), (int64_t)d.lights.size(), (int64_t)d.textures.size(),
                             d.has_camera ? 1 : 0, 0};
        put(f, counts, sizeof(counts));
        put(f, d.skybox, sizeof(d.skybox));
        put(f, &d.camera, sizeof(d.camera));
        put_vec(f, d.kind);
        put_vec(f, d.shape);
        put_vec(f, d.material);
        put_vec(f, d.light);
        put_vec(f, d.tris);
        put_vec(f, d.spheres);
        put_vec(f, d.materials);
        put_vec(f, d.lights);
        for (const TextureData& t : d.textures) {
            int32_t wh[2] = {t.width, t.height};
            put(f, wh, sizeof(wh));
            put_vec(f, t.rgb);
        }
    });
}

int c2_scene_load_binary(const char* path, c2_scene** out) {
    return guard([&] {
        if (!path || !out) fail(C2_ERR_INVALID, "null argument");
        File f(path, "rb");
        if (!f.f) fail(C2_ERR_IO, std::string("cannot open ") + path);
        Reader r{f.f};
        char magic[8];
        r.get(magic, 8);
        if (std::memcmp(magic, kMagic, 8)) fail(C2_ERR_PARSE, std::string(path) + ": not a scene cache file");
        int64_t c[8];
        r.get(c, sizeof(c));
        std::unique_ptr<c2_scene> s(new c2_scene());
        SceneData& d = s->d;
        r.get(d.skybox, sizeof(d.skybox));
        r.get(&d.camera, sizeof(d.camera));
        d.has_camera = c[6] != 0;
        r.vec(d.kind, c[0]);
        r.vec(d.shape, c[0]);
        r.vec(d.material, c[0]);
        r.vec(d.light, c[0]);
        r.vec(d.tris, c[1]);
        r.vec(d.spheres, c[2]);
        r.vec(d.materials, c[3]);
        r.vec(d.lights, c[4]);
        if (c[5] < 0 || c[5] > (1 << 20)) fail(C2_ERR_PARSE, "scene cache: bad texture count");
        d.textures.resize((size_t)c[5]);
        for (TextureData& t : d.textures) {
            int32_t wh[2];
            r.get(wh, sizeof(wh));
            if (wh[0] <= 0 || wh[1] <= 0 || (int64_t)wh[0] * wh[1] > (int64_t)1 << 28)
                fail(C2_ERR_PARSE, "scene cache: bad texture size");
            t.width = wh[0];
            t.height = wh[1];
            r.vec(t.rgb, (int64_t)wh[0] * wh[1] * 3);
        }
        if (c[1] % 24 || c[2] % 22) fail(C2_ERR_PARSE, "scene cache: bad shape arrays");
        // the same invariants the builder keeps
        const int64_t nt = c[1] / 24, ns = c[2] / 22;
        for (int64_t e = 0; e < c[0]; e++) {
            int k = d.kind[e], sh = d.shape[e];
            bool ok = (k == C2_SHAPE_TRIANGLE && sh >= 0 && sh < nt) ||
                      ((k == C2_SHAPE_SPHERE || k == C2_SHAPE_DISK) && sh >= 0 && sh < ns);
            ok = ok && d.material[e] >= -1 && d.material[e] < c[3] && d.light[e] >= -1 && d.light[e] < c[4];
            if (!ok) fail(C2_ERR_PARSE, "scene cache: entity record out of range");
        }
        *out = s.release();
    });
}

}  // extern "C"
