// PNG and baseline-JPEG decoders for textures and the skybox (host side).
//
// Output contract: the bytes stbi_load(path, &w, &h, &c, 3) returns in the reference
// (src/Textures/Texel.cpp:4-13, stb_image v2.23). What that fixes:
//  * PNG: 16-bit samples keep their high byte; gray samples of 1/2/4 bits are scaled by
//    255/(2^d - 1); palette indices are not scaled; alpha is dropped; gray is replicated.
//  * JPEG: the "islow" integer IDCT with 12-bit constants (2 extra bits kept between the
//    passes, +128 level shift folded into the final rounding), the triangle-filter chroma
//    upsampling for 2x1, 1x2 and 2x2 subsampling (nearest for other factors), and the
//    fixed-point YCbCr -> RGB conversion with 20-bit scaled constants. Component ids
//    'R','G','B', or an Adobe marker with transform 0 and no JFIF marker, mean the three
//    components are already RGB. CMYK/YCCK use the 8x8-bit "blinn" product.
// Progressive and arithmetic-coded JPEGs are rejected (C2_ERR_UNSUPPORTED).
#include "c2_image.h"

#include <cstring>
#include <fstream>
#include <iterator>

#include "c2_host.h"

namespace c2 {
namespace {

// ================================================================ inflate (RFC 1950/1951)

struct BitsLSB {
    const uint8_t* p;
    size_t n, pos = 0;
    uint32_t buf = 0;
    int cnt = 0;
    uint32_t get(int k) {
        while (cnt < k) {
            if (pos >= n) fail(C2_ERR_PARSE, "PNG: truncated deflate stream");
            buf |= (uint32_t)p[pos++] << cnt;
            cnt += 8;
        }
        uint32_t v = k ? (buf & ((1u << k) - 1)) : 0;
        buf = k < 32 ? buf >> k : 0;
        cnt -= k;
        return v;
    }
};

// Canonical Huffman code: count of codes per length, symbols in code order.
struct Canon {
    uint16_t count[16];
    uint16_t sym[320];
};

void canon_build(Canon& h, const uint8_t* len, int n) {
    std::memset(h.count, 0, sizeof(h.count));
    for (int i = 0; i < n; i++) h.count[len[i]]++;
    h.count[0] = 0;
    int left = 1;
    for (int l = 1; l < 16; l++) {
        left <<= 1;
        left -= h.count[l];
        if (left < 0) fail(C2_ERR_PARSE, "PNG: over-subscribed Huffman code");
    }
    uint16_t offs[16];
    offs[1] = 0;
    for (int l = 1; l < 15; l++) offs[l + 1] = (uint16_t)(offs[l] + h.count[l]);
    for (int s = 0; s < n; s++)
        if (len[s]) h.sym[offs[len[s]]++] = (uint16_t)s;
}

int canon_decode(BitsLSB& b, const Canon& h) {
    int code = 0, first = 0, index = 0;
    for (int l = 1; l < 16; l++) {
        code |= (int)b.get(1);
        int c = h.count[l];
        if (code - c < first) return h.sym[index + (code - first)];
        index += c;
        first += c;
        first <<= 1;
        code <<= 1;
    }
    fail(C2_ERR_PARSE, "PNG: invalid Huffman code");
}

const uint16_t kLenBase[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                               31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
const uint8_t kLenExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
const uint16_t kDistBase[30] = {1,   2,   3,   4,   5,   7,    9,    13,   17,   25,   33,   49,   65,    97,    129,
                                193, 257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
const uint8_t kDistExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};

std::vector<uint8_t> inflate_zlib(const uint8_t* p, size_t n, size_t expect) {
    if (n < 2) fail(C2_ERR_PARSE, "PNG: empty image data");
    int cmf = p[0], flg = p[1];
    if ((cmf * 256 + flg) % 31 != 0 || (cmf & 15) != 8 || (flg & 32))
        fail(C2_ERR_PARSE, "PNG: bad zlib header");
    BitsLSB b{p + 2, n - 2};
    std::vector<uint8_t> out;
    out.reserve(expect);
    Canon lit, dist;
    int last;
    do {
        last = (int)b.get(1);
        int type = (int)b.get(2);
        if (type == 0) {
            b.get(b.cnt & 7);
            uint32_t len = b.get(16), nlen = b.get(16);
            if ((len ^ 0xFFFFu) != nlen) fail(C2_ERR_PARSE, "PNG: corrupt stored block");
            for (uint32_t i = 0; i < len; i++) out.push_back((uint8_t)b.get(8));
            continue;
        }
        if (type == 3) fail(C2_ERR_PARSE, "PNG: bad deflate block type");
        uint8_t lens[320];
        int nlit, ndist;
        if (type == 1) {
            nlit = 288;
            ndist = 30;
            for (int i = 0; i < 144; i++) lens[i] = 8;
            for (int i = 144; i < 256; i++) lens[i] = 9;
            for (int i = 256; i < 280; i++) lens[i] = 7;
            for (int i = 280; i < 288; i++) lens[i] = 8;
            for (int i = 0; i < 30; i++) lens[288 + i] = 5;
        } else {
            nlit = (int)b.get(5) + 257;
            ndist = (int)b.get(5) + 1;
            int ncl = (int)b.get(4) + 4;
            static const uint8_t order[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
            uint8_t cl[19] = {0};
            for (int i = 0; i < ncl; i++) cl[order[i]] = (uint8_t)b.get(3);
            Canon clh;
            canon_build(clh, cl, 19);
            int k = 0;
            while (k < nlit + ndist) {
                int s = canon_decode(b, clh);
                if (s < 16) {
                    lens[k++] = (uint8_t)s;
                    continue;
                }
                int rep, val = 0;
                if (s == 16) {
                    if (k == 0) fail(C2_ERR_PARSE, "PNG: repeat with no previous length");
                    val = lens[k - 1];
                    rep = 3 + (int)b.get(2);
                } else if (s == 17) {
                    rep = 3 + (int)b.get(3);
                } else {
                    rep = 11 + (int)b.get(7);
                }
                if (k + rep > nlit + ndist) fail(C2_ERR_PARSE, "PNG: too many code lengths");
                while (rep--) lens[k++] = (uint8_t)val;
            }
        }
        canon_build(lit, lens, nlit);
        canon_build(dist, lens + nlit, ndist);
        for (;;) {
            int s = canon_decode(b, lit);
            if (s < 256) {
                out.push_back((uint8_t)s);
            } else if (s == 256) {
                break;
            } else {
                s -= 257;
                if (s >= 29) fail(C2_ERR_PARSE, "PNG: bad length code");
                size_t len = kLenBase[s] + b.get(kLenExtra[s]);
                int ds = canon_decode(b, dist);
                if (ds >= 30) fail(C2_ERR_PARSE, "PNG: bad distance code");
                size_t d = kDistBase[ds] + b.get(kDistExtra[ds]);
                if (d > out.size()) fail(C2_ERR_PARSE, "PNG: distance too far back");
                size_t from = out.size() - d;
                for (size_t i = 0; i < len; i++) out.push_back(out[from + i]);
            }
        }
    } while (!last);
    return out;
}

// ================================================================================ PNG

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

int paeth(int a, int b, int c) {
    int p = a + b - c;
    int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    if (pb <= pc) return b;
    return c;
}

// Undo the per-row filters of one (sub)image in place; rows of `stride` bytes follow
// their filter-type byte in `raw`, results go to `out` (h * stride).
void unfilter(const uint8_t* raw, size_t rawsize, uint8_t* out, int h, size_t stride, int bpp) {
    if ((size_t)h * (stride + 1) > rawsize) fail(C2_ERR_PARSE, "PNG: not enough image data");
    for (int y = 0; y < h; y++) {
        const uint8_t* src = raw + (size_t)y * (stride + 1);
        int ft = src[0];
        src++;
        uint8_t* cur = out + (size_t)y * stride;
        const uint8_t* prev = y ? cur - stride : nullptr;
        for (size_t i = 0; i < stride; i++) {
            int a = i >= (size_t)bpp ? cur[i - bpp] : 0;
            int b = prev ? prev[i] : 0;
            int c = (prev && i >= (size_t)bpp) ? prev[i - bpp] : 0;
            int x = src[i], v;
            switch (ft) {
                case 0: v = x; break;
                case 1: v = x + a; break;
                case 2: v = x + b; break;
                case 3: v = x + ((a + b) >> 1); break;
                case 4: v = x + paeth(a, b, c); break;
                default: fail(C2_ERR_PARSE, "PNG: bad filter type");
            }
            cur[i] = (uint8_t)v;
        }
    }
}

ImageRGB decode_png(const uint8_t* d, size_t n) {
    size_t pos = 8;
    int w = 0, h = 0, depth = 0, color = -1, interlace = 0;
    std::vector<uint8_t> pal, idat;
    bool seen_ihdr = false, seen_iend = false;
    while (pos + 12 <= n && !seen_iend) {
        uint32_t len = be32(d + pos);
        const uint8_t* type = d + pos + 4;
        const uint8_t* body = d + pos + 8;
        if (len > n - pos - 12) fail(C2_ERR_PARSE, "PNG: truncated chunk");
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len != 13) fail(C2_ERR_PARSE, "PNG: bad IHDR");
            w = (int)be32(body);
            h = (int)be32(body + 4);
            depth = body[8];
            color = body[9];
            interlace = body[12];
            if (body[10] || body[11] || interlace > 1) fail(C2_ERR_PARSE, "PNG: bad IHDR");
            seen_ihdr = true;
        } else if (!std::memcmp(type, "CgBI", 4)) {
            fail(C2_ERR_UNSUPPORTED, "PNG: Apple CgBI images are not supported");
        } else if (!std::memcmp(type, "PLTE", 4)) {
            if (len % 3 || len > 768) fail(C2_ERR_PARSE, "PNG: bad palette");
            pal.assign(body, body + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), body, body + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            seen_iend = true;
        } else if (!(type[0] & 32)) {
            fail(C2_ERR_PARSE, "PNG: unknown critical chunk");
        }
        pos += 12 + (size_t)len;
    }
    if (!seen_ihdr || w <= 0 || h <= 0 || w > (1 << 24) || h > (1 << 24)) fail(C2_ERR_PARSE, "PNG: missing or bad IHDR");
    int ch;
    switch (color) {
        case 0: ch = 1; break;
        case 2: ch = 3; break;
        case 3: ch = 1; break;
        case 4: ch = 2; break;
        case 6: ch = 4; break;
        default: fail(C2_ERR_PARSE, "PNG: bad colour type");
    }
    bool okDepth = depth == 8 || (depth == 16 && color != 3) || ((depth == 1 || depth == 2 || depth == 4) && (color == 0 || color == 3));
    if (!okDepth) fail(C2_ERR_PARSE, "PNG: bad bit depth");
    if (color == 3 && pal.empty()) fail(C2_ERR_PARSE, "PNG: palette image without PLTE");
    if (idat.empty()) fail(C2_ERR_PARSE, "PNG: no image data");

    const int bpp = std::max(1, ch * depth / 8);
    auto stride_of = [&](int pw) { return ((size_t)pw * ch * depth + 7) / 8; };
    size_t expect = (size_t)h * (stride_of(w) + 1);
    std::vector<uint8_t> raw = inflate_zlib(idat.data(), idat.size(), expect);

    ImageRGB img;
    img.width = w;
    img.height = h;
    img.rgb.assign((size_t)w * h * 3, 0);
    const int scale = color == 3 ? 1 : (depth == 1 ? 0xFF : depth == 2 ? 0x55 : depth == 4 ? 0x11 : 1);
    // store one unfiltered (sub)image row into the output at pixels x0 + k * dx of row y
    auto emit = [&](const uint8_t* row, int pw, int y, int x0, int dx) {
        for (int k = 0; k < pw; k++) {
            int s[4] = {0, 0, 0, 0};
            for (int c = 0; c < ch; c++) {
                if (depth == 8) {
                    s[c] = row[k * ch + c];
                } else if (depth == 16) {
                    s[c] = row[(k * ch + c) * 2];  // high byte
                } else {
                    size_t bit = (size_t)(k * ch + c) * depth;
                    int v = (row[bit >> 3] >> (8 - depth - (int)(bit & 7))) & ((1 << depth) - 1);
                    s[c] = v * scale;
                }
            }
            uint8_t* o = &img.rgb[((size_t)y * w + x0 + (size_t)k * dx) * 3];
            if (color == 3) {
                size_t e = (size_t)s[0] * 3;
                if (e + 2 < pal.size()) {
                    o[0] = pal[e];
                    o[1] = pal[e + 1];
                    o[2] = pal[e + 2];
                }
            } else if (ch >= 3) {
                o[0] = (uint8_t)s[0];
                o[1] = (uint8_t)s[1];
                o[2] = (uint8_t)s[2];
            } else {
                o[0] = o[1] = o[2] = (uint8_t)s[0];
            }
        }
    };
    if (!interlace) {
        size_t stride = stride_of(w);
        std::vector<uint8_t> pix((size_t)h * stride);
        unfilter(raw.data(), raw.size(), pix.data(), h, stride, bpp);
        for (int y = 0; y < h; y++) emit(pix.data() + (size_t)y * stride, w, y, 0, 1);
    } else {  // Adam7
        static const int xo[7] = {0, 4, 0, 2, 0, 1, 0}, yo[7] = {0, 0, 4, 0, 2, 0, 1};
        static const int xs[7] = {8, 8, 4, 4, 2, 2, 1}, ys[7] = {8, 8, 8, 4, 4, 2, 2};
        size_t off = 0;
        for (int pass = 0; pass < 7; pass++) {
            int pw = (w - xo[pass] + xs[pass] - 1) / xs[pass];
            int ph = (h - yo[pass] + ys[pass] - 1) / ys[pass];
            if (pw <= 0 || ph <= 0) continue;
            size_t stride = stride_of(pw);
            if (off > raw.size()) fail(C2_ERR_PARSE, "PNG: not enough image data");
            std::vector<uint8_t> pix((size_t)ph * stride);
            unfilter(raw.data() + off, raw.size() - off, pix.data(), ph, stride, bpp);
            for (int j = 0; j < ph; j++) emit(pix.data() + (size_t)j * stride, pw, yo[pass] + j * ys[pass], xo[pass], xs[pass]);
            off += (size_t)ph * (stride + 1);
        }
    }
    return img;
}

// =============================================================================== JPEG

// natural-order index of the k-th coefficient in zigzag order; the 15 extra entries keep
// an over-long run inside the block
const uint8_t kDezigzag[64 + 15] = {
    0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13,
    6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31,
    39, 46, 53, 60, 61, 54, 47, 55, 62, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63};

struct JHuff {
    // decoding per ITU T.81 F.2.2.3: for each length, the largest code and where its
    // symbols start
    int32_t maxcode[18];
    int32_t valptr[17];
    int32_t mincode[17];
    uint8_t vals[256];
    bool present = false;
};

struct JComp {
    int id, h, v, tq;
    int hd = 0, ha = 0;  // Huffman table selectors of the current scan
    int x, y, w2, h2;
    int dc_pred = 0;
    std::vector<uint8_t> data;
};

struct Jpeg {
    const uint8_t* d;
    size_t n, pos = 0;
    // entropy-coded segment bit reader
    uint32_t code_buf = 0;
    int code_bits = 0;
    int marker = -1;  // marker met inside the entropy-coded data (-1: none)
    JHuff dc[4], ac[4];
    uint16_t dq[4][64];
    std::vector<JComp> comp;
    int img_x = 0, img_y = 0, hmax = 1, vmax = 1, mcux = 0, mcuy = 0;
    int restart = 0;
    bool jfif = false;
    int app14 = -1;
    bool frame = false;

    int get8() {
        if (pos >= n) fail(C2_ERR_PARSE, "JPEG: unexpected end of file");
        return d[pos++];
    }
    int get16() {
        int hi = get8();
        return hi << 8 | get8();
    }
    void fill() {
        while (code_bits <= 24) {
            int c = 0;
            if (marker < 0 && pos < n) {
                c = d[pos++];
                if (c == 0xFF) {
                    int c2 = pos < n ? d[pos] : 0;
                    while (c2 == 0xFF && pos + 1 < n) c2 = d[++pos];  // fill bytes
                    if (c2 == 0) {
                        pos++;
                    } else {
                        marker = c2;
                        pos++;
                        c = 0;
                        code_buf |= 0u;
                        code_bits += 8;
                        continue;
                    }
                }
            }
            code_buf |= (uint32_t)c << (24 - code_bits);
            code_bits += 8;
        }
    }
    int bits(int k) {  // MSB first
        if (k == 0) return 0;
        if (code_bits < k) fill();
        int v = (int)(code_buf >> (32 - k));
        code_buf <<= k;
        code_bits -= k;
        return v;
    }
    int decode(const JHuff& t) {
        if (!t.present) fail(C2_ERR_PARSE, "JPEG: missing Huffman table");
        if (code_bits < 16) fill();
        int code = 0;
        for (int l = 1; l <= 16; l++) {
            code = code << 1 | (int)(code_buf >> 31);
            code_buf <<= 1;
            code_bits--;
            if (code <= t.maxcode[l]) return t.vals[t.valptr[l] + code - t.mincode[l]];
        }
        fail(C2_ERR_PARSE, "JPEG: bad Huffman code");
    }
    // RECEIVE + EXTEND (T.81 F.2.2.1)
    int receive_extend(int s) {
        int v = bits(s);
        if (s && v < (1 << (s - 1))) v += (int)(~0u << s) + 1;
        return v;
    }
    void reset_bits() {
        code_buf = 0;
        code_bits = 0;
        marker = -1;
        for (auto& c : comp) c.dc_pred = 0;
    }
};

void build_jhuff(JHuff& t, const uint8_t counts[16], const uint8_t* vals, int nvals) {
    std::memcpy(t.vals, vals, (size_t)nvals);
    int code = 0, k = 0;
    for (int l = 1; l <= 16; l++) {
        t.valptr[l] = k;
        t.mincode[l] = code;
        code += counts[l - 1];
        k += counts[l - 1];
        t.maxcode[l] = counts[l - 1] ? code - 1 : -1;
        if (code > (1 << l)) fail(C2_ERR_PARSE, "JPEG: bad Huffman table");
        code <<= 1;
    }
    t.maxcode[17] = 0x7FFFFFFF;
    t.present = true;
}

constexpr int f2f(float x) { return (int)((double)(x * 4096) + 0.5); }  // 12-bit constant

// One 8-point pass of the islow IDCT (jidctint's even/odd decomposition); a..h are the
// inputs in natural order, x0..x3 / t0..t3 the even/odd halves.
struct Idct1 {
    int t0, t1, t2, t3, x0, x1, x2, x3;
    Idct1(int s0, int s1, int s2, int s3, int s4, int s5, int s6, int s7) {
        int p2 = s2, p3 = s6;
        int p1 = (p2 + p3) * f2f(0.5411961f);
        int e2 = p1 + p3 * f2f(-1.847759065f);
        int e3 = p1 + p2 * f2f(0.765366865f);
        p2 = s0;
        p3 = s4;
        int e0 = (p2 + p3) * 4096;
        int e1 = (p2 - p3) * 4096;
        x0 = e0 + e3;
        x3 = e0 - e3;
        x1 = e1 + e2;
        x2 = e1 - e2;
        int o0 = s7, o1 = s5, o2 = s3, o3 = s1;
        int q3 = o0 + o2, q4 = o1 + o3, q1 = o0 + o3, q2 = o1 + o2;
        int p5 = (q3 + q4) * f2f(1.175875602f);
        o0 = o0 * f2f(0.298631336f);
        o1 = o1 * f2f(2.053119869f);
        o2 = o2 * f2f(3.072711026f);
        o3 = o3 * f2f(1.501321110f);
        q1 = p5 + q1 * f2f(-0.899976223f);
        q2 = p5 + q2 * f2f(-2.562915447f);
        q3 = q3 * f2f(-1.961570560f);
        q4 = q4 * f2f(-0.390180644f);
        t3 = o3 + q1 + q4;
        t2 = o2 + q2 + q3;
        t1 = o1 + q2 + q4;
        t0 = o0 + q1 + q3;
    }
};

uint8_t clamp8(int x) { return (uint8_t)(x < 0 ? 0 : x > 255 ? 255 : x); }

void idct_block(uint8_t* out, int stride, const int16_t* in) {
    int v[64];
    for (int i = 0; i < 8; i++) {
        const int16_t* dcol = in + i;
        if (!dcol[8] && !dcol[16] && !dcol[24] && !dcol[32] && !dcol[40] && !dcol[48] && !dcol[56]) {
            int dc = dcol[0] * 4;
            for (int r = 0; r < 8; r++) v[r * 8 + i] = dc;
            continue;
        }
        Idct1 k(dcol[0], dcol[8], dcol[16], dcol[24], dcol[32], dcol[40], dcol[48], dcol[56]);
        // constants carry 12 fraction bits: keep 2 of them for the row pass
        int x0 = k.x0 + 512, x1 = k.x1 + 512, x2 = k.x2 + 512, x3 = k.x3 + 512;
        v[0 * 8 + i] = (x0 + k.t3) >> 10;
        v[7 * 8 + i] = (x0 - k.t3) >> 10;
        v[1 * 8 + i] = (x1 + k.t2) >> 10;
        v[6 * 8 + i] = (x1 - k.t2) >> 10;
        v[2 * 8 + i] = (x2 + k.t1) >> 10;
        v[5 * 8 + i] = (x2 - k.t1) >> 10;
        v[3 * 8 + i] = (x3 + k.t0) >> 10;
        v[4 * 8 + i] = (x3 - k.t0) >> 10;
    }
    for (int r = 0; r < 8; r++) {
        const int* s = v + r * 8;
        Idct1 k(s[0], s[1], s[2], s[3], s[4], s[5], s[6], s[7]);
        // 12 + 2 fraction bits and the two sqrt(8) scales: 17 bits to drop, rounded, with
        // the +128 level shift added before the shift
        const int bias = 65536 + (128 << 17);
        int x0 = k.x0 + bias, x1 = k.x1 + bias, x2 = k.x2 + bias, x3 = k.x3 + bias;
        uint8_t* o = out + (size_t)r * stride;
        o[0] = clamp8((x0 + k.t3) >> 17);
        o[7] = clamp8((x0 - k.t3) >> 17);
        o[1] = clamp8((x1 + k.t2) >> 17);
        o[6] = clamp8((x1 - k.t2) >> 17);
        o[2] = clamp8((x2 + k.t1) >> 17);
        o[5] = clamp8((x2 - k.t1) >> 17);
        o[3] = clamp8((x3 + k.t0) >> 17);
        o[4] = clamp8((x3 - k.t0) >> 17);
    }
}

void decode_block(Jpeg& j, int16_t blk[64], JComp& c) {
    std::memset(blk, 0, 64 * sizeof(int16_t));
    const uint16_t* q = j.dq[c.tq];
    int t = j.decode(j.dc[c.hd]);
    if (t > 16) fail(C2_ERR_PARSE, "JPEG: bad DC code");
    int diff = t ? j.receive_extend(t) : 0;
    int dcv = c.dc_pred + diff;
    c.dc_pred = dcv;
    blk[0] = (int16_t)(dcv * q[0]);
    int k = 1;
    do {
        int rs = j.decode(j.ac[c.ha]);
        int s = rs & 15, r = rs >> 4;
        if (s == 0) {
            if (rs != 0xF0) break;  // end of block
            k += 16;
        } else {
            k += r;
            int zig = kDezigzag[k++];
            blk[zig] = (int16_t)(j.receive_extend(s) * q[zig]);
        }
    } while (k < 64);
}

void decode_scan(Jpeg& j, const std::vector<int>& sc) {
    j.reset_bits();
    int todo = j.restart ? j.restart : 0x7FFFFFFF;
    int16_t blk[64];
    auto restart_check = [&]() {
        if (--todo <= 0) {
            if (j.code_bits < 24) j.fill();
            if (!(j.marker >= 0xD0 && j.marker <= 0xD7)) return false;
            j.reset_bits();
            todo = j.restart ? j.restart : 0x7FFFFFFF;
        }
        return true;
    };
    if (sc.size() == 1) {
        JComp& c = j.comp[sc[0]];
        int bw = (c.x + 7) >> 3, bh = (c.y + 7) >> 3;
        for (int by = 0; by < bh; by++)
            for (int bx = 0; bx < bw; bx++) {
                decode_block(j, blk, c);
                idct_block(&c.data[(size_t)c.w2 * by * 8 + bx * 8], c.w2, blk);
                if (!restart_check()) return;
            }
        return;
    }
    for (int my = 0; my < j.mcuy; my++)
        for (int mx = 0; mx < j.mcux; mx++) {
            for (int ci : sc) {
                JComp& c = j.comp[ci];
                for (int yy = 0; yy < c.v; yy++)
                    for (int xx = 0; xx < c.h; xx++) {
                        int bx = mx * c.h + xx, by = my * c.v + yy;
                        decode_block(j, blk, c);
                        idct_block(&c.data[(size_t)c.w2 * by * 8 + bx * 8], c.w2, blk);
                    }
            }
            if (!restart_check()) return;
        }
}

// --- upsampling of one output row from the component's near and far input rows
const uint8_t* up_1(uint8_t*, const uint8_t* nr, const uint8_t*, int) { return nr; }
const uint8_t* up_v2(uint8_t* out, const uint8_t* nr, const uint8_t* fr, int w) {
    for (int i = 0; i < w; i++) out[i] = (uint8_t)((3 * nr[i] + fr[i] + 2) >> 2);
    return out;
}
const uint8_t* up_h2(uint8_t* out, const uint8_t* in, const uint8_t*, int w) {
    if (w == 1) {
        out[0] = out[1] = in[0];
        return out;
    }
    out[0] = in[0];
    out[1] = (uint8_t)((in[0] * 3 + in[1] + 2) >> 2);
    int i;
    for (i = 1; i < w - 1; i++) {
        int c = 3 * in[i] + 2;
        out[i * 2] = (uint8_t)((c + in[i - 1]) >> 2);
        out[i * 2 + 1] = (uint8_t)((c + in[i + 1]) >> 2);
    }
    out[i * 2] = (uint8_t)((in[w - 2] * 3 + in[w - 1] + 2) >> 2);
    out[i * 2 + 1] = in[w - 1];
    return out;
}
const uint8_t* up_hv2(uint8_t* out, const uint8_t* nr, const uint8_t* fr, int w) {
    if (w == 1) {
        out[0] = out[1] = (uint8_t)((3 * nr[0] + fr[0] + 2) >> 2);
        return out;
    }
    int t1 = 3 * nr[0] + fr[0];
    out[0] = (uint8_t)((t1 + 2) >> 2);
    for (int i = 1; i < w; i++) {
        int t0 = t1;
        t1 = 3 * nr[i] + fr[i];
        out[i * 2 - 1] = (uint8_t)((3 * t0 + t1 + 8) >> 4);
        out[i * 2] = (uint8_t)((3 * t1 + t0 + 8) >> 4);
    }
    out[w * 2 - 1] = (uint8_t)((t1 + 2) >> 2);
    return out;
}

constexpr int fx20(float x) { return ((int)(x * 4096.0f + 0.5f)) << 8; }  // 20-bit constant

void ycc_to_rgb(uint8_t* o, int Y, int cb, int cr) {
    int yf = (Y << 20) + (1 << 19);
    cr -= 128;
    cb -= 128;
    int r = yf + cr * fx20(1.40200f);
    int g = yf + (cr * -fx20(0.71414f)) + (int)((unsigned)(cb * -fx20(0.34414f)) & 0xFFFF0000u);
    int b = yf + cb * fx20(1.77200f);
    o[0] = clamp8(r >> 20);
    o[1] = clamp8(g >> 20);
    o[2] = clamp8(b >> 20);
}

uint8_t blinn8(uint8_t x, uint8_t y) {
    unsigned t = (unsigned)x * y + 128;
    return (uint8_t)((t + (t >> 8)) >> 8);
}

ImageRGB decode_jpeg(const uint8_t* d, size_t n) {
    Jpeg j;
    j.d = d;
    j.n = n;
    j.pos = 2;
    for (auto& t : j.dq)
        for (auto& q : t) q = 0;
    bool done = false;
    int m = -1;
    auto next_marker = [&]() {
        int c = j.get8();
        while (c != 0xFF) c = j.get8();  // skip garbage before a marker
        while (c == 0xFF) c = j.get8();
        return c;
    };
    m = next_marker();
    while (!done) {
        if (m == 0xD9) break;  // EOI
        if (m == 0xC2 || m == 0xC6 || m == 0xCA || m == 0xCE)
            fail(C2_ERR_UNSUPPORTED, "JPEG: progressive images are not supported");
        if ((m >= 0xC3 && m <= 0xCF && m != 0xC4 && m != 0xC8 && m != 0xCC))
            fail(C2_ERR_UNSUPPORTED, "JPEG: lossless or arithmetic-coded images are not supported");
        if (m >= 0xD0 && m <= 0xD7) {  // stray restart marker
            m = next_marker();
            continue;
        }
        if (m == 0xDA) {  // SOS
            if (!j.frame) fail(C2_ERR_PARSE, "JPEG: scan before frame header");
            int len = j.get16();
            int ns = j.get8();
            if (ns < 1 || ns > 4 || ns > (int)j.comp.size() || len != 6 + 2 * ns) fail(C2_ERR_PARSE, "JPEG: bad SOS");
            std::vector<int> sc;
            for (int i = 0; i < ns; i++) {
                int id = j.get8(), tt = j.get8();
                int which = -1;
                for (size_t k = 0; k < j.comp.size(); k++)
                    if (j.comp[k].id == id) which = (int)k;
                if (which < 0) fail(C2_ERR_PARSE, "JPEG: scan names an unknown component");
                j.comp[which].hd = tt >> 4;
                j.comp[which].ha = tt & 15;
                if (j.comp[which].hd > 3 || j.comp[which].ha > 3) fail(C2_ERR_PARSE, "JPEG: bad Huffman selector");
                sc.push_back(which);
            }
            int ss = j.get8(), se = j.get8(), a = j.get8();
            if (ss != 0 || se != 63 || a != 0) fail(C2_ERR_PARSE, "JPEG: bad spectral selection in a baseline scan");
            decode_scan(j, sc);
            if (j.marker >= 0 && !(j.marker >= 0xD0 && j.marker <= 0xD7)) {
                m = j.marker;
            } else {
                // find the marker after the entropy-coded data
                m = -1;
                while (j.pos < j.n) {
                    int c = j.d[j.pos++];
                    if (c == 0xFF && j.pos < j.n) {
                        int c2 = j.d[j.pos];
                        if (c2 != 0 && !(c2 >= 0xD0 && c2 <= 0xD7) && c2 != 0xFF) {
                            j.pos++;
                            m = c2;
                            break;
                        }
                    }
                }
                if (m < 0) break;  // data ends without EOI
            }
            continue;
        }
        int len = j.get16();
        if (len < 2 || j.pos + (size_t)(len - 2) > j.n) fail(C2_ERR_PARSE, "JPEG: truncated marker segment");
        size_t end = j.pos + (size_t)(len - 2);
        if (m == 0xDB) {  // DQT
            while (j.pos < end) {
                int pq = j.get8();
                int p = pq >> 4, t = pq & 15;
                if (p > 1 || t > 3) fail(C2_ERR_PARSE, "JPEG: bad DQT");
                for (int i = 0; i < 64; i++) j.dq[t][kDezigzag[i]] = (uint16_t)(p ? j.get16() : j.get8());
            }
        } else if (m == 0xC4) {  // DHT
            while (j.pos < end) {
                int tc = j.get8();
                int cls = tc >> 4, id = tc & 15;
                if (cls > 1 || id > 3) fail(C2_ERR_PARSE, "JPEG: bad DHT");
                uint8_t counts[16];
                int total = 0;
                for (int i = 0; i < 16; i++) total += counts[i] = (uint8_t)j.get8();
                if (total > 256 || j.pos + (size_t)total > end) fail(C2_ERR_PARSE, "JPEG: bad DHT");
                build_jhuff(cls ? j.ac[id] : j.dc[id], counts, j.d + j.pos, total);
                j.pos += (size_t)total;
            }
        } else if (m == 0xDD) {  // DRI
            j.restart = j.get16();
        } else if (m == 0xC0 || m == 0xC1) {  // SOF0 / SOF1
            if (j.get8() != 8) fail(C2_ERR_UNSUPPORTED, "JPEG: only 8-bit samples are supported");
            j.img_y = j.get16();
            j.img_x = j.get16();
            int nc = j.get8();
            if (j.img_y == 0) fail(C2_ERR_UNSUPPORTED, "JPEG: images with a DNL height are not supported");
            if (j.img_x == 0 || (nc != 1 && nc != 3 && nc != 4)) fail(C2_ERR_PARSE, "JPEG: bad frame header");
            j.comp.assign(nc, JComp{});
            for (auto& c : j.comp) {
                c.id = j.get8();
                int hv = j.get8();
                c.h = hv >> 4;
                c.v = hv & 15;
                c.tq = j.get8();
                if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4 || c.tq > 3) fail(C2_ERR_PARSE, "JPEG: bad component");
                j.hmax = std::max(j.hmax, c.h);
                j.vmax = std::max(j.vmax, c.v);
            }
            // every component's sampling factors must divide the largest: the upsamplers
            // expand a component row by hmax / h (and vmax / v), which must give img_x bytes
            for (auto& c : j.comp)
                if (j.hmax % c.h || j.vmax % c.v) fail(C2_ERR_PARSE, "JPEG: sampling factors do not divide");
            j.mcux = (j.img_x + j.hmax * 8 - 1) / (j.hmax * 8);
            j.mcuy = (j.img_y + j.vmax * 8 - 1) / (j.vmax * 8);
            for (auto& c : j.comp) {
                c.x = (j.img_x * c.h + j.hmax - 1) / j.hmax;
                c.y = (j.img_y * c.v + j.vmax - 1) / j.vmax;
                c.w2 = j.mcux * c.h * 8;
                c.h2 = j.mcuy * c.v * 8;
                c.data.assign((size_t)c.w2 * c.h2, 0);
            }
            j.frame = true;
        } else if (m == 0xE0) {  // APP0: JFIF
            if (len >= 7 && !std::memcmp(j.d + j.pos, "JFIF\0", 5)) j.jfif = true;
        } else if (m == 0xEE) {  // APP14: Adobe colour transform
            if (len >= 14 && !std::memcmp(j.d + j.pos, "Adobe\0", 6)) j.app14 = j.d[j.pos + 11];
        }
        j.pos = end;
        m = next_marker();
    }
    if (!j.frame) fail(C2_ERR_PARSE, "JPEG: no frame header");

    const int nc = (int)j.comp.size();
    int rgbIds = 0;
    if (nc == 3) {
        static const int ids[3] = {'R', 'G', 'B'};
        for (int k = 0; k < 3; k++)
            if (j.comp[k].id == ids[k]) rgbIds++;
    }
    const bool is_rgb = nc == 3 && (rgbIds == 3 || (j.app14 == 0 && !j.jfif));
    struct Res {
        const uint8_t* (*fn)(uint8_t*, const uint8_t*, const uint8_t*, int);
        int hs, vs, ystep, ypos, wl;
        const uint8_t *line0, *line1;
        std::vector<uint8_t> buf;
    };
    std::vector<Res> rs(nc);
    for (int k = 0; k < nc; k++) {
        Res& r = rs[k];
        r.hs = j.hmax / j.comp[k].h;
        r.vs = j.vmax / j.comp[k].v;
        r.ystep = r.vs >> 1;
        r.ypos = 0;
        r.wl = (j.img_x + r.hs - 1) / r.hs;
        r.line0 = r.line1 = j.comp[k].data.data();
        r.buf.assign((size_t)j.img_x + 3 + (size_t)r.hs * 4, 0);
        if (r.hs == 1 && r.vs == 1) r.fn = up_1;
        else if (r.hs == 1 && r.vs == 2) r.fn = up_v2;
        else if (r.hs == 2 && r.vs == 1) r.fn = up_h2;
        else if (r.hs == 2 && r.vs == 2) r.fn = up_hv2;
        else r.fn = nullptr;  // nearest
    }
    ImageRGB img;
    img.width = j.img_x;
    img.height = j.img_y;
    img.rgb.assign((size_t)j.img_x * j.img_y * 3, 0);
    const uint8_t* co[4] = {nullptr, nullptr, nullptr, nullptr};
    for (int y = 0; y < j.img_y; y++) {
        for (int k = 0; k < nc; k++) {
            Res& r = rs[k];
            bool bot = r.ystep >= (r.vs >> 1);
            const uint8_t* nr = bot ? r.line1 : r.line0;
            const uint8_t* fr = bot ? r.line0 : r.line1;
            if (r.fn) {
                co[k] = r.fn(r.buf.data(), nr, fr, r.wl);
            } else {
                for (int i = 0; i < r.wl; i++)
                    for (int t = 0; t < r.hs; t++) r.buf[(size_t)i * r.hs + t] = nr[i];
                co[k] = r.buf.data();
            }
            if (++r.ystep >= r.vs) {
                r.ystep = 0;
                r.line0 = r.line1;
                if (++r.ypos < j.comp[k].y) r.line1 += j.comp[k].w2;
            }
        }
        uint8_t* out = &img.rgb[(size_t)y * j.img_x * 3];
        for (int i = 0; i < j.img_x; i++, out += 3) {
            if (nc == 3) {
                if (is_rgb) {
                    out[0] = co[0][i];
                    out[1] = co[1][i];
                    out[2] = co[2][i];
                } else {
                    ycc_to_rgb(out, co[0][i], co[1][i], co[2][i]);
                }
            } else if (nc == 4) {
                uint8_t k4 = co[3][i];
                if (j.app14 == 0) {  // CMYK
                    out[0] = blinn8(co[0][i], k4);
                    out[1] = blinn8(co[1][i], k4);
                    out[2] = blinn8(co[2][i], k4);
                } else if (j.app14 == 2) {  // YCCK
                    ycc_to_rgb(out, co[0][i], co[1][i], co[2][i]);
                    out[0] = blinn8((uint8_t)(255 - out[0]), k4);
                    out[1] = blinn8((uint8_t)(255 - out[1]), k4);
                    out[2] = blinn8((uint8_t)(255 - out[2]), k4);
                } else {  // YCbCr plus an ignored fourth channel
                    ycc_to_rgb(out, co[0][i], co[1][i], co[2][i]);
                }
            } else {
                out[0] = out[1] = out[2] = co[0][i];
            }
        }
    }
    return img;
}

}  // namespace

ImageRGB decode_image(const uint8_t* data, size_t size) {
    static const uint8_t png_sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (size >= 8 && !std::memcmp(data, png_sig, 8)) return decode_png(data, size);
    if (size >= 3 && data[0] == 0xFF && data[1] == 0xD8) return decode_jpeg(data, size);
    fail(C2_ERR_UNSUPPORTED, "image format not supported (PNG and baseline JPEG are)");
}

ImageRGB load_image(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) fail(C2_ERR_IO, "cannot open image " + path);
    std::vector<uint8_t> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    try {
        return decode_image(bytes.data(), bytes.size());
    } catch (const Error& e) {
        fail(e.code, path + ": " + e.msg);
    }
}

}  // namespace c2
