// Host side of the C-ABI: Crystal2's scene/config formats and the Halton tables.
//
//  * JSON: a parser with the reference JSON library's semantics
//    (deps/TREngine/deps/SJson/SJson.cpp): UTF-8 BOM skipped (:458-459),
//    numbers without fraction/exponent are integers, others accumulate in
//    double as v = v*10 + d, v += d/10^k, v *= 10^e (:185-231, not strtod),
//    object members kept in key order (std::map), errors carry line:col.
//  * Scene assembly in Scene::CreateScene order (src/Core/Scene.cpp:23-199):
//    entities in JSON object order, one area light per triangle (:109-114),
//    lights in creation order.
//  * OBJ: src/Loaders/ObjLoader.cpp:78-208 (v / vn / vt / triangular f with
//    v/vt/vn indices, unshared corners, doubles rounded to float).
//  * Mesh transforms in glm's exact op order (translate / rotate / scale,
//    mat4 * vec4, mat4 inverse), src/Shapes/TriangleMesh.cpp:9-51.
#include "c2_host.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <random>
#include <sstream>

#include "c2_image.h"
#include "c2_vec.h"

namespace c2 {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }

void fail(int code, const std::string& msg) { throw Error{code, msg}; }

// --------------------------------------------------------------------------------------- JSON

struct JNode {
    enum Type { Null, Bool, Int, Float, String, Array, Object } type = Null;
    bool b = false;
    long long i = 0;
    double f = 0.0;
    std::string s;
    std::vector<JNode> arr;
    std::map<std::string, JNode> obj;

    const JNode& member(const std::string& k) const {
        if (type != Object) fail(C2_ERR_PARSE, "JSON: member '" + k + "' requested from a non-object");
        auto it = obj.find(k);
        if (it == obj.end()) fail(C2_ERR_PARSE, "JSON: missing member '" + k + "'");
        return it->second;
    }
    bool has(const std::string& k) const { return type == Object && obj.count(k) != 0; }
    double as_float() const {
        if (type == Int) return (double)i;
        if (type == Float) return f;
        fail(C2_ERR_PARSE, "JSON: value is not a number");
    }
    long long as_int() const {
        if (type == Int) return i;
        if (type == Float) return (long long)f;
        fail(C2_ERR_PARSE, "JSON: value is not a number");
    }
    const std::string& as_string() const {
        if (type != String) fail(C2_ERR_PARSE, "JSON: value is not a string");
        return s;
    }
    const JNode& at(size_t k) const {
        if (type != Array || k >= arr.size()) fail(C2_ERR_PARSE, "JSON: array index out of range");
        return arr[k];
    }
};

class JParser {
public:
    explicit JParser(const std::string& text) : t_(text) {}
    JNode parse() {
        if (t_.size() >= 3 && (unsigned char)t_[0] == 0xEF) p_ = 3;
        line_start_ = p_;
        ws();
        JNode v = value();
        ws();
        if (p_ != t_.size()) err("root not singular");
        return v;
    }

private:
    const std::string& t_;
    size_t p_ = 0, line_start_ = 0;
    int line_ = 1;

    [[noreturn]] void err(const std::string& what) {
        fail(C2_ERR_PARSE, "JSON " + std::to_string(line_) + ":" + std::to_string(p_ - line_start_) + ": " + what);
    }
    void ws() {
        while (p_ < t_.size() && (t_[p_] == ' ' || t_[p_] == '\t' || t_[p_] == '\r' || t_[p_] == '\n')) {
            if (t_[p_] == '\n') {
                line_++;
                line_start_ = p_ + 1;
            }
            p_++;
        }
    }
    void expect(char c) {
        if (p_ >= t_.size() || t_[p_] != c) err(std::string("expected '") + c + "'");
        p_++;
    }
    JNode value() {
        if (p_ >= t_.size()) err("expected value");
        char c = t_[p_];
        if (c == '{') return object();
        if (c == '[') return array();
        if (c == '"') {
            JNode n;
            n.type = JNode::String;
            n.s = str();
            return n;
        }
        if (t_.compare(p_, 4, "true") == 0) { p_ += 4; JNode n; n.type = JNode::Bool; n.b = true; return n; }
        if (t_.compare(p_, 5, "false") == 0) { p_ += 5; JNode n; n.type = JNode::Bool; return n; }
        if (t_.compare(p_, 4, "null") == 0) { p_ += 4; return JNode(); }
        if (c == '-' || (c >= '0' && c <= '9') || c == '.') return number();
        err("unexpected token");
    }
    std::string str() {
        expect('"');
        std::string out;
        while (true) {
            if (p_ >= t_.size()) err("unterminated string");
            char c = t_[p_++];
            if (c == '"') break;
            if (c == '\\') {
                if (p_ >= t_.size()) err("bad escape");
                char e = t_[p_++];
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    default: out += e; break;
                }
            } else {
                out += c;
            }
        }
        return out;
    }
    // SJson.cpp:81-232
    JNode number() {
        bool neg = false;
        if (t_[p_] == '-') { neg = true; p_++; }
        std::string ip, fp, ep;
        bool expNeg = false;
        int stage = 0;
        bool signChecked = false;
        while (p_ < t_.size()) {
            char c = t_[p_];
            if (stage == 0) {
                if (isdigit((unsigned char)c)) ip += c;
                else if (c == '.') stage = 1;
                else if (tolower(c) == 'e') stage = 2;
                else break;
                p_++;
            } else if (stage == 1) {
                if (isdigit((unsigned char)c)) fp += c;
                else if (c == '.') err("invalid value");
                else if (tolower(c) == 'e') stage = 2;
                else break;
                p_++;
            } else {
                if (!signChecked) {
                    if (c == '-') { expNeg = true; p_++; }
                    else if (c == '+') { p_++; }
                    signChecked = true;
                    continue;
                }
                if (isdigit((unsigned char)c)) ep += c;
                else if (c == '.' || tolower(c) == 'e') err("invalid value");
                else break;
                p_++;
            }
        }
        if (ip.empty() && fp.empty() && ep.empty()) err("invalid value");
        if (!ip.empty() && ip[0] == '0' && ip.size() > 1) err("invalid value");
        JNode n;
        if (!ip.empty() && fp.empty() && ep.empty()) {
            long long v = 0;
            for (char d : ip) v = v * 10 + (d - '0');
            n.type = JNode::Int;
            n.i = neg ? -v : v;
            return n;
        }
        double v = 0;
        for (char d : ip) { v *= 10.0; v += d - '0'; }
        if (!fp.empty()) {
            double tt = 10;
            for (char d : fp) { v += (d - '0') / tt; tt *= 10; }
        }
        if (!ep.empty()) {
            long long e = 0;
            for (char d : ep) e = e * 10 + (d - '0');
            v *= std::pow(10.0, (double)(expNeg ? -e : e));
        }
        n.type = JNode::Float;
        n.f = neg ? -v : v;
        return n;
    }
    JNode array() {
        expect('[');
        JNode n;
        n.type = JNode::Array;
        ws();
        if (p_ < t_.size() && t_[p_] == ']') { p_++; return n; }
        while (true) {
            ws();
            n.arr.push_back(value());
            ws();
            if (p_ < t_.size() && t_[p_] == ',') { p_++; continue; }
            expect(']');
            break;
        }
        return n;
    }
    JNode object() {
        expect('{');
        JNode n;
        n.type = JNode::Object;
        ws();
        if (p_ < t_.size() && t_[p_] == '}') { p_++; return n; }
        while (true) {
            ws();
            std::string k = str();
            ws();
            expect(':');
            ws();
            n.obj[k] = value();
            ws();
            if (p_ < t_.size() && t_[p_] == ',') { p_++; continue; }
            expect('}');
            break;
        }
        return n;
    }
};

static std::string read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) fail(C2_ERR_IO, "cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

static JNode load_json(const std::string& path) {
    std::string text = read_file(path);
    try {
        return JParser(text).parse();
    } catch (const Error& e) {
        fail(e.code, path + ": " + e.msg);
    }
}

// parse_vec3 / parse_vec2 (src/Loaders/JsonLoader.h:12-25): doubles rounded to float.
static V3 jvec3(const JNode& n) {
    if (n.type != JNode::Array) fail(C2_ERR_PARSE, "JSON: expected a 3-vector");
    return V3{(float)n.at(0).as_float(), (float)n.at(1).as_float(), (float)n.at(2).as_float()};
}
static V2 jvec2(const JNode& n) {
    if (n.type != JNode::Array) fail(C2_ERR_PARSE, "JSON: expected a 2-vector");
    return V2{(float)n.at(0).as_float(), (float)n.at(1).as_float()};
}

// ------------------------------------------------------------------------------- glm mat4 ops
// Column-major 4x4: m[col][row], as glm.
struct M4 {
    float m[4][4];
};
struct V4 {
    float v[4];
};
static V4 v4_add(const V4& a, const V4& b) { return V4{{a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2], a.v[3] + b.v[3]}}; }
static V4 v4_sub(const V4& a, const V4& b) { return V4{{a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2], a.v[3] - b.v[3]}}; }
static V4 v4_mul(const V4& a, const V4& b) { return V4{{a.v[0] * b.v[0], a.v[1] * b.v[1], a.v[2] * b.v[2], a.v[3] * b.v[3]}}; }
static V4 v4_muls(const V4& a, float s) { return V4{{a.v[0] * s, a.v[1] * s, a.v[2] * s, a.v[3] * s}}; }
static V4 col(const M4& m, int c) { return V4{{m.m[c][0], m.m[c][1], m.m[c][2], m.m[c][3]}}; }
static void setcol(M4& m, int c, const V4& v) { for (int r = 0; r < 4; r++) m.m[c][r] = v.v[r]; }
static M4 m4_identity() {
    M4 r{};
    for (int i = 0; i < 4; i++) r.m[i][i] = 1.f;
    return r;
}
// glm/ext/matrix_transform.inl:10-16
static M4 m4_translate(const M4& m, V3 v) {
    M4 r = m;
    V4 t = v4_add(v4_add(v4_add(v4_muls(col(m, 0), v.x), v4_muls(col(m, 1), v.y)), v4_muls(col(m, 2), v.z)), col(m, 3));
    setcol(r, 3, t);
    return r;
}
// glm/ext/matrix_transform.inl:18-46 (cos/sin: the float libm functions glm calls)
static M4 m4_rotate(const M4& m, float a, V3 axis_in) {
    float c = std::cos(a), s = std::sin(a);
    V3 axis = normalize(axis_in);
    V3 temp = (1.f - c) * axis;
    float R[3][3];
    R[0][0] = c + temp.x * axis.x;
    R[0][1] = temp.x * axis.y + s * axis.z;
    R[0][2] = temp.x * axis.z - s * axis.y;
    R[1][0] = temp.y * axis.x - s * axis.z;
    R[1][1] = c + temp.y * axis.y;
    R[1][2] = temp.y * axis.z + s * axis.x;
    R[2][0] = temp.z * axis.x + s * axis.y;
    R[2][1] = temp.z * axis.y - s * axis.x;
    R[2][2] = c + temp.z * axis.z;
    M4 r;
    for (int k = 0; k < 3; k++) {
        V4 v = v4_add(v4_add(v4_muls(col(m, 0), R[k][0]), v4_muls(col(m, 1), R[k][1])), v4_muls(col(m, 2), R[k][2]));
        setcol(r, k, v);
    }
    setcol(r, 3, col(m, 3));
    return r;
}
// glm/ext/matrix_transform.inl:78-86
static M4 m4_scale(const M4& m, V3 v) {
    M4 r;
    setcol(r, 0, v4_muls(col(m, 0), v.x));
    setcol(r, 1, v4_muls(col(m, 1), v.y));
    setcol(r, 2, v4_muls(col(m, 2), v.z));
    setcol(r, 3, col(m, 3));
    return r;
}
// glm/detail/type_mat4x4.inl:536-573
static V4 m4_mulv(const M4& m, const V4& v) {
    V4 a0 = v4_add(v4_muls(col(m, 0), v.v[0]), v4_muls(col(m, 1), v.v[1]));
    V4 a1 = v4_add(v4_muls(col(m, 2), v.v[2]), v4_muls(col(m, 3), v.v[3]));
    return v4_add(a0, a1);
}
static M4 m4_transpose(const M4& m) {
    M4 r;
    for (int c = 0; c < 4; c++)
        for (int rr = 0; rr < 4; rr++) r.m[c][rr] = m.m[rr][c];
    return r;
}
// glm/detail/func_matrix.inl:293-352
static M4 m4_inverse(const M4& M) {
    auto m = [&](int c, int r) { return M.m[c][r]; };
    float Coef00 = m(2, 2) * m(3, 3) - m(3, 2) * m(2, 3);
    float Coef02 = m(1, 2) * m(3, 3) - m(3, 2) * m(1, 3);
    float Coef03 = m(1, 2) * m(2, 3) - m(2, 2) * m(1, 3);
    float Coef04 = m(2, 1) * m(3, 3) - m(3, 1) * m(2, 3);
    float Coef06 = m(1, 1) * m(3, 3) - m(3, 1) * m(1, 3);
    float Coef07 = m(1, 1) * m(2, 3) - m(2, 1) * m(1, 3);
    float Coef08 = m(2, 1) * m(3, 2) - m(3, 1) * m(2, 2);
    float Coef10 = m(1, 1) * m(3, 2) - m(3, 1) * m(1, 2);
    float Coef11 = m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2);
    float Coef12 = m(2, 0) * m(3, 3) - m(3, 0) * m(2, 3);
    float Coef14 = m(1, 0) * m(3, 3) - m(3, 0) * m(1, 3);
    float Coef15 = m(1, 0) * m(2, 3) - m(2, 0) * m(1, 3);
    float Coef16 = m(2, 0) * m(3, 2) - m(3, 0) * m(2, 2);
    float Coef18 = m(1, 0) * m(3, 2) - m(3, 0) * m(1, 2);
    float Coef19 = m(1, 0) * m(2, 2) - m(2, 0) * m(1, 2);
    float Coef20 = m(2, 0) * m(3, 1) - m(3, 0) * m(2, 1);
    float Coef22 = m(1, 0) * m(3, 1) - m(3, 0) * m(1, 1);
    float Coef23 = m(1, 0) * m(2, 1) - m(2, 0) * m(1, 1);
    V4 Fac0{{Coef00, Coef00, Coef02, Coef03}}, Fac1{{Coef04, Coef04, Coef06, Coef07}};
    V4 Fac2{{Coef08, Coef08, Coef10, Coef11}}, Fac3{{Coef12, Coef12, Coef14, Coef15}};
    V4 Fac4{{Coef16, Coef16, Coef18, Coef19}}, Fac5{{Coef20, Coef20, Coef22, Coef23}};
    V4 Vec0{{m(1, 0), m(0, 0), m(0, 0), m(0, 0)}}, Vec1{{m(1, 1), m(0, 1), m(0, 1), m(0, 1)}};
    V4 Vec2{{m(1, 2), m(0, 2), m(0, 2), m(0, 2)}}, Vec3{{m(1, 3), m(0, 3), m(0, 3), m(0, 3)}};
    V4 Inv0 = v4_add(v4_sub(v4_mul(Vec1, Fac0), v4_mul(Vec2, Fac1)), v4_mul(Vec3, Fac2));
    V4 Inv1 = v4_add(v4_sub(v4_mul(Vec0, Fac0), v4_mul(Vec2, Fac3)), v4_mul(Vec3, Fac4));
    V4 Inv2 = v4_add(v4_sub(v4_mul(Vec0, Fac1), v4_mul(Vec1, Fac3)), v4_mul(Vec3, Fac5));
    V4 Inv3 = v4_add(v4_sub(v4_mul(Vec0, Fac2), v4_mul(Vec1, Fac4)), v4_mul(Vec2, Fac5));
    V4 SignA{{+1.f, -1.f, +1.f, -1.f}}, SignB{{-1.f, +1.f, -1.f, +1.f}};
    M4 Inv;
    setcol(Inv, 0, v4_mul(Inv0, SignA));
    setcol(Inv, 1, v4_mul(Inv1, SignB));
    setcol(Inv, 2, v4_mul(Inv2, SignA));
    setcol(Inv, 3, v4_mul(Inv3, SignB));
    V4 Row0{{Inv.m[0][0], Inv.m[1][0], Inv.m[2][0], Inv.m[3][0]}};
    V4 Dot0 = v4_mul(col(M, 0), Row0);
    float Dot1 = (Dot0.v[0] + Dot0.v[1]) + (Dot0.v[2] + Dot0.v[3]);
    float OneOverDeterminant = 1.f / Dot1;
    M4 r;
    for (int c = 0; c < 4; c++) setcol(r, c, v4_muls(col(Inv, c), OneOverDeterminant));
    return r;
}

// ----------------------------------------------------------------------------------- scene ops

static int add_material(SceneData& d, const c2_material& m) {
    if (m.type < C2_MAT_DIFFUSE || m.type > C2_MAT_ORENNAYAR) fail(C2_ERR_INVALID, "invalid material type");
    if (m.kd_texture < 0 || m.kd_texture > (int32_t)d.textures.size())
        fail(C2_ERR_INVALID, "material names a texture id that does not exist");
    if (m.kd_texture && m.type != C2_MAT_DIFFUSE && m.type != C2_MAT_PHONG)
        fail(C2_ERR_INVALID, "only Diffuse and Phong materials take a Kd texture");
    if ((m.type == C2_MAT_MICRO || m.type == C2_MAT_PLASTIC) && m.ndf != C2_NDF_GGX && m.ndf != C2_NDF_BECKMANN)
        fail(C2_ERR_INVALID, "invalid NDF");
    d.materials.push_back(m);
    return (int)d.materials.size() - 1;
}

static int32_t add_texture(SceneData& d, int32_t w, int32_t h, const uint8_t* rgb) {
    if (w <= 0 || h <= 0 || (int64_t)w * h > (int64_t)1 << 28 || !rgb) fail(C2_ERR_INVALID, "bad texture size");
    TextureData t;
    t.width = w;
    t.height = h;
    t.rgb.assign(rgb, rgb + (size_t)w * h * 3);
    d.textures.push_back(std::move(t));
    return (int32_t)d.textures.size();
}

static void add_triangles(SceneData& d, const float* vd, int ntri, const M4& M, int material, const float* lemit) {
    if (ntri < 0 || (ntri > 0 && !vd)) fail(C2_ERR_INVALID, "add_triangles: bad arguments");
    if (material < -1 || material >= (int)d.materials.size()) fail(C2_ERR_INVALID, "add_triangles: bad material index");
    M4 NT = m4_transpose(m4_inverse(M));
    size_t base = d.tris.size();
    d.tris.resize(base + (size_t)ntri * 24);
    for (int t = 0; t < ntri; t++) {
        for (int k = 0; k < 3; k++) {
            const float* s = vd + ((size_t)t * 3 + k) * 8;
            float* o = d.tris.data() + base + ((size_t)t * 3 + k) * 8;
            V4 p = m4_mulv(M, V4{{s[0], s[1], s[2], 1.f}});
            V4 n = m4_mulv(NT, V4{{s[3], s[4], s[5], 0.f}});
            V3 nn = normalize(V3{n.v[0], n.v[1], n.v[2]});
            o[0] = p.v[0]; o[1] = p.v[1]; o[2] = p.v[2];
            o[3] = nn.x; o[4] = nn.y; o[5] = nn.z;
            o[6] = s[6]; o[7] = s[7];
        }
        int tri_index = (int)((base / 24) + t);
        int entity = (int)d.kind.size();
        d.kind.push_back(C2_SHAPE_TRIANGLE);
        d.shape.push_back(tri_index);
        d.material.push_back(material);
        if (lemit) {
            c2_light L{};
            L.type = C2_LIGHT_AREA;
            L.entity = entity;
            L.radiance[0] = lemit[0]; L.radiance[1] = lemit[1]; L.radiance[2] = lemit[2];
            d.lights.push_back(L);
            d.light.push_back((int)d.lights.size() - 1);
        } else {
            d.light.push_back(-1);
        }
    }
}

// src/Shapes/Sphere.cpp:19-30
// Sphere (Sphere.cpp:19-30) and Disk (Disk.cpp:19-29) share one record: centre, radius,
// local2world = mat3(rotation), world2local = its transpose.
static void add_sphere(SceneData& d, V3 pos, float radius, V3 rot, int material, const float* lemit,
                       int kind = C2_SHAPE_SPHERE) {
    if (material < -1 || material >= (int)d.materials.size()) fail(C2_ERR_INVALID, "add_sphere: bad material index");
    M4 R = m4_identity();
    R = m4_rotate(R, rot.x / 180.f * kPi, V3{1, 0, 0});
    R = m4_rotate(R, rot.y / 180.f * kPi, V3{0, 1, 0});
    R = m4_rotate(R, rot.z / 180.f * kPi, V3{0, 0, 1});
    float rec[22];
    rec[0] = pos.x; rec[1] = pos.y; rec[2] = pos.z; rec[3] = radius;
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) {
            rec[4 + c * 3 + r] = R.m[c][r];       // local2world = mat3(rotMatrix)
            rec[13 + c * 3 + r] = R.m[r][c];      // world2local = transpose
        }
    int sidx = (int)(d.spheres.size() / 22);
    d.spheres.insert(d.spheres.end(), rec, rec + 22);
    int entity = (int)d.kind.size();
    d.kind.push_back(kind);
    d.shape.push_back(sidx);
    d.material.push_back(material);
    if (lemit) {
        c2_light L{};
        L.type = C2_LIGHT_AREA;
        L.entity = entity;
        L.radiance[0] = lemit[0]; L.radiance[1] = lemit[1]; L.radiance[2] = lemit[2];
        d.lights.push_back(L);
        d.light.push_back((int)d.lights.size() - 1);
    } else {
        d.light.push_back(-1);
    }
}

// ------------------------------------------------------------------------------------ OBJ

// src/Loaders/ObjLoader.cpp:78-208. Returns unshared corner data [ntri][3][8].
static std::vector<float> load_obj(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) fail(C2_ERR_IO, "cannot open OBJ " + path);
    std::vector<V3> P, N;
    std::vector<V2> T;
    std::vector<float> out;
    std::vector<char> buf(100005);
    int lineno = 0;
    auto readInt = [](const char* s, int& idx, int& num) {
        num = 0;
        while (s[idx] == ' ') idx++;
        bool neg = false;
        if (s[idx] == '-') { neg = true; idx++; }
        else if (s[idx] == '+') idx++;
        while (isdigit((unsigned char)s[idx])) { num = num * 10 + (s[idx] - '0'); idx++; }
        if (neg) num = -num;
    };
    while (std::fgets(buf.data(), (int)buf.size(), f)) {
        lineno++;
        char tok[64] = {0};
        int consumed = 0;
        if (std::sscanf(buf.data(), "%63s%n", tok, &consumed) != 1) continue;
        const char* rest = buf.data() + consumed;
        if (!std::strcmp(tok, "v") || !std::strcmp(tok, "vn")) {
            double x = 0, y = 0, z = 0;
            std::sscanf(rest, "%lf%lf%lf", &x, &y, &z);
            (tok[1] ? N : P).push_back(V3{(float)x, (float)y, (float)z});
        } else if (!std::strcmp(tok, "vt")) {
            double x = 0, y = 0;
            std::sscanf(rest, "%lf%lf", &x, &y);
            T.push_back(V2{(float)x, (float)y});
        } else if (!std::strcmp(tok, "f")) {
            int vd[3] = {0, 0, 0}, vt[3] = {0, 0, 0}, vn[3] = {0, 0, 0};
            int count = 0;
            char fv[128];
            int off = 0, n = 0;
            while (std::sscanf(rest + off, "%127s%n", fv, &n) == 1) {
                if (count >= 3)
                    fail(C2_ERR_PARSE, path + ":" + std::to_string(lineno) + ": only triangular faces are supported");
                int id = 0;
                readInt(fv, id, vd[count]);
                if (fv[id] == '/') { id++; readInt(fv, id, vt[count]); }
                if (fv[id] == '/') { id++; readInt(fv, id, vn[count]); }
                count++;
                off += n;
            }
            if (count != 3) fail(C2_ERR_PARSE, path + ":" + std::to_string(lineno) + ": face with fewer than 3 vertices");
            if (vn[0] == 0)
                fail(C2_ERR_PARSE, path + ":" + std::to_string(lineno) + ": face without vertex normals (vn)");
            for (int k = 0; k < 3; k++) {
                if (vd[k] < 1 || vd[k] > (int)P.size() || vn[k] < 1 || vn[k] > (int)N.size() ||
                    (vt[0] != 0 && (vt[k] < 1 || vt[k] > (int)T.size())))
                    fail(C2_ERR_PARSE, path + ":" + std::to_string(lineno) + ": face index out of range");
                V3 p = P[vd[k] - 1], nn = N[vn[k] - 1];
                V2 uv = vt[0] != 0 ? T[vt[k] - 1] : V2{0.f, 0.f};
                float rec[8] = {p.x, p.y, p.z, nn.x, nn.y, nn.z, uv.x, uv.y};
                out.insert(out.end(), rec, rec + 8);
            }
        }
    }
    std::fclose(f);
    return out;
}

static std::string join_path(const std::string& base, const std::string& rel) {
    if (base.empty() || (!rel.empty() && rel[0] == '/')) return rel;
    return base + "/" + rel;
}

// src/Materials/Material.cpp:22-126
static int32_t parse_ndf(const std::string& ndf) {  // Material.cpp:52-64
    if (ndf == "GGX") return C2_NDF_GGX;
    if (ndf == "Beckmann") return C2_NDF_BECKMANN;
    fail(C2_ERR_PARSE, "Cannot find NDF Type: " + ndf);
}

static c2_material parse_material(const JNode& n, const std::map<std::string, int32_t>& texByName) {
    c2_material m{};
    const std::string& type = n.member("Type").as_string();
    auto color = [&]() {
        V3 c = jvec3(n.member("Color"));
        m.color[0] = c.x; m.color[1] = c.y; m.color[2] = c.z;
    };
    // "Kd" names a texture for Diffuse and Phong, else "Color" (Material.cpp:25-35,91-102)
    auto kd_or_color = [&]() {
        if (!n.has("Kd")) return color();
        const std::string& name = n.member("Kd").as_string();
        auto it = texByName.find(name);
        if (it == texByName.end()) fail(C2_ERR_PARSE, "Texture name doesn't exist: " + name);  // Scene.cpp:71-79
        m.kd_texture = it->second;
    };
    if (type == "Diffuse") {
        m.type = C2_MAT_DIFFUSE;
        kd_or_color();
    } else if (type == "Mirror") {
        m.type = C2_MAT_MIRROR;
        color();
    } else if (type == "Glass") {
        m.type = C2_MAT_GLASS;
        color();
        V2 r = jvec2(n.member("Roughness"));
        m.roughness[0] = r.x; m.roughness[1] = r.y;
        m.ior = (float)n.member("IOR").as_float();
    } else if (type == "Micro") {
        m.type = C2_MAT_MICRO;
        color();
        V2 r = jvec2(n.member("Roughness"));
        m.roughness[0] = r.x; m.roughness[1] = r.y;
        m.ior = (float)n.member("IOR").as_float();
        m.ndf = parse_ndf(n.member("NDF").as_string());
    } else if (type == "Plastic") {  // Material.cpp:70-89: Kd and Ks are colours here
        m.type = C2_MAT_PLASTIC;
        V3 kd = jvec3(n.member("Kd")), ks = jvec3(n.member("Ks"));
        m.color[0] = kd.x; m.color[1] = kd.y; m.color[2] = kd.z;
        m.ks[0] = ks.x; m.ks[1] = ks.y; m.ks[2] = ks.z;
        V2 r = jvec2(n.member("Roughness"));
        m.roughness[0] = r.x; m.roughness[1] = r.y;
        m.ndf = parse_ndf(n.member("NDF").as_string());
    } else if (type == "Phong") {
        m.type = C2_MAT_PHONG;
        m.phong_k = (int32_t)n.member("K").as_int();
        kd_or_color();
    } else if (type == "OrenNayar") {  // Material.cpp:104-120 (a Kd texture is "Not implemented")
        m.type = C2_MAT_ORENNAYAR;
        if (n.has("Kd")) fail(C2_ERR_UNSUPPORTED, "OrenNayar with a Kd texture is not implemented");
        V3 c = jvec3(n.member("Color"));
        m.color[0] = c.x; m.color[1] = c.y; m.color[2] = c.z;
        m.sigma = (float)n.member("Sigma").as_float();
    } else {
        fail(C2_ERR_PARSE, "Invalid material name");
    }
    return m;
}

static void load_scene_json(SceneData& d, const std::string& path, const std::string& base) {
    JNode root = load_json(path);
    const JNode& sc = root.member("Scene");
    const JNode& tex = sc.member("Textures");  // required (Scene.cpp:31)
    if (tex.type != JNode::Object) fail(C2_ERR_PARSE, "\"Textures\" must be an object");
    // Scene::loadTextures (Scene.cpp:84-92): every named image is loaded, used or not
    std::map<std::string, int32_t> texByName;
    for (const auto& kv : tex.obj) {
        ImageRGB im = load_image(join_path(base, kv.second.member("Path").as_string()));
        texByName[kv.first] = add_texture(d, im.width, im.height, im.rgb.data());
    }
    const JNode& objs = sc.member("Objects");
    if (objs.type != JNode::Array) fail(C2_ERR_PARSE, "\"Objects\" must be an array");
    for (const JNode& o : objs.arr) {
        const std::string& type = o.member("Type").as_string();
        if (type == "Geometry") {
            const JNode& sh = o.member("Shape");
            const std::string& st = sh.member("Type").as_string();
            int material = -1;
            if (o.has("Material")) material = add_material(d, parse_material(o.member("Material"), texByName));
            float lemit[3];
            bool hasLight = o.has("AreaLight");
            if (hasLight) {
                const JNode& al = o.member("AreaLight");
                const std::string& lt = al.member("Type").as_string();
                if (lt != "Diffused") fail(C2_ERR_PARSE, "Invalid area light name");
                V3 le = jvec3(al.member("Lemit"));
                (void)al.member("Samples").as_int();
                lemit[0] = le.x; lemit[1] = le.y; lemit[2] = le.z;
            }
            if (st == "Sphere") {
                V3 pos = jvec3(sh.member("Position"));
                float r = (float)sh.member("Radius").as_float();
                V3 rot{0, 0, 0};
                if (sh.has("Rotation")) rot = jvec3(sh.member("Rotation"));
                add_sphere(d, pos, r, rot, material, hasLight ? lemit : nullptr);
            } else if (st == "TriangleMesh") {
                std::vector<float> vd = load_obj(join_path(base, sh.member("ObjFile").as_string()));
                V3 t = jvec3(sh.member("Translation"));
                V3 s = jvec3(sh.member("Scale"));
                V3 r = jvec3(sh.member("Rotation"));
                const float k = 0.01745329251994329576923690768489f;  // glm::radians
                r = V3{r.x * k, r.y * k, r.z * k};
                M4 M = m4_identity();
                M = m4_translate(M, t);
                M = m4_rotate(M, r.x, V3{1, 0, 0});
                M = m4_rotate(M, r.y, V3{0, 1, 0});
                M = m4_rotate(M, r.z, V3{0, 0, 1});
                M = m4_scale(M, s);
                add_triangles(d, vd.data(), (int)(vd.size() / 24), M, material, hasLight ? lemit : nullptr);
            } else if (st == "Disk") {  // Disk::CreateDisk, Disk.cpp:7-17
                V3 pos = jvec3(sh.member("Position"));
                float r = (float)sh.member("Radius").as_float();
                V3 rot{0, 0, 0};
                if (sh.has("Rotation")) rot = jvec3(sh.member("Rotation"));
                add_sphere(d, pos, r, rot, material, hasLight ? lemit : nullptr, C2_SHAPE_DISK);
            } else {
                fail(C2_ERR_PARSE, "Invalid Shape Type!");
            }
        } else if (type == "Light") {
            const std::string& lt = o.member("LightType").as_string();
            if (lt != "Point") fail(C2_ERR_PARSE, "Invalid light name");
            V3 p = jvec3(o.member("Position"));
            V3 fl = jvec3(o.member("Flux"));
            c2_light L{};
            L.type = C2_LIGHT_POINT;
            L.entity = -1;
            L.radiance[0] = fl.x; L.radiance[1] = fl.y; L.radiance[2] = fl.z;
            L.position[0] = p.x; L.position[1] = p.y; L.position[2] = p.z;
            d.lights.push_back(L);
        } else {
            fail(C2_ERR_PARSE, "Invalid Object Type!");
        }
    }
    // Scene::loadSkybox (Scene.cpp:128-143): <Path>/posx.jpg ... negz.jpg
    if (sc.has("Skybox") && sc.member("Skybox").type != JNode::Null) {
        const std::string dir = sc.member("Skybox").member("Path").as_string();
        static const char* faces[6] = {"posx.jpg", "negx.jpg", "posy.jpg", "negy.jpg", "posz.jpg", "negz.jpg"};
        for (int f = 0; f < 6; f++) {
            ImageRGB im = load_image(join_path(base, dir + "/" + faces[f]));
            d.skybox[f] = add_texture(d, im.width, im.height, im.rgb.data());
        }
    }
    // Camera::CreateCamera (src/Cameras/Camera.cpp:7-20)
    const JNode& cam = root.member("Camera");
    if (cam.member("Type").as_string() != "Normal") fail(C2_ERR_PARSE, "Invalid Camera Type!");
    V3 look = jvec3(cam.member("LookAt")), up = jvec3(cam.member("Up")), pos = jvec3(cam.member("Position"));
    double fov = cam.member("FOV").as_float() / 180.f * (double)kPi;
    d.camera.fov = (float)fov;
    d.camera.position[0] = pos.x; d.camera.position[1] = pos.y; d.camera.position[2] = pos.z;
    d.camera.look_at[0] = look.x; d.camera.look_at[1] = look.y; d.camera.look_at[2] = look.z;
    d.camera.up[0] = up.x; d.camera.up[1] = up.y; d.camera.up[2] = up.z;
    d.has_camera = true;
}

// --------------------------------------------------------------------------------- Halton

static std::vector<int32_t> first_primes(int n) {
    std::vector<int32_t> p;
    for (int c = 2; (int)p.size() < n; c++) {
        bool prime = true;
        for (int q : p) {
            if (q * q > c) break;
            if (c % q == 0) { prime = false; break; }
        }
        if (prime) p.push_back(c);
    }
    return p;
}

const std::vector<int32_t>& halton_primes() {
    static const std::vector<int32_t> p = first_primes(100);
    return p;
}

struct PermTables {
    std::vector<uint16_t> data;
    std::vector<int32_t> offsets;
};
static const PermTables& perm_tables() {
    static const PermTables t = [] {
        PermTables r;
        std::mt19937 mt;  // default seed, consumed in dimension order
        for (int p : halton_primes()) {
            std::vector<int> v(p);
            for (int j = 0; j < p; j++) v[j] = j;
            std::shuffle(v.begin(), v.end(), mt);
            r.offsets.push_back((int32_t)r.data.size());
            for (int x : v) r.data.push_back((uint16_t)x);
        }
        r.offsets.push_back((int32_t)r.data.size());
        return r;
    }();
    return t;
}
const std::vector<uint16_t>& halton_perm_data() { return perm_tables().data; }
const std::vector<int32_t>& halton_perm_offsets() { return perm_tables().offsets; }

static void extgcd(uint64_t a, uint64_t b, uint64_t& d, uint64_t& x, uint64_t& y) {
    if (!b) { d = a; x = 1; y = 0; }
    else { extgcd(b, a % b, d, y, x); y -= x * (a / b); }
}
static uint64_t get_inverse(uint64_t A, uint64_t M) {
    uint64_t d, x, y;
    extgcd(A, M, d, x, y);
    return d == 1 ? (x + M) % M : (uint64_t)-1;
}

HaltonParams halton_params(int width, int height) {
    HaltonParams p{};
    int res[2] = {width, height};
    for (int i = 0; i < 2; i++) {
        int base = i == 0 ? 2 : 3;
        int scale = 1, e = 0;
        while (scale < std::min(res[i], 128)) { scale *= base; ++e; }
        p.base_scales[i] = scale;
        p.base_exponents[i] = e;
    }
    p.stride = p.base_scales[0] * p.base_scales[1];
    p.mult_inverse[0] = (int32_t)get_inverse((uint64_t)p.base_scales[1], (uint64_t)p.base_scales[0]);
    p.mult_inverse[1] = (int32_t)get_inverse((uint64_t)p.base_scales[0], (uint64_t)p.base_scales[1]);
    return p;
}

static uint64_t reverse_digits(uint64_t v, int base, int n) {
    uint64_t r = 0;
    for (int i = 0; i < n; i++) { r = r * base + (v % base); v /= base; }
    return r;
}

int32_t halton_index(const HaltonParams& p, int x, int y, int k) {
    int64_t off = 0;
    if (p.stride > 1) {
        int pm[2] = {x % 128, y % 128};
        for (int i = 0; i < 2; i++) {
            uint64_t rem = reverse_digits((uint64_t)pm[i], i == 0 ? 2 : 3, p.base_exponents[i]);
            off += (int64_t)(rem * (uint64_t)(int64_t)p.base_scales[!i] * (uint64_t)(int64_t)p.mult_inverse[i]);
        }
    }
    return (int32_t)(off + (int64_t)(k * p.stride));
}

static uint32_t reverse_bits32(uint32_t n) {
    n = (n << 16) | (n >> 16);
    n = ((n & 0x00ff00ffu) << 8) | ((n & 0xff00ff00u) >> 8);
    n = ((n & 0x0f0f0f0fu) << 4) | ((n & 0xf0f0f0f0u) >> 4);
    n = ((n & 0x33333333u) << 2) | ((n & 0xccccccccu) >> 2);
    n = ((n & 0x55555555u) << 1) | ((n & 0xaaaaaaaau) >> 1);
    return n;
}

float halton_value(const HaltonParams& p, int32_t idx, int32_t dim) {
    const auto& primes = halton_primes();
    const auto& pd = halton_perm_data();
    const auto& po = halton_perm_offsets();
    if (dim == 0) {
        uint64_t v = (uint64_t)(int64_t)(idx >> p.base_exponents[0]);
        uint64_t rev = ((uint64_t)reverse_bits32((uint32_t)v) << 32) | reverse_bits32((uint32_t)(v >> 32));
        return (float)((double)rev * 0x1p-64);
    }
    if (dim >= 100) {
        float u = (float)hash_u32((uint32_t)idx, (uint32_t)dim) * 2.3283064365386963e-10f;
        return std::min(kOneMinusEpsilon, u);
    }
    int base = primes[dim];
    float invBase = 1.f / base;
    float result = 0, invN = 1.f;
    if (dim == 1) {
        uint64_t v = (uint64_t)(int64_t)(idx / p.base_scales[1]);
        while (v) {
            uint64_t digit = v % base;
            result = result * base + digit;
            invN *= invBase;
            v /= base;
        }
        return std::min(result * invN, kOneMinusEpsilon);
    }
    const uint16_t* perm = pd.data() + po[dim];
    uint64_t v = (uint64_t)(int64_t)idx;
    while (v) {
        uint64_t digit = perm[v % base];
        result = result * base + digit;
        invN *= invBase;
        v /= base;
    }
    return std::min((result + invBase * (int)perm[0] / (1 - invBase)) * invN, kOneMinusEpsilon);
}

}  // namespace c2

using namespace c2;

// ==================================================================================== C-ABI

extern "C" {

const char* c2_last_error(void) { return g_last_error.c_str(); }

int c2_config_load(const char* path, c2_config* out) {
    return guard([&] {
        if (!path || !out) fail(C2_ERR_INVALID, "c2_config_load: null argument");
        JNode n = load_json(path);
        c2_config c{};
        c.width = (int32_t)n.member("Width").as_int();
        c.height = (int32_t)n.member("Height").as_int();
        c.spp = (int32_t)n.member("SamplesPerPixel").as_int();
        c.threads = (int32_t)n.member("Threads").as_int();
        std::snprintf(c.accel, sizeof c.accel, "%s", n.member("Acceleration").as_string().c_str());
        const JNode& in = n.member("Integrator");
        std::snprintf(c.integrator, sizeof c.integrator, "%s", in.member("Type").as_string().c_str());
        const JNode& sm = in.member("Sampler");
        std::snprintf(c.sampler, sizeof c.sampler, "%s", sm.member("Type").as_string().c_str());
        (void)sm.member("Seed").as_int();  // required by Sampler::LoadSampler even for Halton
        c.max_depth = (int32_t)in.member("Depth").as_int();
        if (c.width <= 0 || c.height <= 0 || c.spp <= 0) fail(C2_ERR_INVALID, "config: non-positive size");
        *out = c;
    });
}

int c2_scene_create(c2_scene** out) {
    return guard([&] {
        if (!out) fail(C2_ERR_INVALID, "null output");
        *out = new c2_scene();
    });
}

void c2_scene_destroy(c2_scene* s) { delete s; }

int c2_scene_load(const char* path, const char* base_dir, c2_scene** out) {
    return guard([&] {
        if (!path || !out) fail(C2_ERR_INVALID, "c2_scene_load: null argument");
        std::unique_ptr<c2_scene> s(new c2_scene());
        load_scene_json(s->d, path, base_dir ? std::string(base_dir) : std::string());
        *out = s.release();
    });
}

int c2_scene_add_material(c2_scene* s, const c2_material* m, int32_t* index) {
    return guard([&] {
        if (!s || !m) fail(C2_ERR_INVALID, "null argument");
        int i = add_material(s->d, *m);
        if (index) *index = i;
    });
}

int c2_scene_add_triangles(c2_scene* s, const float* vd, int32_t ntri, const float* transform, int32_t material,
                           const float* lemit) {
    return guard([&] {
        if (!s) fail(C2_ERR_INVALID, "null scene");
        M4 M = m4_identity();
        if (transform) std::memcpy(M.m, transform, sizeof(M.m));
        add_triangles(s->d, vd, ntri, M, material, lemit);
    });
}

int c2_scene_add_sphere(c2_scene* s, const float center[3], float radius, const float rot[3], int32_t material,
                        const float* lemit) {
    return guard([&] {
        if (!s || !center) fail(C2_ERR_INVALID, "null argument");
        V3 r{0, 0, 0};
        if (rot) r = V3{rot[0], rot[1], rot[2]};
        add_sphere(s->d, V3{center[0], center[1], center[2]}, radius, r, material, lemit);
    });
}

int c2_scene_add_disk(c2_scene* s, const float center[3], float radius, const float rot[3], int32_t material,
                      const float* lemit) {
    return guard([&] {
        if (!s || !center) fail(C2_ERR_INVALID, "null argument");
        V3 r{0, 0, 0};
        if (rot) r = V3{rot[0], rot[1], rot[2]};
        add_sphere(s->d, V3{center[0], center[1], center[2]}, radius, r, material, lemit, C2_SHAPE_DISK);
    });
}

int c2_scene_add_point_light(c2_scene* s, const float pos[3], const float flux[3]) {
    return guard([&] {
        if (!s || !pos || !flux) fail(C2_ERR_INVALID, "null argument");
        c2_light L{};
        L.type = C2_LIGHT_POINT;
        L.entity = -1;
        for (int i = 0; i < 3; i++) { L.radiance[i] = flux[i]; L.position[i] = pos[i]; }
        s->d.lights.push_back(L);
    });
}

int c2_scene_add_texture(c2_scene* s, int32_t w, int32_t h, const uint8_t* rgb, int32_t* id) {
    return guard([&] {
        if (!s || !rgb) fail(C2_ERR_INVALID, "null argument");
        int32_t k = add_texture(s->d, w, h, rgb);
        if (id) *id = k;
    });
}

int c2_scene_set_skybox(c2_scene* s, const int32_t ids[6]) {
    return guard([&] {
        if (!s || !ids) fail(C2_ERR_INVALID, "null argument");
        int set = 0;
        for (int f = 0; f < 6; f++) {
            if (ids[f] < 0 || ids[f] > (int32_t)s->d.textures.size()) fail(C2_ERR_INVALID, "skybox face id out of range");
            set += ids[f] > 0;
        }
        if (set != 0 && set != 6) fail(C2_ERR_INVALID, "a skybox needs all six faces");
        for (int f = 0; f < 6; f++) s->d.skybox[f] = ids[f];
    });
}

int c2_image_load(const char* path, int32_t* w, int32_t* h, uint8_t** rgb) {
    return guard([&] {
        if (!path || !w || !h || !rgb) fail(C2_ERR_INVALID, "null argument");
        ImageRGB im = load_image(path);
        uint8_t* buf = (uint8_t*)std::malloc(im.rgb.size());
        if (!buf) fail(C2_ERR_NOMEM, "image buffer");
        std::memcpy(buf, im.rgb.data(), im.rgb.size());
        *w = im.width;
        *h = im.height;
        *rgb = buf;
    });
}

void c2_image_free(uint8_t* rgb) { std::free(rgb); }

int c2_scene_set_camera(c2_scene* s, const c2_camera* cam) {
    return guard([&] {
        if (!s || !cam) fail(C2_ERR_INVALID, "null argument");
        s->d.camera = *cam;
        s->d.has_camera = true;
    });
}

int c2_scene_get_desc(const c2_scene* s, c2_scene_desc* out) {
    return guard([&] {
        if (!s || !out) fail(C2_ERR_INVALID, "null argument");
        const SceneData& d = s->d;
        c2_scene_desc r{};
        r.num_entities = (int32_t)d.kind.size();
        r.entity_shape_kind = d.kind.data();
        r.entity_shape_index = d.shape.data();
        r.entity_material = d.material.data();
        r.entity_light = d.light.data();
        r.num_triangles = (int32_t)(d.tris.size() / 24);
        r.triangles = d.tris.data();
        r.num_spheres = (int32_t)(d.spheres.size() / 22);
        r.spheres = d.spheres.data();
        r.num_materials = (int32_t)d.materials.size();
        r.materials = d.materials.data();
        r.num_lights = (int32_t)d.lights.size();
        r.lights = d.lights.data();
        r.camera = d.camera;
        d.texture_desc.resize(d.textures.size());
        for (size_t i = 0; i < d.textures.size(); i++)
            d.texture_desc[i] = c2_texture{d.textures[i].width, d.textures[i].height, d.textures[i].rgb.data()};
        r.num_textures = (int32_t)d.textures.size();
        r.textures = d.texture_desc.data();
        for (int f = 0; f < 6; f++) r.skybox[f] = d.skybox[f];
        *out = r;
    });
}

int c2_halton_index(int32_t w, int32_t h, int32_t x, int32_t y, int32_t k, int32_t* index) {
    return guard([&] {
        if (w <= 0 || h <= 0 || !index) fail(C2_ERR_INVALID, "bad arguments");
        *index = halton_index(halton_params(w, h), x, y, k);
    });
}

int c2_halton_value(int32_t w, int32_t h, int32_t index, int32_t dim, float* value) {
    return guard([&] {
        if (w <= 0 || h <= 0 || dim < 0 || index < 0 || !value) fail(C2_ERR_INVALID, "bad arguments");
        *value = halton_value(halton_params(w, h), index, dim);
    });
}

}  // extern "C"
