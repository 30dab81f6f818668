// gltf.cpp -- binary glTF 2.0 (.glb) mesh loader for the host Scene, in place
// of the reference's assimp import (Scene::load, CRT/Scene.cpp:20-221).
//
// What it reads, as the reference's import uses it:
//   * the default scene's node tree; a mesh's vertices are placed by the
//     node's accumulated transform, normals by its inverse transpose and
//     renormalised (CRT/Scene.cpp:119-139);
//   * triangle primitives: POSITION, NORMAL (smooth area-weighted normals are
//     generated when absent, as aiProcess_GenSmoothNormals fills them in),
//     TEXCOORD_0 with v flipped (aiProcess_FlipUVs), and indices;
//   * materials by name (or index), base colour from baseColorFactor;
//     roughness 1 and metallic 0 as the reference's loader forces them
//     (CRT/Scene.cpp:203-204);
//   * embedded PNG images (png.cpp) of baseColorTexture (RGB) and of
//     metallicRoughnessTexture (its green channel, the roughness, and its
//     blue channel, the metalness, which the shading reads in place of the
//     factors as get_roughness / get_metallic do, CRT/dMaterial.cu:25-59);
//     samples are scaled by 1/255 and rows stored bottom first, matching the
//     flipped v; normalTexture (RGB, read by get_normal, CRT/dMaterial.cu:61-87)
//     with per-vertex tangent frames from TANGENT, or computed from the
//     texture coordinates when the file has none.
// Node transforms compose parent * local, the glTF rule (the reference's
// accumulation multiplies in the other order, CRT/Scene.cpp:88; the two
// agree for one level of nesting).  Not read: JPEG or external images,
// occlusion / emissive maps (the reference's shading reads neither; its
// emission is zero, CRT/dScene.cpp:196), cameras and lights (the reference ignores both,
// CRT/Scene.cpp:230-258), sparse accessors, skins, morph targets,
// non-triangle primitives; a file using sparse accessors or other primitive
// modes is rejected.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <utility>
#include <vector>

#include "firefly.h"

namespace firefly {
namespace {

// ------------------------------------------------------------- JSON ----
struct JVal {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;

    const JVal* get(const char* key) const {
        if (kind != OBJ) return nullptr;
        for (const auto& kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    double num_or(const char* key, double d) const {
        const JVal* v = get(key);
        return v && v->kind == NUM ? v->num : d;
    }
    // This value as a non-negative integer no larger than `limit` (< 2^53):
    // fractions, negatives, non-numbers and non-finite values are rejected
    // before any conversion, so no cast below is undefined.
    size_t as_uint(size_t limit, const char* what) const {
        if (kind != NUM || !(num >= 0.0) || num > (double)limit || num != std::floor(num))
            throw Error(FF_ERR_INVALID, std::string("glTF: ") + what + " is not an integer in [0, " +
                                            std::to_string(limit) + "]");
        return (size_t)num;
    }
    // field `key` as a non-negative integer <= limit, `d` when absent
    size_t uint_or(const char* key, size_t d, size_t limit) const {
        const JVal* v = get(key);
        return v ? v->as_uint(limit, key) : d;
    }
    // field `key` as a 32-bit signed integer, `d` when absent
    int int_or(const char* key, int d) const {
        const JVal* v = get(key);
        if (!v) return d;
        if (v->kind != NUM || !(v->num >= -2147483648.0 && v->num <= 2147483647.0) || v->num != std::floor(v->num))
            throw Error(FF_ERR_INVALID, std::string("glTF: ") + key + " is not a 32-bit integer");
        return (int)v->num;
    }
    // this value as an index into a top-level array (checked again by at())
    int index(const char* what) const { return (int)as_uint(2147483647u, what); }
    const JVal& at(size_t i) const {
        if (kind != ARR || i >= arr.size()) throw Error(FF_ERR_INVALID, "glTF: index out of range");
        return arr[i];
    }
};

class JParser {
public:
    JParser(const char* b, const char* e) : p(b), end(e) {}
    JVal parse() {
        JVal v = value();
        ws();
        if (p != end) fail("trailing characters");
        return v;
    }

private:
    const char* p;
    const char* end;
    int depth = 0;  // nesting bound: a hostile file cannot exhaust the stack
    [[noreturn]] void fail(const char* what) { throw Error(FF_ERR_INVALID, std::string("glTF JSON: ") + what); }
    void ws() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) p++;
    }
    bool lit(const char* s) {
        size_t n = std::strlen(s);
        if ((size_t)(end - p) >= n && std::memcmp(p, s, n) == 0) {
            p += n;
            return true;
        }
        return false;
    }
    static void utf8(std::string& o, uint32_t c) {
        if (c < 0x80) {
            o += (char)c;
        } else if (c < 0x800) {
            o += (char)(0xC0 | (c >> 6));
            o += (char)(0x80 | (c & 0x3F));
        } else if (c < 0x10000) {
            o += (char)(0xE0 | (c >> 12));
            o += (char)(0x80 | ((c >> 6) & 0x3F));
            o += (char)(0x80 | (c & 0x3F));
        } else {
            o += (char)(0xF0 | (c >> 18));
            o += (char)(0x80 | ((c >> 12) & 0x3F));
            o += (char)(0x80 | ((c >> 6) & 0x3F));
            o += (char)(0x80 | (c & 0x3F));
        }
    }
    uint32_t hex4() {
        if (end - p < 4) fail("short \\u escape");
        uint32_t c = 0;
        for (int k = 0; k < 4; k++) {
            char h = *p++;
            c <<= 4;
            if (h >= '0' && h <= '9') c |= (uint32_t)(h - '0');
            else if (h >= 'a' && h <= 'f') c |= (uint32_t)(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') c |= (uint32_t)(h - 'A' + 10);
            else fail("bad \\u escape");
        }
        return c;
    }
    std::string string() {
        if (p >= end || *p != '"') fail("expected a string");
        p++;
        std::string o;
        while (true) {
            if (p >= end) fail("unterminated string");
            char c = *p++;
            if (c == '"') break;
            if (c != '\\') {
                o += c;
                continue;
            }
            if (p >= end) fail("unterminated escape");
            char e = *p++;
            switch (e) {
                case '"': o += '"'; break;
                case '\\': o += '\\'; break;
                case '/': o += '/'; break;
                case 'b': o += '\b'; break;
                case 'f': o += '\f'; break;
                case 'n': o += '\n'; break;
                case 'r': o += '\r'; break;
                case 't': o += '\t'; break;
                case 'u': {
                    uint32_t c1 = hex4();
                    if (c1 >= 0xD800 && c1 < 0xDC00 && end - p >= 6 && p[0] == '\\' && p[1] == 'u') {
                        p += 2;
                        uint32_t c2 = hex4();
                        c1 = 0x10000 + ((c1 - 0xD800) << 10) + (c2 - 0xDC00);
                    }
                    utf8(o, c1);
                    break;
                }
                default: fail("bad escape");
            }
        }
        return o;
    }
    struct Nest {
        int& d;
        explicit Nest(int& x) : d(x) { ++d; }
        ~Nest() { --d; }
    };
    JVal value() {
        ws();
        if (p >= end) fail("unexpected end");
        Nest guard(depth);
        if (depth > 256) fail("nesting deeper than 256");
        JVal v;
        char c = *p;
        if (c == '{') {
            p++;
            v.kind = JVal::OBJ;
            ws();
            if (p < end && *p == '}') {
                p++;
                return v;
            }
            while (true) {
                ws();
                std::string k = string();
                ws();
                if (p >= end || *p != ':') fail("expected ':'");
                p++;
                v.obj.emplace_back(std::move(k), value());
                ws();
                if (p < end && *p == ',') {
                    p++;
                    continue;
                }
                if (p < end && *p == '}') {
                    p++;
                    return v;
                }
                fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            p++;
            v.kind = JVal::ARR;
            ws();
            if (p < end && *p == ']') {
                p++;
                return v;
            }
            while (true) {
                v.arr.push_back(value());
                ws();
                if (p < end && *p == ',') {
                    p++;
                    continue;
                }
                if (p < end && *p == ']') {
                    p++;
                    return v;
                }
                fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = JVal::STR;
            v.str = string();
            return v;
        }
        if (lit("true")) {
            v.kind = JVal::BOOL;
            v.b = true;
            return v;
        }
        if (lit("false")) {
            v.kind = JVal::BOOL;
            return v;
        }
        if (lit("null")) return v;
        // number: copy the token so strtod never reads past the chunk
        const char* s = p;
        while (p < end && (std::strchr("+-0123456789.eE", *p) != nullptr)) p++;
        if (p == s) fail("unexpected character");
        std::string tok(s, p);
        char* stop = nullptr;
        v.kind = JVal::NUM;
        v.num = std::strtod(tok.c_str(), &stop);
        if (stop != tok.c_str() + tok.size()) fail("bad number");
        return v;
    }
};

// ------------------------------------------------------------ glTF -----
struct Doc {
    JVal j;
    std::vector<uint8_t> bin;
    // a top-level array the file must have once something refers to it
    const JVal& list(const char* key) const {
        const JVal* v = j.get(key);
        if (!v || v->kind != JVal::ARR) throw Error(FF_ERR_INVALID, std::string("glTF: missing array ") + key);
        return *v;
    }
};

uint32_t rd32(const uint8_t* b) { return (uint32_t)b[0] | (uint32_t)b[1] << 8 | (uint32_t)b[2] << 16 | (uint32_t)b[3] << 24; }

int comps_of(const std::string& t) {
    if (t == "SCALAR") return 1;
    if (t == "VEC2") return 2;
    if (t == "VEC3") return 3;
    if (t == "VEC4") return 4;
    if (t == "MAT4") return 16;
    throw Error(FF_ERR_INVALID, "glTF: accessor type " + t);
}
int bytes_of(int ct) {
    switch (ct) {
        case 5120: case 5121: return 1;
        case 5122: case 5123: return 2;
        case 5125: case 5126: return 4;
        default: throw Error(FF_ERR_INVALID, "glTF: component type " + std::to_string(ct));
    }
}

// The accessor's elements as doubles (float) or integers (index types),
// each bounds-checked against its buffer view and the binary chunk.
struct Acc {
    int comps = 0, ct = 0;
    size_t count = 0;
    std::vector<double> v;  // count * comps
};
Acc read_accessor(const Doc& d, int idx) {
    const JVal& a = d.list("accessors").at((size_t)idx);
    if (a.get("sparse")) throw Error(FF_ERR_UNSUPPORTED, "glTF: sparse accessors are not supported");
    Acc r;
    const JVal* type = a.get("type");
    if (!type || type->kind != JVal::STR) throw Error(FF_ERR_INVALID, "glTF: accessor without type");
    r.comps = comps_of(type->str);
    r.ct = a.int_or("componentType", 0);
    int es = bytes_of(r.ct);
    bool norm = a.get("normalized") && a.get("normalized")->b;
    const JVal* bvi = a.get("bufferView");
    const std::string outside = "glTF: accessor " + std::to_string(idx) + " outside its buffer";
    if (!bvi) {  // all zeros; bounded so that a count alone cannot exhaust memory
        r.count = a.uint_or("count", 0, (size_t)1 << 24);
        r.v.assign(r.count * (size_t)r.comps, 0.0);
        return r;
    }
    const size_t nb = d.bin.size();
    const JVal& bv = d.list("bufferViews").at((size_t)bvi->index("bufferView"));
    if (bv.int_or("buffer", 0) != 0) throw Error(FF_ERR_UNSUPPORTED, "glTF: only the GLB binary buffer is read");
    // every offset and length is at most the binary chunk's size, and each
    // check is written as a subtraction of quantities already known to be in
    // range, so nothing can wrap
    size_t vo = bv.uint_or("byteOffset", 0, nb), vl = bv.uint_or("byteLength", 0, nb);
    if (vl > nb - vo) throw Error(FF_ERR_INVALID, outside);
    size_t stride = bv.uint_or("byteStride", 0, 255);
    size_t elem = (size_t)es * (size_t)r.comps;
    if (stride == 0) stride = elem;
    size_t ao = a.uint_or("byteOffset", 0, nb);
    r.count = a.uint_or("count", 0, nb);
    if (r.count && (ao > vl || elem > vl - ao || r.count - 1 > (vl - ao - elem) / stride))
        throw Error(FF_ERR_INVALID, outside);
    r.v.assign(r.count * (size_t)r.comps, 0.0);  // count <= the chunk's size, comps <= 16
    const uint8_t* base = d.bin.data() + vo + ao;
    for (size_t i = 0; i < r.count; i++)
        for (int c = 0; c < r.comps; c++) {
            const uint8_t* e = base + i * stride + (size_t)c * (size_t)es;
            double x;
            switch (r.ct) {
                case 5126: { float f; std::memcpy(&f, e, 4); x = f; break; }
                case 5125: x = (double)rd32(e); break;
                case 5123: x = (double)((uint32_t)e[0] | (uint32_t)e[1] << 8); break;
                case 5122: { int16_t s; std::memcpy(&s, e, 2); x = norm ? std::fmax(s / 32767.0, -1.0) : s; break; }
                case 5121: x = (double)e[0]; break;
                default: { int8_t s; std::memcpy(&s, e, 1); x = norm ? std::fmax(s / 127.0, -1.0) : s; break; }
            }
            if (norm && r.ct == 5123) x /= 65535.0;
            if (norm && r.ct == 5121) x /= 255.0;
            r.v[i * (size_t)r.comps + (size_t)c] = x;
        }
    return r;
}

// column-major 4x4
struct M4 {
    double m[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
};
M4 mul(const M4& a, const M4& b) {
    M4 r;
    for (int c = 0; c < 4; c++)
        for (int row = 0; row < 4; row++) {
            double s = 0.0;
            for (int k = 0; k < 4; k++) s += a.m[k * 4 + row] * b.m[c * 4 + k];
            r.m[c * 4 + row] = s;
        }
    return r;
}
M4 local_of(const JVal& node) {
    M4 r;
    if (const JVal* m = node.get("matrix")) {
        for (int k = 0; k < 16; k++) r.m[k] = m->at((size_t)k).num;
        return r;
    }
    double t[3] = {0, 0, 0}, q[4] = {0, 0, 0, 1}, s[3] = {1, 1, 1};
    if (const JVal* v = node.get("translation"))
        for (int k = 0; k < 3; k++) t[k] = v->at((size_t)k).num;
    if (const JVal* v = node.get("rotation"))
        for (int k = 0; k < 4; k++) q[k] = v->at((size_t)k).num;
    if (const JVal* v = node.get("scale"))
        for (int k = 0; k < 3; k++) s[k] = v->at((size_t)k).num;
    double x = q[0], y = q[1], z = q[2], w = q[3];
    double R[9] = {1 - 2 * (y * y + z * z), 2 * (x * y + z * w), 2 * (x * z - y * w),
                   2 * (x * y - z * w), 1 - 2 * (x * x + z * z), 2 * (y * z + x * w),
                   2 * (x * z + y * w), 2 * (y * z - x * w), 1 - 2 * (x * x + y * y)};  // columns
    for (int c = 0; c < 3; c++)
        for (int row = 0; row < 3; row++) r.m[c * 4 + row] = R[c * 3 + row] * s[c];
    for (int k = 0; k < 3; k++) r.m[12 + k] = t[k];
    return r;
}

struct Loader {
    const Doc& d;
    Scene& scene;
    Model model;
    int tris = 0;
    std::vector<std::pair<long, int>> tex_done;  // (image * 2 + roughness?, Scene texture index)

    // glTF texture `ti` as a Scene texture: RGB base colour, or roughness
    // from the green channel; -1 when the material names none
    // channel < 0: the RGB texture of a base colour; 1 or 2: a one-channel
    // texture of the green (roughness) or blue (metalness) samples
    int texture(const JVal* info, int channel) {
        if (!info) return -1;
        const JVal* ti = info->get("index");
        if (!ti) return -1;
        const JVal& tj = d.list("textures").at((size_t)ti->index("texture index"));
        const JVal* src = tj.get("source");
        if (!src) throw Error(FF_ERR_UNSUPPORTED, "glTF: texture without a source image");
        const int src_i = src->index("texture source");
        long key = (long)src_i * 3 + (channel < 0 ? 0 : channel);
        for (const auto& kv : tex_done)
            if (kv.first == key) return kv.second;
        const JVal& im = d.list("images").at((size_t)src_i);
        const JVal* bvi = im.get("bufferView");
        const JVal* mime = im.get("mimeType");
        if (!bvi || (mime && mime->str != "image/png"))
            throw Error(FF_ERR_UNSUPPORTED, "glTF: only PNG images embedded in the binary chunk are read");
        const JVal& bv = d.list("bufferViews").at((size_t)bvi->index("bufferView"));
        const size_t nb = d.bin.size();
        size_t off = bv.uint_or("byteOffset", 0, nb), len = bv.uint_or("byteLength", 0, nb);
        if (len > nb - off) throw Error(FF_ERR_INVALID, "glTF: image outside the binary chunk");
        int w = 0, h = 0, c = 0;
        std::vector<uint8_t> px;
        decode_png(d.bin.data() + off, len, w, h, c, px);
        Texture t;
        t.filepath = "image" + std::to_string(src_i);
        t.width = w;
        t.height = h;
        t.nrComponents = channel < 0 ? 3 : 1;
        t.data.resize((size_t)w * h * (size_t)t.nrComponents);
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                const uint8_t* s = &px[((size_t)(h - 1 - y) * w + x) * (size_t)c];  // bottom row first
                float* o = &t.data[((size_t)y * w + x) * (size_t)t.nrComponents];
                if (channel >= 0) {
                    o[0] = (c >= 3 ? s[channel] : s[0]) / 255.f;  // gray images: the one sample
                } else {
                    for (int k = 0; k < 3; k++) o[k] = (c >= 3 ? s[k] : s[0]) / 255.f;
                }
            }
        scene.textures_loaded.push_back(std::move(t));
        int idx = (int)scene.textures_loaded.size() - 1;
        tex_done.emplace_back(key, idx);
        return idx;
    }

    std::string material_name(int mi) {
        if (mi < 0) {
            if (!scene.materials_loaded.count("default")) {
                Material m;
                m.name = "default";
                m.baseColorFactor = vec3(1.f, 1.f, 1.f);
                scene.add_material(m);
            }
            return "default";
        }
        const JVal& mj = d.list("materials").at((size_t)mi);
        const JVal* nm = mj.get("name");
        std::string name = nm && nm->kind == JVal::STR && !nm->str.empty() ? nm->str : std::to_string(mi);
        if (!scene.materials_loaded.count(name)) {
            Material m;
            m.name = name;
            m.baseColorFactor = vec3(1.f, 1.f, 1.f);
            if (const JVal* pbr = mj.get("pbrMetallicRoughness")) {
                if (const JVal* bc = pbr->get("baseColorFactor"))
                    m.baseColorFactor = vec3((float)bc->at(0).num, (float)bc->at(1).num, (float)bc->at(2).num);
                m.baseColorTexture = texture(pbr->get("baseColorTexture"), -1);
                m.roughnessTexture = texture(pbr->get("metallicRoughnessTexture"), 1);
                m.metallicTexture = texture(pbr->get("metallicRoughnessTexture"), 2);
            }
            m.normalTexture = texture(mj.get("normalTexture"), -1);
            if (const JVal* e = mj.get("emissiveFactor"))
                m.emissiveColorFactor = vec3((float)e->at(0).num, (float)e->at(1).num, (float)e->at(2).num);
            m.roughnessFactor = 1.f;  // forced, as CRT/Scene.cpp:203-204
            m.metallicFactor = 0.f;
            scene.add_material(m);
        }
        return name;
    }

    void mesh(int mi, const M4& world) {
        const JVal& mj = d.list("meshes").at((size_t)mi);
        const JVal* prims = mj.get("primitives");
        if (!prims) return;
        // inverse transpose of the upper 3x3 (cofactors / determinant)
        const double* m = world.m;
        double a = m[0], b = m[4], c = m[8], dd = m[1], e = m[5], f = m[9], g = m[2], h = m[6], i = m[10];
        double det = a * (e * i - f * h) - b * (dd * i - f * g) + c * (dd * h - e * g);
        double N[9] = {(e * i - f * h) / det, -(dd * i - f * g) / det, (dd * h - e * g) / det,
                       -(b * i - c * h) / det, (a * i - c * g) / det, -(a * h - b * g) / det,
                       (b * f - c * e) / det, -(a * f - c * dd) / det, (a * e - b * dd) / det};  // rows of M^-T (cofactors / det)
        for (size_t pi = 0; pi < prims->arr.size(); pi++) {
            const JVal& p = prims->arr[pi];
            if (p.int_or("mode", 4) != 4) throw Error(FF_ERR_UNSUPPORTED, "glTF: only triangle primitives are read");
            const JVal* at = p.get("attributes");
            const JVal* pos_i = at ? at->get("POSITION") : nullptr;
            if (!pos_i) throw Error(FF_ERR_INVALID, "glTF: primitive without POSITION");
            Acc P = read_accessor(d, pos_i->index("POSITION"));
            if (P.comps != 3 || P.ct != 5126) throw Error(FF_ERR_INVALID, "glTF: POSITION must be float VEC3");
            Mesh me;
            const JVal* nm = mj.get("name");
            me.name = (nm && nm->kind == JVal::STR ? nm->str : "mesh" + std::to_string(mi)) + "/" + std::to_string(pi);
            me.material = material_name(p.int_or("material", -1));
            size_t nv = P.count;
            me.positions.resize(nv);
            for (size_t v = 0; v < nv; v++) {
                double x = P.v[3 * v], y = P.v[3 * v + 1], z = P.v[3 * v + 2];
                me.positions[v] = vec3((float)(((m[0] * x + m[4] * y) + m[8] * z) + m[12]),
                                       (float)(((m[1] * x + m[5] * y) + m[9] * z) + m[13]),
                                       (float)(((m[2] * x + m[6] * y) + m[10] * z) + m[14]));
            }
            if (const JVal* ii = p.get("indices")) {
                Acc I = read_accessor(d, ii->index("indices"));
                if (I.comps != 1 || (I.ct != 5121 && I.ct != 5123 && I.ct != 5125))
                    throw Error(FF_ERR_INVALID, "glTF: indices must be unsigned scalars");
                me.indices.resize(I.count);
                for (size_t k = 0; k < I.count; k++) {
                    if (I.v[k] >= (double)nv) throw Error(FF_ERR_INVALID, "glTF: index out of range in " + me.name);
                    me.indices[k] = (uint32_t)I.v[k];
                }
            } else {
                me.indices.resize(nv);
                for (size_t k = 0; k < nv; k++) me.indices[k] = (uint32_t)k;
            }
            me.indices.resize(me.indices.size() / 3 * 3);
            if (const JVal* ni = at->get("NORMAL")) {
                Acc Nn = read_accessor(d, ni->index("NORMAL"));
                if (Nn.comps != 3 || Nn.count != nv) throw Error(FF_ERR_INVALID, "glTF: NORMAL must be VEC3 per vertex");
                me.normals.resize(nv);
                for (size_t v = 0; v < nv; v++) {
                    double x = Nn.v[3 * v], y = Nn.v[3 * v + 1], z = Nn.v[3 * v + 2];
                    double X = (N[0] * x + N[1] * y) + N[2] * z, Y = (N[3] * x + N[4] * y) + N[5] * z,
                           Z = (N[6] * x + N[7] * y) + N[8] * z;
                    double l = std::sqrt((X * X + Y * Y) + Z * Z);
                    me.normals[v] = l > 0 ? vec3((float)(X / l), (float)(Y / l), (float)(Z / l)) : vec3(0.f, 1.f, 0.f);
                }
            } else {  // smooth normals: area-weighted face normals summed per vertex, in world space
                std::vector<double> acc(3 * nv, 0.0);
                for (size_t k = 0; k + 2 < me.indices.size(); k += 3) {
                    const vec3& p0 = me.positions[me.indices[k]];
                    const vec3& p1 = me.positions[me.indices[k + 1]];
                    const vec3& p2 = me.positions[me.indices[k + 2]];
                    double ux = (double)p1.x - p0.x, uy = (double)p1.y - p0.y, uz = (double)p1.z - p0.z;
                    double vx = (double)p2.x - p0.x, vy = (double)p2.y - p0.y, vz = (double)p2.z - p0.z;
                    double fx = uy * vz - uz * vy, fy = uz * vx - ux * vz, fz = ux * vy - uy * vx;
                    for (int q = 0; q < 3; q++) {
                        size_t vi = me.indices[k + (size_t)q];
                        acc[3 * vi] += fx;
                        acc[3 * vi + 1] += fy;
                        acc[3 * vi + 2] += fz;
                    }
                }
                me.normals.resize(nv);
                for (size_t v = 0; v < nv; v++) {
                    double X = acc[3 * v], Y = acc[3 * v + 1], Z = acc[3 * v + 2];
                    double l = std::sqrt((X * X + Y * Y) + Z * Z);
                    me.normals[v] = l > 0 ? vec3((float)(X / l), (float)(Y / l), (float)(Z / l)) : vec3(0.f, 1.f, 0.f);
                }
            }
            if (const JVal* ti = at->get("TEXCOORD_0")) {
                Acc T = read_accessor(d, ti->index("TEXCOORD_0"));
                if (T.comps != 2 || T.count != nv) throw Error(FF_ERR_INVALID, "glTF: TEXCOORD_0 must be VEC2 per vertex");
                me.texcoords.resize(nv);
                for (size_t v = 0; v < nv; v++) {
                    float u = (float)T.v[2 * v], w = (float)T.v[2 * v + 1];
                    me.texcoords[v] = vec2{u, 1.f - w};  // aiProcess_FlipUVs
                }
            }
            const JVal* tgi = at->get("TANGENT");
            const JVal* nmi = at->get("NORMAL");
            if (tgi && nmi) {
                // TANGENT xyz and bitangent cross(NORMAL, xyz) * w (the glTF rule), both
                // in mesh space, then placed as the reference places them:
                // normalize(mat * vec4(v, 1)), translation included (CRT/Scene.cpp:143-153)
                Acc T = read_accessor(d, tgi->index("TANGENT"));
                Acc Nn = read_accessor(d, nmi->index("NORMAL"));
                if (T.comps != 4 || T.count != nv) throw Error(FF_ERR_INVALID, "glTF: TANGENT must be VEC4 per vertex");
                auto place = [&](double x, double y, double z) {
                    double X = ((m[0] * x + m[4] * y) + m[8] * z) + m[12], Y = ((m[1] * x + m[5] * y) + m[9] * z) + m[13],
                           Z = ((m[2] * x + m[6] * y) + m[10] * z) + m[14];
                    double l = std::sqrt((X * X + Y * Y) + Z * Z);
                    return l > 0 ? vec3((float)(X / l), (float)(Y / l), (float)(Z / l)) : vec3(1.f, 0.f, 0.f);
                };
                me.tangents.resize(nv);
                me.bitangents.resize(nv);
                for (size_t v = 0; v < nv; v++) {
                    double tx = T.v[4 * v], ty = T.v[4 * v + 1], tz = T.v[4 * v + 2], tw = T.v[4 * v + 3];
                    double nx = Nn.v[3 * v], ny = Nn.v[3 * v + 1], nz = Nn.v[3 * v + 2];
                    me.tangents[v] = place(tx, ty, tz);
                    me.bitangents[v] = place((ny * tz - nz * ty) * tw, (nz * tx - nx * tz) * tw, (nx * ty - ny * tx) * tw);
                }
            } else if (!me.texcoords.empty() && scene.materials_loaded[me.material].normalTexture >= 0) {
                // from the texture coordinates (the per-face vectors of assimp's
                // aiProcess_CalcTangentSpace): tangent w*sy - v*ty, bitangent
                // w*sx - v*tx, signed by the uv winding, made orthogonal to each
                // vertex normal, normalised, summed per vertex, normalised.  The
                // import runs that step before aiProcess_FlipUVs; with v already
                // flipped here the same bitangent is -(w*sx - v*tx).
                std::vector<double> at3(3 * nv, 0.0), ab3(3 * nv, 0.0);
                for (size_t k = 0; k + 2 < me.indices.size(); k += 3) {
                    uint32_t i0 = me.indices[k], i1 = me.indices[k + 1], i2 = me.indices[k + 2];
                    const vec3 &p0 = me.positions[i0], &p1 = me.positions[i1], &p2 = me.positions[i2];
                    double vx = (double)p1.x - p0.x, vy = (double)p1.y - p0.y, vz = (double)p1.z - p0.z;
                    double wx = (double)p2.x - p0.x, wy = (double)p2.y - p0.y, wz = (double)p2.z - p0.z;
                    const vec2 &t0 = me.texcoords[i0], &t1 = me.texcoords[i1], &t2 = me.texcoords[i2];
                    double sx = (double)t1.x - t0.x, sy = (double)t1.y - t0.y;
                    double tx = (double)t2.x - t0.x, ty = (double)t2.y - t0.y;
                    double dir = (tx * sy - ty * sx) < 0.0 ? -1.0 : 1.0;
                    if (sx * ty == sy * tx) { sx = 0.0; sy = 1.0; tx = 1.0; ty = 0.0; }
                    double T3[3] = {(wx * sy - vx * ty) * dir, (wy * sy - vy * ty) * dir, (wz * sy - vz * ty) * dir};
                    double B3[3] = {-(wx * sx - vx * tx) * dir, -(wy * sx - vy * tx) * dir, -(wz * sx - vz * tx) * dir};
                    for (uint32_t vi : {i0, i1, i2}) {
                        const vec3& n = me.normals[vi];
                        for (int which = 0; which < 2; which++) {
                            const double* src = which ? B3 : T3;
                            double dn = (src[0] * n.x + src[1] * n.y) + src[2] * n.z;
                            double x = src[0] - n.x * dn, y = src[1] - n.y * dn, z = src[2] - n.z * dn;
                            double l = std::sqrt((x * x + y * y) + z * z);
                            if (!(l > 0)) continue;
                            double* acc = which ? &ab3[3 * vi] : &at3[3 * vi];
                            acc[0] += x / l;
                            acc[1] += y / l;
                            acc[2] += z / l;
                        }
                    }
                }
                auto unit = [](const double* a, vec3 fallback) {
                    double l = std::sqrt((a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]);
                    return l > 0 ? vec3((float)(a[0] / l), (float)(a[1] / l), (float)(a[2] / l)) : fallback;
                };
                me.tangents.resize(nv);
                me.bitangents.resize(nv);
                for (size_t v = 0; v < nv; v++) {
                    me.tangents[v] = unit(&at3[3 * v], vec3(1.f, 0.f, 0.f));
                    me.bitangents[v] = unit(&ab3[3 * v], vec3(0.f, 0.f, 1.f));
                }
            }
            tris += me.get_nmb_of_triangles();
            model.meshes.push_back(std::move(me));
        }
    }

    void node(int ni, const M4& parent, int depth) {
        if (depth > 256) throw Error(FF_ERR_INVALID, "glTF: node hierarchy too deep (a cycle?)");
        const JVal& nj = d.list("nodes").at((size_t)ni);
        M4 world = mul(parent, local_of(nj));
        if (const JVal* mi = nj.get("mesh")) mesh(mi->index("mesh"), world);
        if (const JVal* ch = nj.get("children"))
            for (const JVal& c : ch->arr) node(c.index("child node"), world, depth + 1);
    }
};

}  // namespace

int Scene::load(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error(FF_ERR_INVALID, "Scene::load: cannot open " + path);
    std::vector<uint8_t> file((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    if (file.size() < 20 || rd32(file.data()) != 0x46546C67u)  // "glTF"
        throw Error(FF_ERR_INVALID, "Scene::load: not a binary glTF file: " + path);
    if (rd32(file.data() + 4) != 2) throw Error(FF_ERR_UNSUPPORTED, "Scene::load: glTF version is not 2");
    size_t total = std::min<size_t>(rd32(file.data() + 8), file.size());
    Doc d;
    bool have_json = false;
    for (size_t off = 12; off + 8 <= total;) {
        size_t len = rd32(file.data() + off);
        uint32_t type = rd32(file.data() + off + 4);
        if (off + 8 + len > total) throw Error(FF_ERR_INVALID, "Scene::load: chunk outside the file");
        const uint8_t* data = file.data() + off + 8;
        if (type == 0x4E4F534Au && !have_json) {  // "JSON"
            d.j = JParser((const char*)data, (const char*)data + len).parse();
            have_json = true;
        } else if (type == 0x004E4942u && d.bin.empty()) {  // "BIN\0"
            d.bin.assign(data, data + len);
        }
        off += 8 + ((len + 3) & ~(size_t)3);
    }
    if (!have_json) throw Error(FF_ERR_INVALID, "Scene::load: no JSON chunk");
    Loader L{d, *this, {}, 0};
    const JVal* scenes = d.j.get("scenes");
    if (scenes && scenes->kind == JVal::ARR && !scenes->arr.empty()) {
        const JVal& sc = scenes->at((size_t)d.j.int_or("scene", 0));
        if (const JVal* roots = sc.get("nodes"))
            for (const JVal& r : roots->arr) L.node(r.index("root node"), M4{}, 0);
    } else if (const JVal* meshes = d.j.get("meshes")) {  // no scene: every mesh untransformed
        for (size_t k = 0; k < meshes->arr.size(); k++) L.mesh((int)k, M4{});
    }
    if (!L.model.meshes.empty()) add_model(std::move(L.model));
    return L.tris;
}

}  // namespace firefly
