// firefly.cpp -- C++ host interface over the C ABI (see firefly.h).
#include "firefly.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>

namespace firefly {

void check(ff_status st, const char* what) {
    if (st != FF_OK) throw Error(st, std::string(what) + ": " + ff_last_error());
}

namespace {

constexpr double kPi = 3.14159265358979323846;

vec3 vnorm(vec3 v) {
    double l = std::sqrt((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z);
    return vec3((float)(v.x / l), (float)(v.y / l), (float)(v.z / l));
}
vec3 vcross(vec3 a, vec3 b) { return vec3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

// column-major 4x4 helpers in double: m[c*4 + r]
void mat_mul(const double* a, const double* b, double* out) {
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) {
            double s = 0;
            for (int k = 0; k < 4; k++) s += a[k * 4 + r] * b[c * 4 + k];
            out[c * 4 + r] = s;
        }
}
bool mat_inv(const double* m, double* out) {
    double a[4][8];
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) {
            a[r][c] = m[c * 4 + r];
            a[r][c + 4] = r == c ? 1.0 : 0.0;
        }
    for (int col = 0; col < 4; col++) {
        int piv = col;
        for (int r = col + 1; r < 4; r++)
            if (std::fabs(a[r][col]) > std::fabs(a[piv][col])) piv = r;
        if (a[piv][col] == 0.0) return false;
        if (piv != col)
            for (int k = 0; k < 8; k++) std::swap(a[piv][k], a[col][k]);
        double inv = 1.0 / a[col][col];
        for (int k = 0; k < 8; k++) a[col][k] *= inv;
        for (int r = 0; r < 4; r++)
            if (r != col) {
                double f = a[r][col];
                for (int k = 0; k < 8; k++) a[r][k] -= f * a[col][k];
            }
    }
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) out[c * 4 + r] = a[r][c + 4];
    return true;
}

}  // namespace

// ---------------------------------------------------------------- Camera --
void Camera::updateCameraVectors() {
    double y = yaw * kPi / 180.0, p = pitch * kPi / 180.0;
    front = vnorm(vec3((float)(std::cos(y) * std::cos(p)), (float)std::sin(p), (float)(std::sin(y) * std::cos(p))));
    right = vnorm(vcross(front, worldUp));
    up = vnorm(vcross(right, front));
    // right-handed lookAt(position, position + front, up)
    vec3 f = front, s = right, u = up;
    float m[16] = {s.x, u.x, -f.x, 0.f, s.y, u.y, -f.y, 0.f, s.z, u.z, -f.z, 0.f, 0.f, 0.f, 0.f, 1.f};
    m[12] = -(s.x * position.x + s.y * position.y + s.z * position.z);
    m[13] = -(u.x * position.x + u.y * position.y + u.z * position.z);
    m[14] = f.x * position.x + f.y * position.y + f.z * position.z;
    std::memcpy(lookat_mat, m, sizeof(m));
}

void Camera::look_at(vec3 target) {
    vec3 d = vnorm(vec3(target.x - position.x, target.y - position.y, target.z - position.z));
    pitch = (float)(std::asin(std::max(-1.f, std::min(1.f, d.y))) * 180.0 / kPi);
    yaw = (float)(std::atan2(d.z, d.x) * 180.0 / kPi);
    updateCameraVectors();
}

void Camera::processMouseMovement(float xoffset, float yoffset, bool constrainPitch) {
    yaw += xoffset * 0.1f;
    pitch += yoffset * 0.1f;
    if (constrainPitch) pitch = std::max(-89.f, std::min(89.f, pitch));
    updateCameraVectors();
}

void Camera::processMouseScroll(float yoffset) {
    zoom = std::max(1.f, std::min(120.f, zoom - yoffset));
    updateCameraVectors();
}

void Camera::inv_view_proj(float out[16]) const {
    double p[16], v[16], pv[16], inv[16];
    for (int i = 0; i < 16; i++) { p[i] = proj_mat[i]; v[i] = lookat_mat[i]; }
    mat_mul(p, v, pv);
    if (!mat_inv(pv, inv)) throw Error(FF_ERR_INVALID, "camera: singular proj * view");
    for (int i = 0; i < 16; i++) out[i] = (float)inv[i];
}

PerspectiveCamera::PerspectiveCamera() : aspectRatio(1.6f), yfov(90.f) {
    znear = 1.f;
    zfar = 100.f;
    set_projection();
    updateCameraVectors();
}
PerspectiveCamera::PerspectiveCamera(float a, float fov, float zn, float zf) : aspectRatio(a), yfov(fov) {
    znear = zn;
    zfar = zf;
    set_projection();
    updateCameraVectors();
}
PerspectiveCamera::PerspectiveCamera(vec3 pos, float a, float fov, float zn, float zf) : aspectRatio(a), yfov(fov) {
    position = pos;
    znear = zn;
    zfar = zf;
    set_projection();
    updateCameraVectors();
}
void PerspectiveCamera::set_projection() {
    // right-handed perspective, clip z in [-1, 1]; yfov in degrees
    double f = 1.0 / std::tan(yfov * kPi / 360.0);
    std::memset(proj_mat, 0, sizeof(proj_mat));
    proj_mat[0] = (float)(f / aspectRatio);
    proj_mat[5] = (float)f;
    proj_mat[10] = (float)(-(zfar + znear) / (double)(zfar - znear));
    proj_mat[11] = -1.f;
    proj_mat[14] = (float)(-(2.0 * zfar * znear) / (double)(zfar - znear));
}

// ------------------------------------------------------ EnvironmentLight --
EnvironmentLight::EnvironmentLight(vec3 c) : color(c), width(2), height(2), rgb(12) {
    for (int i = 0; i < 4; i++) { rgb[3 * i] = c.x; rgb[3 * i + 1] = c.y; rgb[3 * i + 2] = c.z; }
}

EnvironmentLight::EnvironmentLight(int w, int h, std::vector<float> data) : width(w), height(h), rgb(std::move(data)) {
    if (w <= 0 || h <= 0 || rgb.size() != (size_t)w * h * 3)
        throw Error(FF_ERR_INVALID, "EnvironmentLight: texel count does not match the size");
}

// Radiance RGBE reader (the format of hrdi/*.hdr, which the reference loads
// through stb_image, CRT/dTexture.cu:189-200): a "#?RADIANCE" (or "#?RGBE")
// first line, header lines up to an empty line with
// FORMAT=32-bit_rle_rgbe, the "-Y h +X w" resolution line, then h scanlines,
// each either new-style run-length encoded (bytes 2, 2, w >> 8, w & 255,
// then the four byte planes one after another, each as runs "128 + n, value"
// and dumps "n, n bytes") or flat (w four-byte pixels; always flat when
// w < 8 or w >= 32768).  Texel = mantissa * 2^(exponent - 136), zero for a
// zero exponent.  Rows are stored bottom first, as the reference's loader
// flips them (stbi_set_flip_vertically_on_load(true)).  Every byte read is
// checked: a truncated file, a run or dump past the end of its scanline, a
// zero-length dump or an image above 2^28 texels is an error.
EnvironmentLight::EnvironmentLight(const std::string& path) : filepath(path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error(FF_ERR_INVALID, "EnvironmentLight: cannot open " + path);
    std::string line;
    if (!std::getline(f, line) || (line.rfind("#?RADIANCE", 0) != 0 && line.rfind("#?RGBE", 0) != 0))
        throw Error(FF_ERR_INVALID, "EnvironmentLight: not a Radiance file: " + path);
    bool fmt = false, ended = false;
    while (std::getline(f, line)) {
        if (line.empty()) {
            ended = true;
            break;
        }
        if (line.rfind("FORMAT=32-bit_rle_rgbe", 0) == 0) fmt = true;
    }
    if (!fmt || !ended) throw Error(FF_ERR_INVALID, "EnvironmentLight: not a 32-bit_rle_rgbe file: " + path);
    if (!std::getline(f, line)) throw Error(FF_ERR_INVALID, "EnvironmentLight: no resolution line: " + path);
    int w = 0, h = 0, used = 0;
    if (std::sscanf(line.c_str(), "-Y %d +X %d%n", &h, &w, &used) != 2 || (size_t)used != line.size() || w <= 0 ||
        h <= 0 || w > 65536 || h > 65536 || (size_t)w * (size_t)h > ((size_t)1 << 28))
        throw Error(FF_ERR_INVALID, "EnvironmentLight: unsupported resolution line: " + line);
    width = w;
    height = h;
    rgb.assign((size_t)w * h * 3, 0.f);
    std::vector<unsigned char> scan((size_t)w * 4);
    auto byte = [&f]() -> unsigned char {
        int c = f.get();
        if (c == EOF) throw Error(FF_ERR_INVALID, "EnvironmentLight: truncated file");
        return (unsigned char)c;
    };
    for (int y = 0; y < h; y++) {
        unsigned char hd[4];
        if (!f.read((char*)hd, 4)) throw Error(FF_ERR_INVALID, "EnvironmentLight: truncated file");
        if (w >= 8 && w < 32768 && hd[0] == 2 && hd[1] == 2 && ((hd[2] << 8) | hd[3]) == w) {
            for (int ch = 0; ch < 4; ch++) {
                int x = 0;
                while (x < w) {
                    int c = byte();
                    if (c > 128) {
                        int cnt = c - 128;
                        if (x + cnt > w) throw Error(FF_ERR_INVALID, "EnvironmentLight: bad RLE run");
                        unsigned char v = byte();
                        for (int k = 0; k < cnt; k++) scan[(size_t)(x++) * 4 + ch] = v;
                    } else {
                        if (c == 0 || x + c > w) throw Error(FF_ERR_INVALID, "EnvironmentLight: bad RLE dump");
                        for (int k = 0; k < c; k++) scan[(size_t)(x++) * 4 + ch] = byte();
                    }
                }
            }
        } else {
            std::memcpy(scan.data(), hd, 4);
            if (!f.read((char*)scan.data() + 4, (std::streamsize)w * 4 - 4))
                throw Error(FF_ERR_INVALID, "EnvironmentLight: truncated file");
        }
        int row = h - 1 - y;  // bottom-up, as loaded with vertical flip
        for (int x = 0; x < w; x++) {
            const unsigned char* p = &scan[(size_t)x * 4];
            float s = p[3] ? std::ldexp(1.f, (int)p[3] - 136) : 0.f;
            float* o = &rgb[((size_t)row * w + x) * 3];
            o[0] = p[0] * s;
            o[1] = p[1] * s;
            o[2] = p[2] * s;
        }
    }
}

// ----------------------------------------------------------------- Scene --
int Scene::get_nmb_of_triangles() const {
    int n = 0;
    for (const Model& m : models)
        for (const Mesh& me : m.meshes) n += me.get_nmb_of_triangles();
    return n;
}

// ---------------------------------------------------------------- dScene --
dScene::dScene(Scene* s, int device, int max_prims_in_node, int builder) : h_scene(s) {
    if (!s || !s->camera || !s->environment_light) throw Error(FF_ERR_INVALID, "dScene: scene needs a camera and an environment light");
    // material table (dScene::load_materials, CRT/dScene.cpp:172-214)
    std::map<std::string, int> index;
    std::vector<float> mats;
    auto push_mat = [&](const Material& m) {
        float r[FF_MAT_FLOATS] = {(float)m.type, m.baseColorFactor.x, m.baseColorFactor.y, m.baseColorFactor.z,
                                  m.roughnessFactor, m.metallicFactor, m.fresnel.x, m.fresnel.y, m.fresnel.z, m.ior,
                                  (float)m.baseColorTexture, (float)m.roughnessTexture, (float)m.metallicTexture,
                                  (float)m.normalTexture, 0.f, 0.f};
        mats.insert(mats.end(), r, r + FF_MAT_FLOATS);
    };
    push_mat(Material{});  // index 0: default material for unknown names
    for (auto& kv : s->materials_loaded) {
        index[kv.first] = (int)(mats.size() / FF_MAT_FLOATS);
        push_mat(kv.second);
    }
    // triangles of every mesh in model order (dScene::load_models, CRT/dScene.cpp:82-170)
    std::vector<float> pos, nrm, uv, tan;
    std::vector<int32_t> mid;
    bool any_normal_map = false;
    for (auto& kv : s->materials_loaded) any_normal_map |= kv.second.normalTexture >= 0;
    for (const Model& m : s->models)
        for (const Mesh& me : m.meshes) {
            auto it = index.find(me.material);
            int mi = it == index.end() ? 0 : it->second;
            bool has_n = me.normals.size() == me.positions.size();
            bool has_t = me.texcoords.size() == me.positions.size();
            bool has_tan = me.tangents.size() == me.positions.size() && me.bitangents.size() == me.positions.size();
            if (any_normal_map && mi > 0 && mats[(size_t)mi * FF_MAT_FLOATS + 13] >= 0.f && !has_tan)
                throw Error(FF_ERR_INVALID, "dScene: mesh " + me.name + " has a normal texture but no tangents");
            for (size_t k = 0; k + 2 < me.indices.size(); k += 3) {
                for (int v = 0; v < 3; v++) {
                    uint32_t vi = me.indices[k + v];
                    if (vi >= me.positions.size()) throw Error(FF_ERR_INVALID, "dScene: index out of range in " + me.name);
                    const vec3& p = me.positions[vi];
                    pos.insert(pos.end(), {p.x, p.y, p.z});
                    vec3 n = has_n ? me.normals[vi] : vec3(0.f, 1.f, 0.f);
                    nrm.insert(nrm.end(), {n.x, n.y, n.z});
                    vec2 t = has_t ? me.texcoords[vi] : vec2{};
                    uv.insert(uv.end(), {t.x, t.y});
                    if (any_normal_map) {
                        vec3 a = has_tan ? me.tangents[vi] : vec3(1.f, 0.f, 0.f);
                        vec3 b = has_tan ? me.bitangents[vi] : vec3(0.f, 0.f, 1.f);
                        tan.insert(tan.end(), {a.x, a.y, a.z, b.x, b.y, b.z});
                    }
                }
                mid.push_back(mi);
            }
        }
    std::vector<const float*> tex_ptr;
    std::vector<int32_t> tw, th, tc;
    for (const Texture& t : s->textures_loaded) {
        tex_ptr.push_back(t.data.data());
        tw.push_back(t.width);
        th.push_back(t.height);
        tc.push_back(t.nrComponents);
    }
    ff_scene_desc d{};
    d.n_tris = (int32_t)mid.size();
    d.positions = pos.data();
    d.normals = nrm.data();
    d.uvs = uv.data();
    d.material_ids = mid.data();
    d.n_materials = (int32_t)(mats.size() / FF_MAT_FLOATS);
    d.materials = mats.data();
    d.n_textures = (int32_t)tex_ptr.size();
    d.textures = tex_ptr.data();
    d.tex_w = tw.data();
    d.tex_h = th.data();
    d.tex_c = tc.data();
    d.max_prims_in_node = max_prims_in_node;
    d.builder = builder;
    d.tangents = any_normal_map ? tan.data() : nullptr;
    check(ff_context_create(device, &ctx), "dScene: context");
    try {
        check(ff_scene_create(ctx, &d, &scene), "dScene: scene");
        const EnvironmentLight* e = s->environment_light;
        check(ff_envmap_create(ctx, e->get_texels().data(), e->get_tex_width(), e->get_tex_height(), &env),
              "dScene: environment light");
        update();
    } catch (...) {
        if (env) ff_envmap_destroy(env);
        if (scene) ff_scene_destroy(scene);
        ff_context_destroy(ctx);
        throw;
    }
}

dScene::~dScene() {
    if (env) ff_envmap_destroy(env);
    if (scene) ff_scene_destroy(scene);
    if (ctx) ff_context_destroy(ctx);
}

int dScene::get_nmb_nodes() const {
    int32_t n = 0;
    check(ff_scene_info(scene, nullptr, &n, nullptr), "dScene::get_nmb_nodes");
    return n;
}

void dScene::update() {
    const Camera* c = h_scene->camera;
    c->inv_view_proj(camera.inv_view_proj);
    camera.lens_radius = c->lens_radius;
    camera.focal_distance = c->focal_distance;
    camera.exposure = c->exposure;
    camera.pad = 0.f;
}

// ------------------------------------------------------------ PathTracer --
PathTracer::PathTracer(int w, int h) : width(w), height(h) {
    if (w <= 0 || h <= 0) throw Error(FF_ERR_INVALID, "PathTracer: bad size");
    if (hipMalloc((void**)&d_film, sizeof(float) * 4 * (size_t)w * h) != hipSuccess)
        throw Error(FF_ERR_ALLOC, "PathTracer: film allocation");
    clear_buffer();
}

PathTracer::~PathTracer() {
    if (d_film) (void)hipFree(d_film);
}

void PathTracer::clear_buffer() {
    if (hipMemset(d_film, 0, sizeof(float) * 4 * (size_t)width * height) != hipSuccess)
        throw Error(FF_ERR_HIP, "PathTracer::clear_buffer");
    n_samples = 0;
}

void PathTracer::draw(dScene* s) {
    if (!s) throw Error(FF_ERR_INVALID, "PathTracer::draw: NULL scene");
    s->update();
    ff_render_req r{};
    r.width = width;
    r.height = height;
    r.spp = samples_per_draw;
    r.max_depth = (int32_t)max_path_length;
    r.seed = seed;
    r.spp_batch = samples_in_flight;
    r.sample_offset = n_samples;
    r.flags = 0;
    check(ff_render(s->get_context(), s->get_triangles(), s->get_environment_light(), s->get_camera(), &r, d_film,
                    nullptr),
          "PathTracer::draw");
    n_samples += samples_per_draw;
}

std::vector<float> PathTracer::get_image(dScene* s) const {
    float* d_avg = nullptr;
    size_t n = (size_t)width * height;
    if (hipMalloc((void**)&d_avg, sizeof(float) * 3 * n) != hipSuccess) throw Error(FF_ERR_ALLOC, "get_image");
    std::vector<float> out(3 * n);
    ff_status st = ff_film_resolve(d_film, width, height, s->get_camera()->exposure, d_avg, nullptr, nullptr);
    hipError_t e = st == FF_OK ? hipMemcpy(out.data(), d_avg, sizeof(float) * 3 * n, hipMemcpyDeviceToHost) : hipSuccess;
    (void)hipFree(d_avg);
    check(st, "PathTracer::get_image");
    if (e != hipSuccess) throw Error(FF_ERR_HIP, "PathTracer::get_image: copy");
    return out;
}

std::vector<uint8_t> PathTracer::get_rgba8(dScene* s) const {
    uint8_t* d_rgba = nullptr;
    size_t n = (size_t)width * height;
    if (hipMalloc((void**)&d_rgba, 4 * n) != hipSuccess) throw Error(FF_ERR_ALLOC, "get_rgba8");
    std::vector<uint8_t> out(4 * n);
    ff_status st = ff_film_resolve(d_film, width, height, s->get_camera()->exposure, nullptr, d_rgba, nullptr);
    hipError_t e = st == FF_OK ? hipMemcpy(out.data(), d_rgba, 4 * n, hipMemcpyDeviceToHost) : hipSuccess;
    (void)hipFree(d_rgba);
    check(st, "PathTracer::get_rgba8");
    if (e != hipSuccess) throw Error(FF_ERR_HIP, "PathTracer::get_rgba8: copy");
    return out;
}

namespace {
uint32_t crc_table[256];
bool crc_ready = false;
uint32_t crc32(const uint8_t* p, size_t n, uint32_t c = 0xffffffffu) {
    if (!crc_ready) {
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t v = i;
            for (int k = 0; k < 8; k++) v = (v & 1u) ? 0xedb88320u ^ (v >> 1) : v >> 1;
            crc_table[i] = v;
        }
        crc_ready = true;
    }
    for (size_t i = 0; i < n; i++) c = crc_table[(c ^ p[i]) & 0xffu] ^ (c >> 8);
    return c;
}
void put_be32(std::vector<uint8_t>& o, uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) o.push_back((uint8_t)(v >> s));
}
void png_chunk(std::vector<uint8_t>& o, const char* type, const std::vector<uint8_t>& data) {
    put_be32(o, (uint32_t)data.size());
    size_t start = o.size();
    o.insert(o.end(), type, type + 4);
    o.insert(o.end(), data.begin(), data.end());
    put_be32(o, crc32(o.data() + start, o.size() - start) ^ 0xffffffffu);
}
}  // namespace

void write_png(const std::string& path, const uint8_t* rgba, int width, int height) {
    if (!rgba || width <= 0 || height <= 0) throw Error(FF_ERR_INVALID, "write_png: empty image");
    std::vector<uint8_t> raw;  // filter byte 0 + row
    size_t row = (size_t)width * 4;
    raw.reserve((row + 1) * (size_t)height);
    for (int y = 0; y < height; y++) {
        raw.push_back(0);
        raw.insert(raw.end(), rgba + (size_t)y * row, rgba + (size_t)(y + 1) * row);
    }
    std::vector<uint8_t> z = {0x78, 0x01};  // zlib header: deflate, 32 KiB window, no dictionary
    uint32_t a = 1, b = 0;
    for (uint8_t v : raw) {
        a = (a + v) % 65521u;
        b = (b + a) % 65521u;
    }
    for (size_t off = 0; off < raw.size() || off == 0;) {
        size_t n = std::min<size_t>(65535, raw.size() - off);
        bool last = off + n == raw.size();
        z.push_back(last ? 1 : 0);  // BFINAL, BTYPE = 00 (stored)
        z.push_back((uint8_t)(n & 0xff));
        z.push_back((uint8_t)(n >> 8));
        z.push_back((uint8_t)(~n & 0xff));
        z.push_back((uint8_t)((~n >> 8) & 0xff));
        z.insert(z.end(), raw.begin() + (std::ptrdiff_t)off, raw.begin() + (std::ptrdiff_t)(off + n));
        off += n;
        if (last) break;
    }
    put_be32(z, (b << 16) | a);
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, (uint32_t)width);
    put_be32(ihdr, (uint32_t)height);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});  // 8 bits, RGBA, deflate, filter 0, no interlace
    png_chunk(out, "IHDR", ihdr);
    png_chunk(out, "IDAT", z);
    png_chunk(out, "IEND", {});
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw Error(FF_ERR_INVALID, "write_png: cannot write " + path);
    size_t wr = std::fwrite(out.data(), 1, out.size(), f);
    std::fclose(f);
    if (wr != out.size()) throw Error(FF_ERR_INVALID, "write_png: short write to " + path);
}

}  // namespace firefly
