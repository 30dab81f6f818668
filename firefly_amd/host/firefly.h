// firefly.h -- C++ host interface of the MI355X wavefront path tracer.
//
// Keeps the reference's host types and method shapes (JakeKurtz/Firefly,
// CUDA-RayTracer/ = CRT/) without OpenGL, assimp or glm, and drives the
// gfx950 kernels only through the C ABI (include/ff/ff_abi.h):
//   PathTracer(w, h), draw(dScene*), clear_buffer()   CRT/PathTracer.h:11-62
//   dScene(Scene*), get_camera/get_environment_light,
//     update()                                         CRT/dScene.h:13-65
//   Scene: models, materials_loaded, environment_light,
//     set_camera / set_environment_light,
//     get_nmb_of_triangles                             CRT/Scene.h:10-47
//   Camera / PerspectiveCamera(pos, aspect, yfov,
//     znear, zfar)                                     CRT/Camera.h:17-62, CRT/PerspectiveCamera.h:4-14
//   Material                                           CRT/Material.h:10-37
//   EnvironmentLight(color) / EnvironmentLight(path),
//     get_texture_filepath/get_color/get_tex_width/
//     get_tex_height                                   CRT/EnvironmentLight.h:306-431
// Failures throw firefly::Error carrying the ff_status and message (the
// reference exits the process through checkCudaErrors, CRT/CudaHelpers.cpp:3-11).
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ff/ff_abi.h"

namespace firefly {

struct vec2 { float x = 0.f, y = 0.f; };
struct vec3 {
    float x = 0.f, y = 0.f, z = 0.f;
    vec3() = default;
    vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    explicit vec3(float s) : x(s), y(s), z(s) {}
};

class Error : public std::runtime_error {
public:
    Error(ff_status st, const std::string& msg) : std::runtime_error(msg), status(st) {}
    ff_status status;
};
// throws Error when st != FF_OK
void check(ff_status st, const char* what);

// Material (CRT/Material.h:10-37) plus the material class selecting the
// shading queue.  Defaults follow the reference loader: roughness 1,
// metallic 0, fresnel 0.04 (CRT/Scene.cpp:203-204, CRT/dScene.cpp:192-200).
struct Material {
    std::string name;
    int type = FF_MAT_MIX;
    vec3 baseColorFactor{0.5f, 0.5f, 0.5f};
    vec3 emissiveColorFactor{0.f, 0.f, 0.f};
    vec3 fresnel{0.04f, 0.04f, 0.04f};
    float roughnessFactor = 1.f;
    float metallicFactor = 0.f;
    float ior = 1.5f;
    int baseColorTexture = -1;  // index into Scene::textures_loaded
    int roughnessTexture = -1;
    int metallicTexture = -1;
    int normalTexture = -1;     // tangent-space normal map (needs mesh tangents)
};

// Float texture in host memory ([h][w][channels]).
struct Texture {
    std::string filepath;
    int width = 0, height = 0, nrComponents = 0;
    std::vector<float> data;
};

// Indexed triangle mesh (CRT/Mesh.h) with one material.
struct Mesh {
    std::string name;
    std::vector<vec3> positions, normals;
    std::vector<vec2> texcoords;
    std::vector<vec3> tangents, bitangents;  // per vertex (Vertex, CRT/Mesh.h:16-17), may be empty
    std::vector<uint32_t> indices;
    std::string material;
    int get_nmb_of_triangles() const { return (int)(indices.size() / 3); }
};

struct Model {
    std::vector<Mesh> meshes;
};

// Camera (CRT/Camera.h:17-62): yaw/pitch camera producing view and projection
// matrices (column-major float[16]).
class Camera {
public:
    virtual ~Camera() = default;
    std::string name;
    float zfar = 100.f, znear = 0.1f;
    vec3 position{0.f, 0.f, 0.f};
    vec3 front{0.f, 0.f, -1.f}, up{0.f, 1.f, 0.f}, right{1.f, 0.f, 0.f};
    vec3 worldUp{0.f, 1.f, 0.f};
    float lookat_mat[16];
    float proj_mat[16];
    float yaw = -90.f, pitch = 0.f;
    float zoom = 53.f;
    float exposure = 1.f;
    float lens_radius = 0.f;
    float focal_distance = 35.f;
    float d = 50.f;

    const float* getViewMatrix() const { return lookat_mat; }
    const float* getProjMatrix() const { return proj_mat; }
    // Points the camera at `target` (sets yaw/pitch, recomputes the view).
    void look_at(vec3 target);
    void processMouseMovement(float xoffset, float yoffset, bool constrainPitch = true);
    void processMouseScroll(float yoffset);
    // inverse(proj * view), column-major: the dCamera matrix (CRT/dScene.cpp:381-391)
    void inv_view_proj(float out[16]) const;

protected:
    void updateCameraVectors();
};

// PerspectiveCamera (CRT/PerspectiveCamera.h:4-14); yfov in degrees.
class PerspectiveCamera : public Camera {
public:
    PerspectiveCamera();
    PerspectiveCamera(float aspectRatio, float yfov, float znear, float zfar);
    PerspectiveCamera(vec3 position, float aspectRatio, float yfov, float znear, float zfar);

private:
    float aspectRatio;
    float yfov;
    void set_projection();
};

// EnvironmentLight (CRT/EnvironmentLight.h:306-431) without the GL IBL
// pre-filtering: a constant colour or an equirectangular float map.
class EnvironmentLight {
public:
    explicit EnvironmentLight(vec3 color);
    // Radiance RGBE (.hdr) file; rows are stored bottom-up as the reference
    // loads them (stbi_set_flip_vertically_on_load, CRT/dTexture.cu:193).
    explicit EnvironmentLight(const std::string& path);
    EnvironmentLight(int width, int height, std::vector<float> rgb);

    std::string get_texture_filepath() const { return filepath; }
    vec3 get_color() const { return color; }
    int get_tex_width() const { return width; }
    int get_tex_height() const { return height; }
    const std::vector<float>& get_texels() const { return rgb; }

private:
    std::string filepath;
    vec3 color{1.f, 1.f, 1.f};
    int width = 0, height = 0;
    std::vector<float> rgb;
};

// Writes 8-bit RGBA rows (top row first) as a PNG file: colour type 6, no
// filtering, the zlib stream in stored (uncompressed) deflate blocks, CRC-32
// per chunk and Adler-32 over the image data (PNG specification, RFC 1950,
// RFC 1951).  The step after PathTracer::get_rgba8 in place of the
// reference's GL surface (CRT/kernel.cu:447-454).
void write_png(const std::string& path, const uint8_t* rgba, int width, int height);
// Decodes a PNG image (bit depth 8 or 16, colour types 0/2/3/4/6, no
// interlace) into 8-bit samples, top row first; channels: 1-4 as stored
// (palette images expand to RGB or RGBA).  Throws Error on a malformed or
// unsupported file.
void decode_png(const uint8_t* data, size_t n, int& width, int& height, int& channels, std::vector<uint8_t>& pixels);

// Scene (CRT/Scene.h:10-47): models, materials, textures, camera, light.
class Scene {
public:
    std::vector<Model> models;
    EnvironmentLight* environment_light = nullptr;
    Camera* camera = nullptr;
    std::vector<Texture> textures_loaded;
    std::map<std::string, Material> materials_loaded;

    void set_camera(Camera* c) { camera = c; }
    void set_environment_light(EnvironmentLight* e) { environment_light = e; }
    void add_model(Model m) { models.push_back(std::move(m)); }
    void add_material(const Material& m) { materials_loaded[m.name] = m; }
    // Adds the meshes and materials of a binary glTF 2.0 file (.glb), as the
    // reference's assimp import does (Scene::load, CRT/Scene.cpp:20-221):
    // vertices placed by their node transforms, v flipped, materials by name
    // with the loader's forced roughness 1 / metallic 0, embedded PNG base
    // colour and roughness textures.  Returns the number of triangles added;
    // throws Error on a malformed or unsupported file.
    int load(const std::string& path);
    int get_nmb_of_triangles() const;
};

// dScene (CRT/dScene.h:13-65): device copy of a Scene -- host SAH BVH,
// triangles in BVH order, materials, textures, FP64 environment CDF and the
// camera.  Owns a device context.
class dScene {
public:
    // builder: FF_BVH_SAH (the reference's host SAH tree) or FF_BVH_LBVH (built on the GPU)
    explicit dScene(Scene* scene, int device = 0, int max_prims_in_node = 8, int builder = FF_BVH_SAH);
    ~dScene();
    dScene(const dScene&) = delete;
    dScene& operator=(const dScene&) = delete;

    const ff_camera* get_camera() const { return &camera; }
    ff_scene* get_triangles() const { return scene; }
    ff_envmap* get_environment_light() const { return env; }
    ff_context* get_context() const { return ctx; }
    int get_nmb_nodes() const;
    // re-reads the host camera (CRT/dScene.cpp:55-58)
    void update();

private:
    Scene* h_scene;
    ff_context* ctx = nullptr;
    ff_scene* scene = nullptr;
    ff_envmap* env = nullptr;
    ff_camera camera{};
};

// PathTracer (CRT/PathTracer.h:11-62): progressive film; each draw() adds
// samples_per_draw samples per pixel (paths of at most max_path_length
// vertices), continuing the sample sequence of earlier draws.
class PathTracer {
public:
    PathTracer(int w, int h);
    ~PathTracer();
    PathTracer(const PathTracer&) = delete;
    PathTracer& operator=(const PathTracer&) = delete;

    void draw(dScene* s);
    void clear_buffer();

    uint32_t seed = 0;
    int samples_per_draw = 1;
    int samples_in_flight = 1;   // film layers (spp_batch)
    uint32_t max_path_length = 3;  // MAX_PATH_LENGTH (CRT/PathTracer.h:21)

    int samples() const { return n_samples; }
    // device film [h][w][4]: radiance sums and sample counts
    const float* film() const { return d_film; }
    // averaged radiance (host, [h][w][3]) and the tonemapped RGBA8 image of
    // the reference's resolve (CRT/kernel.cu:447-454)
    std::vector<float> get_image(dScene* s) const;
    std::vector<uint8_t> get_rgba8(dScene* s) const;

private:
    int width, height;
    int n_samples = 0;
    float* d_film = nullptr;
};

}  // namespace firefly
