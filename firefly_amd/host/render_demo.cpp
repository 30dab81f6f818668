// render_demo.cpp -- renders a scene file through the C++ host interface
// (Scene -> dScene -> PathTracer::draw, as the reference's main loop drives
// it, CRT/main.cpp:245-275) and writes the averaged image as a PFM.
//
// usage: render_demo <scene.ffs> <out.pfm|out.png> <spp> <max_path_length> [samples_per_draw] [model.glb]
//        render_demo --dump-glb <model.glb> <out.bin>   (no GPU: the loaded meshes and materials)
//        render_demo --dump-hdr <env.hdr> <out.bin>     (no GPU: the environment map as loaded)
//
// A model.glb argument adds that file's meshes (Scene::load) to the scene's.
// The dump: int32 n_meshes; per mesh int32 name length, name, int32 material
// name length, name, int32 n_vertices, n_indices, has_texcoords, has_tangents,
// float positions[n][3], normals[n][3], (texcoords[n][2]), (tangents[n][3],
// bitangents[n][3]), uint32 indices[]; then int32 n_materials; per material
// int32 name length, name, float base[3], roughness, metallic, int32 base
// colour, roughness, metallic and normal texture (-1 none);
// then int32 n_textures; per texture int32 width, height, channels and
// float texels[height][width][channels] (bottom row first).
// The environment dump: int32 width, height, float texels[height][width][3]
// (bottom row first, as EnvironmentLight(path) stores them).
//
// Scene file (little endian): "FFS1", int32 n_meshes, n_materials, env_w,
// env_h, width, height, uint32 seed, float eye[3], target[3], yfov_deg,
// float materials[n_materials][16], float env[env_h][env_w][3], then per mesh
// int32 material, int32 n_tris, float pos[n][9], nrm[n][9], uv[n][6].
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "firefly.h"

using namespace firefly;

namespace {

template <typename T>
void rd(std::ifstream& f, T* p, size_t n) {
    if (!f.read((char*)p, (std::streamsize)(sizeof(T) * n))) throw Error(FF_ERR_INVALID, "scene file truncated");
}

void put_str(FILE* o, const std::string& s) {
    int32_t n = (int32_t)s.size();
    std::fwrite(&n, 4, 1, o);
    std::fwrite(s.data(), 1, s.size(), o);
}

int dump_glb(const char* glb, const char* out) {
    Scene scene;
    int tris = scene.load(glb);
    FILE* o = std::fopen(out, "wb");
    if (!o) throw Error(FF_ERR_INVALID, std::string("cannot write ") + out);
    int32_t nm = 0;
    for (const Model& m : scene.models) nm += (int32_t)m.meshes.size();
    std::fwrite(&nm, 4, 1, o);
    for (const Model& m : scene.models)
        for (const Mesh& me : m.meshes) {
            put_str(o, me.name);
            put_str(o, me.material);
            int32_t h[4] = {(int32_t)me.positions.size(), (int32_t)me.indices.size(), me.texcoords.empty() ? 0 : 1,
                            me.tangents.empty() ? 0 : 1};
            std::fwrite(h, 4, 4, o);
            std::fwrite(me.positions.data(), sizeof(vec3), me.positions.size(), o);
            std::fwrite(me.normals.data(), sizeof(vec3), me.normals.size(), o);
            if (h[2]) std::fwrite(me.texcoords.data(), sizeof(vec2), me.texcoords.size(), o);
            if (h[3]) {
                std::fwrite(me.tangents.data(), sizeof(vec3), me.tangents.size(), o);
                std::fwrite(me.bitangents.data(), sizeof(vec3), me.bitangents.size(), o);
            }
            std::fwrite(me.indices.data(), 4, me.indices.size(), o);
        }
    int32_t nmat = (int32_t)scene.materials_loaded.size();
    std::fwrite(&nmat, 4, 1, o);
    for (const auto& kv : scene.materials_loaded) {
        put_str(o, kv.first);
        float r[5] = {kv.second.baseColorFactor.x, kv.second.baseColorFactor.y, kv.second.baseColorFactor.z,
                      kv.second.roughnessFactor, kv.second.metallicFactor};
        std::fwrite(r, 4, 5, o);
        int32_t t[4] = {kv.second.baseColorTexture, kv.second.roughnessTexture, kv.second.metallicTexture,
                        kv.second.normalTexture};
        std::fwrite(t, 4, 4, o);
    }
    int32_t ntex = (int32_t)scene.textures_loaded.size();
    std::fwrite(&ntex, 4, 1, o);
    for (const Texture& t : scene.textures_loaded) {
        int32_t h[3] = {t.width, t.height, t.nrComponents};
        std::fwrite(h, 4, 3, o);
        std::fwrite(t.data.data(), 4, t.data.size(), o);
    }
    std::fclose(o);
    std::printf("loaded %s: %d meshes, %d triangles, %d materials\n", glb, nm, tris, nmat);
    return 0;
}

int dump_hdr(const char* hdr, const char* out) {
    EnvironmentLight env{std::string(hdr)};
    FILE* o = std::fopen(out, "wb");
    if (!o) throw Error(FF_ERR_INVALID, std::string("cannot write ") + out);
    int32_t wh[2] = {env.get_tex_width(), env.get_tex_height()};
    std::fwrite(wh, 4, 2, o);
    std::fwrite(env.get_texels().data(), 4, env.get_texels().size(), o);
    std::fclose(o);
    std::printf("loaded %s: %dx%d\n", hdr, wh[0], wh[1]);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc == 4 && (std::strcmp(argv[1], "--dump-glb") == 0 || std::strcmp(argv[1], "--dump-hdr") == 0)) {
        try {
            return argv[1][7] == 'g' ? dump_glb(argv[2], argv[3]) : dump_hdr(argv[2], argv[3]);
        } catch (const Error& e) {
            std::fprintf(stderr, "render_demo: %s (status %d)\n", e.what(), (int)e.status);
            return 1;
        }
    }
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s scene.ffs out.pfm spp max_path_length [samples_per_draw]\n", argv[0]);
        return 2;
    }
    try {
        std::ifstream f(argv[1], std::ios::binary);
        if (!f) throw Error(FF_ERR_INVALID, std::string("cannot open ") + argv[1]);
        char magic[4];
        rd(f, magic, 4);
        if (std::memcmp(magic, "FFS1", 4) != 0) throw Error(FF_ERR_INVALID, "not an FFS1 scene file");
        int32_t hdr[6];
        rd(f, hdr, 6);
        int n_meshes = hdr[0], n_mats = hdr[1], env_w = hdr[2], env_h = hdr[3], width = hdr[4], height = hdr[5];
        uint32_t seed;
        rd(f, &seed, 1);
        float cam[7];
        rd(f, cam, 7);
        std::vector<float> mats((size_t)n_mats * 16);
        rd(f, mats.data(), mats.size());
        std::vector<float> env((size_t)env_w * env_h * 3);
        rd(f, env.data(), env.size());

        Scene scene;
        for (int m = 0; m < n_mats; m++) {
            const float* r = &mats[(size_t)m * 16];
            Material mt;
            mt.name = "material" + std::to_string(1000 + m);
            mt.type = (int)r[0];
            mt.baseColorFactor = vec3(r[1], r[2], r[3]);
            mt.roughnessFactor = r[4];
            mt.metallicFactor = r[5];
            mt.fresnel = vec3(r[6], r[7], r[8]);
            mt.ior = r[9];
            scene.add_material(mt);
        }
        Model model;
        for (int k = 0; k < n_meshes; k++) {
            int32_t mi[2];
            rd(f, mi, 2);
            int n = mi[1];
            std::vector<float> pos((size_t)n * 9), nrm((size_t)n * 9), uv((size_t)n * 6);
            rd(f, pos.data(), pos.size());
            rd(f, nrm.data(), nrm.size());
            rd(f, uv.data(), uv.size());
            Mesh me;
            me.name = "mesh" + std::to_string(k);
            me.material = "material" + std::to_string(1000 + mi[0]);
            for (int t = 0; t < 3 * n; t++) {
                me.positions.push_back(vec3(pos[3 * t], pos[3 * t + 1], pos[3 * t + 2]));
                me.normals.push_back(vec3(nrm[3 * t], nrm[3 * t + 1], nrm[3 * t + 2]));
                me.texcoords.push_back(vec2{uv[2 * t], uv[2 * t + 1]});
                me.indices.push_back((uint32_t)t);
            }
            model.meshes.push_back(std::move(me));
        }
        scene.add_model(std::move(model));
        if (argc > 6) scene.load(argv[6]);
        EnvironmentLight light(env_w, env_h, env);
        scene.set_environment_light(&light);
        PerspectiveCamera camera(vec3(cam[0], cam[1], cam[2]), (float)width / (float)height, cam[6], 0.1f, 100.f);
        camera.look_at(vec3(cam[3], cam[4], cam[5]));
        scene.set_camera(&camera);

        dScene ds(&scene);
        PathTracer pt(width, height);
        pt.seed = seed;
        pt.max_path_length = (uint32_t)std::atoi(argv[4]);
        pt.samples_per_draw = argc > 5 ? std::atoi(argv[5]) : 1;
        int spp = std::atoi(argv[3]);
        while (pt.samples() < spp) pt.draw(&ds);
        std::string out_path = argv[2];
        if (out_path.size() > 4 && out_path.compare(out_path.size() - 4, 4, ".png") == 0) {
            std::vector<uint8_t> rgba = pt.get_rgba8(&ds);  // the reference's resolve + tone map
            write_png(out_path, rgba.data(), width, height);
            std::printf("rendered %dx%d, %d spp into %s\n", width, height, pt.samples(), out_path.c_str());
            return 0;
        }
        std::vector<float> img = pt.get_image(&ds);

        FILE* o = std::fopen(argv[2], "wb");
        if (!o) throw Error(FF_ERR_INVALID, std::string("cannot write ") + argv[2]);
        std::fprintf(o, "PF\n%d %d\n-1.0\n", width, height);
        for (int y = height - 1; y >= 0; y--)  // PFM rows run bottom to top
            std::fwrite(&img[(size_t)y * width * 3], sizeof(float), (size_t)width * 3, o);
        std::fclose(o);
        std::printf("rendered %dx%d, %d spp, %d tris, %d BVH nodes\n", width, height, pt.samples(),
                    scene.get_nmb_of_triangles(), ds.get_nmb_nodes());
        return 0;
    } catch (const Error& e) {
        std::fprintf(stderr, "render_demo: %s (status %d)\n", e.what(), (int)e.status);
        return 1;
    }
}
