// vlb/light_baker.hpp -- C++ facade with the reference baker's interface.
//
// Mirrors vlb::LightBaker (src/baker/light_baker.hpp:16-54):
//   LightBaker(std::string& sceneName)   -> scene upload + BVH build + grid
//   probePositionsFromBoudingBox(bounds) -> light_baker.cpp:80-101
//   bake()                               -> light_baker.cpp:287-328
//   serialize()                          -> light_baker.cpp:375-401
// on top of the C ABI in vlb_bake.h.  Errors are rethrown as
// std::runtime_error, as the reference's main (src/baker/main.cpp:26-40)
// expects.  Defaults reproduce the reference bake: 7x7x7 probes over the
// scene bounds (which, as in the reference's zero-initialised bounds
// (scene_manager.hpp:170), always include the origin), breadth-first probe
// order, the 3141 x 1000 equirect ray grid of setEnvShpereRadius(500), 16
// coefficients, env_map.rchit shading with sRGB + RGBA8 quantisation.
// Unlike the reference, every probe starts from zero (its SSBO is zeroed
// once and accumulates across probes, light_baker.cpp:110,121).
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../vlb_bake.h"

namespace vlb {

struct vec3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
};

// Bake settings of LightBaker (the reference hard-codes them, see below).
struct BakeSettings {
    std::array<uint32_t, 3> probesCount3D{7, 7, 7};
    int32_t order = VLB_ORDER_REFERENCE_BFS;
    int32_t dirMode = VLB_DIRS_EQUIRECT;
    uint32_t rays = 4096;           // Fibonacci rays per probe
    uint32_t equirectW = 3141;      // pi * 500 * 2 (env_map_generator.cpp:24-28)
    uint32_t equirectH = 1000;
    uint32_t shOrder = 3;           // 16 coefficients (sh.comp:35)
    int32_t shConstants = VLB_SH_LITERAL;
    int32_t precision = VLB_PREC_FP32;
    int32_t shadeMode = VLB_SHADE_REFERENCE;
    vlb_ref_shading ref{{1.0f, 10.0f, 1.0f}, 0.005f, 0.0f, 0.5f, 0.5f, 16.0f,
                        VLB_REF_SRGB | VLB_REF_UNORM8};
    std::vector<vlb_light> lights;
    float shadowBias = 1e-3f;
    vlb_sky sky{{2.0f, 2.0f, 2.0f}, {0.5f, 0.5f, 0.5f}, {0.0f, 1.0f, 0.0f}, {50.0f, 50.0f, 50.0f}, 200.0f};
    float tMin = 0.001f, tMax = 10000.0f;  // env_map.rgen:22-23
};

class LightBaker {
public:
    using Settings = BakeSettings;

    explicit LightBaker(const std::string& sceneName, int device = 0, Settings s = Settings())
        : name_(sceneName), settings_(std::move(s))
    {
        check(vlb_ctx_create(device, 0, &ctx_), "vlb_ctx_create");
        // scene bounds as Scene_t::getBounds (scene_manager.cpp:497-507):
        // zero-initialised, so they always contain the origin
        std::array<vec3, 2> bounds;
        if (!load_gltf(sceneName, bounds)) {
            load_scene(sceneName);
            vlb_scene_info info{};
            check(vlb_scene_info_get(ctx_, scene_, &info), "vlb_scene_info_get");
            bounds[0] = {std::min(0.0f, info.bounds_min[0]), std::min(0.0f, info.bounds_min[1]),
                         std::min(0.0f, info.bounds_min[2])};
            bounds[1] = {std::max(0.0f, info.bounds_max[0]), std::max(0.0f, info.bounds_max[1]),
                         std::max(0.0f, info.bounds_max[2])};
        }
        probePositions_ = probePositionsFromBoudingBox(bounds);
    }
    LightBaker(const LightBaker&) = delete;
    LightBaker& operator=(const LightBaker&) = delete;
    ~LightBaker()
    {
        if (ctx_) vlb_ctx_destroy(ctx_);
    }

    std::vector<vec3> probePositionsFromBoudingBox(std::array<vec3, 2> bounds)
    {
        const float lo[3] = {bounds[0].x, bounds[0].y, bounds[0].z};
        const float hi[3] = {bounds[1].x, bounds[1].y, bounds[1].z};
        const auto& c = settings_.probesCount3D;
        std::vector<float> xyz((size_t)c[0] * c[1] * c[2] * 3);
        float step[3];
        check(vlb_probe_grid(lo, hi, c.data(), settings_.order, xyz.data(), step), "vlb_probe_grid");
        gridStep_ = {step[0], step[1], step[2]};
        std::vector<vec3> out(xyz.size() / 3);
        for (size_t i = 0; i < out.size(); ++i) out[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
        return out;
    }

    void bake()
    {
        vlb_bake_params p = params();
        const size_t K = settings_.shOrder == 2 ? 9 : 16;
        std::vector<float> out(probePositions_.size() * K * 3);
        check(vlb_bake(ctx_, scene_, &p, out.data(), out.size()), "vlb_bake");
        coeffs_.resize(out.size() * sizeof(float));
        std::memcpy(coeffs_.data(), out.data(), coeffs_.size());
        numCoefs_ = (uint32_t)K;
    }

    // Writes "baked_" + sceneName (light_baker.cpp:399): the input glTF JSON
    // (or an empty asset for synthetic scenes) with the coefficient buffer.
    void serialize()
    {
        if (coeffs_.empty()) throw std::runtime_error("serialize() before bake()");
        const char* json = gltfText_.c_str();
        const float step[3] = {gridStep_.x, gridStep_.y, gridStep_.z};
        size_t need = 0;
        const float* c = reinterpret_cast<const float*>(coeffs_.data());
        const uint32_t P = (uint32_t)probePositions_.size();
        check(vlb_gltf_bake_json(json, c, P, numCoefs_, step, nullptr, 0, &need), "vlb_gltf_bake_json");
        std::string text(need + 1, '\0');
        check(vlb_gltf_bake_json(json, c, P, numCoefs_, step, &text[0], text.size(), &need),
              "vlb_gltf_bake_json");
        text.resize(need);
        std::ofstream o(outputName());
        if (!o) throw std::runtime_error("cannot write " + outputName());
        o << text;
    }

    std::string outputName() const
    {
        std::string base = name_;
        for (char& ch : base)
            if (ch == ':' || ch == '/') ch = '_';
        if (base.find(".gltf") == std::string::npos) base += ".gltf";
        return "baked_" + base;
    }

    const std::vector<uint8_t>& coefficients() const { return coeffs_; }
    const std::vector<vec3>& probePositions() const { return probePositions_; }
    vec3 gridStep() const { return gridStep_; }
    Settings& settings() { return settings_; }
    vlb_ctx* context() { return ctx_; }
    uint32_t scene() const { return scene_; }

private:
    static void check(vlb_status s, const char* what, vlb_ctx* ctx = nullptr)
    {
        if (s != VLB_OK) {
            std::ostringstream m;
            m << what << " failed with status " << s;
            if (ctx) m << ": " << vlb_last_error(ctx);
            throw std::runtime_error(m.str());
        }
    }

    // LightBaker::LightBaker (light_baker.cpp:34-60): a .gltf / .glb asset
    bool load_gltf(const std::string& name, std::array<vec3, 2>& bounds)
    {
        if (name.find(".gltf") == std::string::npos && name.find(".glb") == std::string::npos) return false;
        vlb_host_scene* hs = nullptr;
        char* json = nullptr;
        check(vlb_gltf_load(name.c_str(), &hs, &json), ("vlb_gltf_load(" + name + ")").c_str());
        vlb_status st = vlb_scene_upload(ctx_, hs->xyz, hs->nverts, hs->idx, hs->ntris, hs->normals,
                                         hs->albedo, &scene_);
        bounds[0] = {hs->bounds_min[0], hs->bounds_min[1], hs->bounds_min[2]};
        bounds[1] = {hs->bounds_max[0], hs->bounds_max[1], hs->bounds_max[2]};
        gltfText_ = json;
        vlb_text_free(json);
        vlb_host_scene_free(hs);
        check(st, "vlb_scene_upload", ctx_);
        return true;
    }

    void load_scene(const std::string& name)
    {
        // synthetic "<kind>[:<seed>]"
        std::string kind = name, seed_s = "0";
        const size_t colon = name.find(':');
        if (colon != std::string::npos) {
            kind = name.substr(0, colon);
            seed_s = name.substr(colon + 1);
        }
        int32_t k;
        if (kind == "room") k = VLB_SCENE_ROOM;
        else if (kind == "terrain") k = VLB_SCENE_TERRAIN;
        else if (kind == "city") k = VLB_SCENE_CITY;
        else if (kind == "cube") k = VLB_SCENE_CUBE;
        else if (kind == "spheres") k = VLB_SCENE_SPHERES;
        else throw std::runtime_error("Select scene to bake: a .gltf/.glb file or room|terrain|city|cube|spheres[:seed] (got '" + name + "')");
        vlb_host_scene* hs = nullptr;
        check(vlb_scene_generate(k, std::stoull(seed_s), 0, 0, &hs), "vlb_scene_generate");
        vlb_status st = vlb_scene_upload(ctx_, hs->xyz, hs->nverts, hs->idx, hs->ntris, hs->normals,
                                         hs->albedo, &scene_);
        vlb_host_scene_free(hs);
        check(st, "vlb_scene_upload", ctx_);
        gltfText_ = "{\"asset\": {\"generator\": \"vlb-mi355x synthetic scene\", \"version\": \"2.0\"}}";
    }

    vlb_bake_params params()
    {
        vlb_bake_params p{};
        p.num_probes = (uint32_t)probePositions_.size();
        p.probe_xyz = reinterpret_cast<const float*>(probePositions_.data());
        p.dir_mode = settings_.dirMode;
        p.rays_per_probe = settings_.rays;
        p.equirect_w = settings_.equirectW;
        p.equirect_h = settings_.equirectH;
        p.t_min = settings_.tMin;
        p.t_max = settings_.tMax;
        p.sh_order = settings_.shOrder;
        p.sh_constants = settings_.shConstants;
        p.precision = settings_.precision;
        p.shade_mode = settings_.shadeMode;
        p.num_lights = (uint32_t)settings_.lights.size();
        p.lights = settings_.lights.data();
        p.shadow_bias = settings_.shadowBias;
        p.sky = settings_.sky;
        p.ref = settings_.ref;
        return p;
    }

    std::string name_;
    Settings settings_;
    vlb_ctx* ctx_ = nullptr;
    uint32_t scene_ = 0;
    std::vector<vec3> probePositions_;
    vec3 gridStep_;
    std::vector<uint8_t> coeffs_;
    uint32_t numCoefs_ = 16;
    std::string gltfText_;
};

}  // namespace vlb
