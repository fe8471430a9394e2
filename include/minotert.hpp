// minotert.hpp — C++ render-stage API over the C ABI of minotert.h.
//
// Mirrors the reference's host classes: Camera (src/gfx/camera.ixx:8-65) and the render-stage
// class Pathtracer (src/gfx/modules/pathtracer.ixx:18-200) with its two stages
//     auto primaryRays(uvec2 size, Camera const& camera, Camera const& prevCamera) -> GBuffer;
//     auto secondaryRays(GBuffer, Camera const&, ...) -> image;
// and the post-processing classes Denoiser (denoiser.ixx:16-97) and Tonemapper
// (tonemapper.ixx:15-373).
// Errors are thrown as exceptions (the reference throws runtime_error_fmt,
// src/stx/except.ixx:12-20); the C layer underneath returns status codes. Device memory is
// owned by RAII buffers; every call is ordered on the context's stream.
#pragma once
#include "minotert.h"

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace minote {

struct uvec2 { uint32_t x, y; };
struct vec3 { float x, y, z; };

class Error : public std::runtime_error {
public:
    Error(mrt_status s, std::string const& msg) : std::runtime_error(msg), status(s) {}
    mrt_status status;
};

inline void check(mrt_status s, mrt_ctx const* ctx, char const* what) {
    if (s != MRT_OK) throw Error(s, std::string(what) + ": " + mrt_last_error(ctx));
}

// Camera with the reference's fields (camera.ixx:8-22); the movement fields are not needed.
struct Camera {
    uvec2 viewport{0, 0};
    float verticalFov = 1.0471976f;  // 60 degrees (app.ixx:25)
    float nearPlane = 0.001f;        // app.ixx:26
    vec3 position{0.0f, 0.0f, 0.0f};
    float yaw = 0.0f;
    float pitch = 0.0f;

    // Unit view direction (camera.ixx:26-32).
    vec3 direction() const {
        return vec3{std::cos(pitch) * std::cos(yaw), std::cos(pitch) * std::sin(yaw), std::sin(pitch)};
    }
    mrt_camera abi() const {
        mrt_camera c{};
        c.position[0] = position.x; c.position[1] = position.y; c.position[2] = position.z;
        c.yaw = yaw; c.pitch = pitch; c.vertical_fov = verticalFov; c.near_plane = nearPlane;
        c.width = viewport.x; c.height = viewport.y;
        return c;
    }
};

class Context {
public:
    explicit Context(int device = 0) {
        mrt_opts o{device, 0};
        check(mrt_create(&h_, &o), nullptr, "mrt_create");
    }
    ~Context() { if (h_) mrt_destroy(h_); }
    Context(Context const&) = delete;
    Context& operator=(Context const&) = delete;
    mrt_ctx* get() const { return h_; }
    void sync() const { check(mrt_sync(h_), h_, "mrt_sync"); }

private:
    mrt_ctx* h_ = nullptr;
};

// One device allocation of n elements of T.
template <typename T>
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    DeviceBuffer(Context const& ctx, size_t n) : ctx_(ctx.get()), n_(n) {
        check(mrt_alloc(ctx_, n * sizeof(T), reinterpret_cast<void**>(&p_)), ctx_, "mrt_alloc");
    }
    ~DeviceBuffer() { reset(); }
    DeviceBuffer(DeviceBuffer&& o) noexcept { *this = std::move(o); }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
        if (this != &o) { reset(); ctx_ = o.ctx_; p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
        return *this;
    }
    DeviceBuffer(DeviceBuffer const&) = delete;
    DeviceBuffer& operator=(DeviceBuffer const&) = delete;
    T* data() const { return p_; }
    size_t size() const { return n_; }
    explicit operator bool() const { return p_ != nullptr; }
    void upload(std::vector<T> const& h) {
        if (h.size() != n_) throw Error(MRT_ERR_INVALID, "upload: size mismatch");
        check(mrt_upload(ctx_, p_, h.data(), n_ * sizeof(T)), ctx_, "mrt_upload");
    }
    std::vector<T> download() const {
        std::vector<T> h(n_);
        check(mrt_download(ctx_, h.data(), p_, n_ * sizeof(T)), ctx_, "mrt_download");
        return h;
    }

private:
    void reset() { if (p_) mrt_free(ctx_, p_); p_ = nullptr; }
    mrt_ctx* ctx_ = nullptr;
    T* p_ = nullptr;
    size_t n_ = 0;
};

// Triangle soup + materials on the device with its BVH.
class Scene {
public:
    Scene(Context const& ctx, std::vector<float> const& verts, std::vector<uint32_t> const& material,
          std::vector<float> const& albedo, mrt_bvh_opts const& opts = mrt_bvh_opts{4, 16})
        : ctx_(ctx.get()) {
        check(mrt_scene_create(ctx_, verts.data(), (uint32_t)(verts.size() / 9), material.data(), albedo.data(),
                               (uint32_t)(albedo.size() / 3), &opts, &h_),
              ctx_, "mrt_scene_create");
    }
    static Scene generate(Context const& ctx, mrt_scene_desc const& d) {
        uint32_t n = mrt_scene_desc_tri_count(&d);
        std::vector<float> v((size_t)n * 9), a(((size_t)d.n_spheres + (d.ground ? 1 : 0)) * 3);
        std::vector<uint32_t> m(n);
        check(mrt_scene_generate(&d, v.data(), m.data(), a.data()), nullptr, "mrt_scene_generate");
        return Scene(ctx, v, m, a);
    }
    ~Scene() { if (h_) mrt_scene_destroy(ctx_, h_); }
    Scene(Scene&& o) noexcept : ctx_(o.ctx_), h_(o.h_) { o.h_ = nullptr; }
    Scene(Scene const&) = delete;
    Scene& operator=(Scene const&) = delete;
    mrt_scene* get() const { return h_; }
    mrt_scene_stats stats() const {
        mrt_scene_stats s{};
        check(mrt_scene_stats_get(ctx_, h_, &s), ctx_, "mrt_scene_stats_get");
        return s;
    }

private:
    mrt_ctx* ctx_ = nullptr;
    mrt_scene* h_ = nullptr;
};

// Miss radiance and directional light; stands in for the reference's Atmosphere + sky-view
// LUT inputs of secondaryRays (pathtracer.ixx:118-120, skyColor secondaryRays.comp:36-58).
struct Environment {
    vec3 sky{1.0f, 1.0f, 1.0f};
    vec3 sunDirection{-0.435286462f, 0.818654716f, 0.374606609f};  // secondaryRays.comp:43
    vec3 sunIlluminance{2.0f, 2.0f, 2.0f};
};

// The reference hard-codes 8 samples and 8 bounces (secondaryRays.comp:128-129); here they
// are settings.
struct RenderSettings {
    uint32_t frameSeed = 0;
    uint32_t sampleBase = 0;
    uint32_t samples = 1;
    uint32_t bounces = 4;
    uint32_t rouletteStart = 3;
    bool sortRays = false;
};

class Pathtracer {
public:
    // G-buffer planes (pathtracer.ixx:22-27): hit records (visibility + t + barycentrics),
    // reverse-Z depth, normal and motion in FP32.
    struct GBuffer {
        DeviceBuffer<mrt_hit> visibility;
        DeviceBuffer<float> depth;
        DeviceBuffer<float> normal;
        DeviceBuffer<float> motion;
        uvec2 size{0, 0};
        mrt_gbuffer view() const {
            return mrt_gbuffer{visibility.data(), depth.data(), normal.data(), motion.data(), size.x, size.y};
        }
    };

    Pathtracer(Context const& ctx, Scene const& scene) : ctx_(ctx), scene_(scene) {}

    auto primaryRays(uvec2 size, Camera const& camera, Camera const& prevCamera) -> GBuffer {
        if (size.x != camera.viewport.x || size.y != camera.viewport.y)
            throw Error(MRT_ERR_INVALID, "primaryRays: size does not match the camera viewport");
        size_t n = (size_t)size.x * size.y;
        GBuffer gb{DeviceBuffer<mrt_hit>(ctx_, n), DeviceBuffer<float>(ctx_, n), DeviceBuffer<float>(ctx_, 4 * n),
                   DeviceBuffer<float>(ctx_, 2 * n), size};
        mrt_camera c = camera.abi(), p = prevCamera.abi();
        mrt_gbuffer v = gb.view();
        check(mrt_primary_rays(ctx_.get(), scene_.get(), &c, &p, &v), ctx_.get(), "mrt_primary_rays");
        return gb;
    }

    auto secondaryRays(GBuffer const& gbuffer, Camera const& camera, Environment const& env,
                       RenderSettings const& rs = RenderSettings{}) -> DeviceBuffer<float> {
        DeviceBuffer<float> rgb(ctx_, (size_t)gbuffer.size.x * gbuffer.size.y * 3);
        mrt_render_opts o{};
        o.frame_seed = rs.frameSeed; o.sample_base = rs.sampleBase; o.spp = rs.samples;
        o.bounces = rs.bounces; o.rr_start = rs.rouletteStart; o.flags = rs.sortRays ? MRT_RENDER_SORT_RAYS : 0u;
        o.sky[0] = env.sky.x; o.sky[1] = env.sky.y; o.sky[2] = env.sky.z;
        o.sun_dir[0] = env.sunDirection.x; o.sun_dir[1] = env.sunDirection.y; o.sun_dir[2] = env.sunDirection.z;
        o.sun_e[0] = env.sunIlluminance.x; o.sun_e[1] = env.sunIlluminance.y; o.sun_e[2] = env.sunIlluminance.z;
        mrt_camera c = camera.abi();
        mrt_gbuffer v = gbuffer.view();
        check(mrt_secondary_rays(ctx_.get(), scene_.get(), &c, &v, &o, rgb.data()), ctx_.get(), "mrt_secondary_rays");
        return rgb;
    }

    // Config-0 shading: Lambert with one shadow ray towards the light.
    auto shadeLambert(GBuffer const& gbuffer, Camera const& camera, vec3 lightDir, vec3 lightE) -> DeviceBuffer<float> {
        DeviceBuffer<float> rgb(ctx_, (size_t)gbuffer.size.x * gbuffer.size.y * 3);
        mrt_camera c = camera.abi();
        mrt_gbuffer v = gbuffer.view();
        float l[3] = {lightDir.x, lightDir.y, lightDir.z}, e[3] = {lightE.x, lightE.y, lightE.z};
        check(mrt_shade_lambert(ctx_.get(), scene_.get(), &c, &v, l, e, rgb.data()), ctx_.get(), "mrt_shade_lambert");
        return rgb;
    }

private:
    Context const& ctx_;
    Scene const& scene_;
};

// class Denoiser (src/gfx/modules/denoiser.ixx:16-97): bilateral(color, depth, normal, camera,
// params) filters the path-traced image with its G-buffer.
class Denoiser {
public:
    struct BilateralParams {  // denoiser.ixx:20-31
        float sigma = 5.0f;
        float kSigma = 2.0f;
        float threshold = 0.12f;
    };

    explicit Denoiser(Context const& ctx) : ctx_(ctx) {}

    auto bilateral(DeviceBuffer<float> const& color, Pathtracer::GBuffer const& gbuffer, Camera const& camera)
        -> DeviceBuffer<float> {
        return bilateral(color, gbuffer, camera, BilateralParams{}, 0);
    }
    // kmPerUnit: kilometres per scene unit (the reference's world unit is the km, which its depth
    // term assumes; 1e-3 for a scene modelled in metres).
    auto bilateral(DeviceBuffer<float> const& color, Pathtracer::GBuffer const& gbuffer, Camera const& camera,
                   BilateralParams const& params, uint32_t frame, float kmPerUnit = 1.0f) -> DeviceBuffer<float> {
        DeviceBuffer<float> out(ctx_, (size_t)gbuffer.size.x * gbuffer.size.y * 3);
        mrt_bilateral_params p{params.sigma, params.kSigma, params.threshold, camera.nearPlane * kmPerUnit, frame};
        check(mrt_denoise_bilateral(ctx_.get(), color.data(), gbuffer.depth.data(), gbuffer.normal.data(), gbuffer.size.x,
                                    gbuffer.size.y, &p, out.data()),
              ctx_.get(), "mrt_denoise_bilateral");
        return out;
    }

private:
    Context const& ctx_;
};

// class Tonemapper (src/gfx/modules/tonemapper.ixx:15-373): one method per operator, each
// applying exposure, the curve and the sRGB encode; the output is RGBA8 UNORM (W*H*4 bytes).
class Tonemapper {
public:
    struct UchimuraParams {  // tonemapper.ixx:19-36
        float maxBrightness = 1.0f, contrast = 1.0f, linearStart = 0.22f, linearLength = 0.4f,
              blackTightness = 1.33f, pedestal = 0.0f;
    };
    struct AMDParams {  // tonemapper.ixx:38-55
        float hdrMax = 16.0f, contrast = 2.0f, shoulder = 1.0f, midIn = 0.18f, midOut = 0.18f;
    };

    explicit Tonemapper(Context const& ctx) : ctx_(ctx) {}

    auto linear(DeviceBuffer<float> const& in, float exposure) { return run(in, base(MRT_TONEMAP_LINEAR, exposure)); }
    auto reinhard(DeviceBuffer<float> const& in, float exposure, float hdrMax) {
        mrt_tonemap_params p = base(MRT_TONEMAP_REINHARD, exposure);
        p.reinhard_max = hdrMax;
        return run(in, p);
    }
    auto hable(DeviceBuffer<float> const& in, float exposure) { return run(in, base(MRT_TONEMAP_HABLE, exposure)); }
    auto aces(DeviceBuffer<float> const& in, float exposure) { return run(in, base(MRT_TONEMAP_ACES, exposure)); }
    auto uchimura(DeviceBuffer<float> const& in, float exposure) { return uchimura(in, exposure, UchimuraParams{}); }
    auto uchimura(DeviceBuffer<float> const& in, float exposure, UchimuraParams const& u) -> DeviceBuffer<uint8_t> {
        mrt_tonemap_params p = base(MRT_TONEMAP_UCHIMURA, exposure);
        const float v[6] = {u.maxBrightness, u.contrast, u.linearStart, u.linearLength, u.blackTightness, u.pedestal};
        for (int k = 0; k < 6; ++k) p.uchimura[k] = v[k];
        return run(in, p);
    }
    auto amd(DeviceBuffer<float> const& in, float exposure) { return amd(in, exposure, AMDParams{}); }
    auto amd(DeviceBuffer<float> const& in, float exposure, AMDParams const& a) -> DeviceBuffer<uint8_t> {
        mrt_tonemap_params p = base(MRT_TONEMAP_AMD, exposure);
        const float v[5] = {a.hdrMax, a.contrast, a.shoulder, a.midIn, a.midOut};
        for (int k = 0; k < 5; ++k) p.amd[k] = v[k];
        return run(in, p);
    }

private:
    static mrt_tonemap_params base(uint32_t op, float exposure) {
        mrt_tonemap_params p;
        mrt_tonemap_defaults(&p);
        p.op = op;
        p.exposure = exposure;
        return p;
    }
    DeviceBuffer<uint8_t> run(DeviceBuffer<float> const& in, mrt_tonemap_params const& p) {
        const uint32_t n = (uint32_t)(in.size() / 3);
        DeviceBuffer<uint8_t> out(ctx_, (size_t)n * 4);
        check(mrt_tonemap(ctx_.get(), in.data(), n, &p, nullptr, out.data()), ctx_.get(), "mrt_tonemap");
        return out;
    }
    Context const& ctx_;
};

}  // namespace minote
