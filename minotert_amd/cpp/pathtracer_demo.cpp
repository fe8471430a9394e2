// pathtracer_demo.cpp — the C++ render-stage API (include/minotert.hpp) end to end:
// generates a scene, runs Pathtracer::primaryRays and then Lambert shading or path tracing,
// and writes the RGB image as raw little-endian float32 (W*H*3) to the given file; with a
// tenth argument it also runs Denoiser::bilateral and the AMD Tonemapper on the image and
// writes the RGBA8 result (W*H*4 bytes) there.
// Usage: pathtracer_demo <out.f32> <spheres> <level> <box> <width> <height> <mode: lambert|path>
//        [frame_seed] [bounces] [out.rgba8]
#include "minotert.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

int main(int argc, char** argv) {
    if (argc < 8) {
        std::fprintf(stderr, "usage: %s out.f32 spheres level box width height lambert|path [seed] [bounces]\n", argv[0]);
        return 2;
    }
    try {
        minote::Context ctx(0);
        mrt_scene_desc d{};
        d.seed = 0x5EED0000u;
        d.n_spheres = (uint32_t)std::atoi(argv[2]);
        d.level = (uint32_t)std::atoi(argv[3]);
        d.box = (float)std::atof(argv[4]);
        d.r_min = 0.5f; d.r_max = 2.0f; d.alb_min = 0.2f; d.alb_max = 0.9f;
        d.ground = 1; d.ground_z = -d.box - 2.0f; d.ground_half = 4.0f * d.box; d.ground_albedo = 0.5f;
        minote::Scene scene = minote::Scene::generate(ctx, d);
        minote::Camera cam;
        cam.viewport = {(uint32_t)std::atoi(argv[5]), (uint32_t)std::atoi(argv[6])};
        cam.position = {0.0f, -2.2f * d.box, 0.35f * d.box};
        cam.yaw = 1.5707963267948966f;
        cam.pitch = (float)-std::atan2(0.35, 2.2);
        minote::Pathtracer pt(ctx, scene);
        auto gb = pt.primaryRays(cam.viewport, cam, cam);
        minote::Environment env;
        minote::DeviceBuffer<float> rgb;
        if (std::strcmp(argv[7], "lambert") == 0) {
            rgb = pt.shadeLambert(gb, cam, env.sunDirection, minote::vec3{1.0f, 1.0f, 1.0f});
        } else {
            minote::RenderSettings rs;
            rs.frameSeed = argc > 8 ? (uint32_t)std::strtoul(argv[8], nullptr, 0) : 0u;
            rs.bounces = argc > 9 ? (uint32_t)std::atoi(argv[9]) : 4u;
            rgb = pt.secondaryRays(gb, cam, env, rs);
        }
        std::vector<float> img = rgb.download();
        FILE* f = std::fopen(argv[1], "wb");
        if (!f) { std::perror("fopen"); return 1; }
        std::fwrite(img.data(), sizeof(float), img.size(), f);
        std::fclose(f);
        if (argc > 10) {
            minote::Denoiser denoiser(ctx);
            minote::Tonemapper tonemapper(ctx);
            auto denoised = denoiser.bilateral(rgb, gb, cam);
            std::vector<uint8_t> q = tonemapper.amd(denoised, 1.0f).download();
            FILE* g = std::fopen(argv[10], "wb");
            if (!g) { std::perror("fopen"); return 1; }
            std::fwrite(q.data(), 1, q.size(), g);
            std::fclose(g);
        }
        double sum = 0.0;
        for (float x : img) sum += x;
        mrt_scene_stats st = scene.stats();
        std::printf("{\"pixels\": %u, \"mean\": %.9g, \"tris\": %u, \"bvh_nodes\": %u}\n", cam.viewport.x * cam.viewport.y,
                    sum / (double)img.size(), st.n_tris, st.n_nodes);
        return 0;
    } catch (minote::Error const& e) {
        std::fprintf(stderr, "minote error %d: %s\n", (int)e.status, e.what());
        return 1;
    }
}
