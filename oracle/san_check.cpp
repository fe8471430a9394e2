// san_check.cpp — host-side checks run under AddressSanitizer and UndefinedBehaviorSanitizer
// (`make -C oracle SAN=1 check`). TEST INFRASTRUCTURE ONLY: it exercises every entry point of the
// CPU oracle and the product's host-only translation units (scene generator and camera basis of
// minotert_amd/csrc/scene_host.cpp, host BVH4 builder of minotert_amd/csrc/bvh_host.cpp) on small
// inputs and checks their results, so that a crash, an out-of-bounds access or undefined
// behaviour anywhere on these paths fails the run. Exit status 0 when every check passes.
#include "oracle.h"

#include "../minotert_amd/csrc/internal.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

namespace {

int g_fail = 0;
#define CHECK(cond, ...)                                                   \
    do {                                                                   \
        if (!(cond)) {                                                     \
            std::fprintf(stderr, "san_check FAIL %s:%d: ", __FILE__, __LINE__); \
            std::fprintf(stderr, __VA_ARGS__);                             \
            std::fprintf(stderr, "\n");                                    \
            ++g_fail;                                                      \
        }                                                                  \
    } while (0)

struct Soup {
    std::vector<float> verts;
    std::vector<uint32_t> mat;
    std::vector<float> alb;
    uint32_t n = 0;
};

mrt_scene_desc desc(uint32_t seed, uint32_t spheres, uint32_t level, uint32_t ground, float box) {
    mrt_scene_desc d{};
    d.seed = seed; d.n_spheres = spheres; d.level = level; d.ground = ground; d.box = box;
    d.r_min = 0.5f; d.r_max = 2.0f; d.alb_min = 0.2f; d.alb_max = 0.9f;
    d.ground_z = ground ? -box - 2.0f : 0.0f; d.ground_half = ground ? 4.0f * box : 0.0f; d.ground_albedo = 0.5f;
    return d;
}

// Scene generator of the product against the oracle's, byte for byte.
Soup scene(const mrt_scene_desc& d) {
    orc_scene_desc od;
    static_assert(sizeof(od) == sizeof(d), "same descriptor layout");
    std::memcpy(&od, &d, sizeof(d));
    Soup s;
    s.n = mrt_scene_desc_tri_count(&d);
    CHECK(s.n == orc_scene_tri_count(&od), "triangle counts differ");
    const uint32_t nm = d.n_spheres + (d.ground ? 1u : 0u);
    s.verts.resize((size_t)s.n * 9); s.mat.resize(s.n); s.alb.resize((size_t)nm * 3);
    std::vector<float> v2(s.verts.size()), a2(s.alb.size());
    std::vector<uint32_t> m2(s.n);
    CHECK(mrt_scene_generate(&d, s.verts.data(), s.mat.data(), s.alb.data()) == MRT_OK, "generate");
    CHECK(orc_scene_generate(&od, v2.data(), m2.data(), a2.data()) == (int)s.n, "oracle generate (triangles written)");
    CHECK(std::memcmp(v2.data(), s.verts.data(), v2.size() * 4) == 0, "vertices differ from the oracle's");
    CHECK(std::memcmp(m2.data(), s.mat.data(), m2.size() * 4) == 0, "materials differ from the oracle's");
    CHECK(std::memcmp(a2.data(), s.alb.data(), a2.size() * 4) == 0, "albedo differs from the oracle's");
    return s;
}

// Structural check of a host BVH4: a tree from node 0, every leaf record covered once, record ids
// a permutation whose vertices equal the soup's, leaves of 1..max_leaf, child boxes inside their
// slot boxes, leaf triangles inside theirs.
void check_bvh4(const HostBVH& h, const float* verts, uint32_t n, uint32_t max_leaf) {
    CHECK(!h.nodes.empty() && h.nodes.size() <= std::max<size_t>(n, 1), "node count %zu for %u triangles", h.nodes.size(), n);
    CHECK(h.tri4.size() == (size_t)n * 12, "leaf record count");
    std::vector<int> seen(h.nodes.size(), 0), cover(n, 0), ids(n, 0);
    struct Item { uint32_t node; float lo[3], hi[3]; };
    const float inf = INFINITY;
    std::vector<Item> stack{{0u, {-inf, -inf, -inf}, {inf, inf, inf}}};
    while (!stack.empty()) {
        const Item it = stack.back();
        stack.pop_back();
        CHECK(it.node < h.nodes.size(), "child index past the node array");
        if (it.node >= h.nodes.size()) continue;
        CHECK(++seen[it.node] == 1, "node %u reached twice", it.node);
        const WNode& w = h.nodes[it.node];
        for (int c = 0; c < 4; ++c) {
            const uint32_t cw = w.child[c];
            if (cw == kEmptyChild) continue;
            const float lo[3] = {w.lox[c], w.loy[c], w.loz[c]}, hi[3] = {w.hix[c], w.hiy[c], w.hiz[c]};
            for (int k = 0; k < 3; ++k)
                CHECK(lo[k] >= it.lo[k] && hi[k] <= it.hi[k], "child box outside its parent slot (node %u)", it.node);
            if (cw & kLeafBit) {
                const uint32_t first = (cw >> 3) & 0x0FFFFFFFu, count = (cw & 7u) + 1u;
                CHECK(count <= max_leaf, "leaf of %u triangles", count);
                CHECK(first + count <= n, "leaf range past the records");
                for (uint32_t r = first; r < first + count && r < n; ++r) {
                    ++cover[r];
                    const float* rec = &h.tri4[(size_t)r * 12];
                    uint32_t id;
                    std::memcpy(&id, &rec[3], 4);
                    CHECK(id < n, "record id");
                    if (id >= n) continue;
                    ++ids[id];
                    for (int v = 0; v < 3; ++v)
                        for (int k = 0; k < 3; ++k) {
                            const float x = rec[4 * v + k];
                            CHECK(std::memcmp(&x, &verts[(size_t)id * 9 + 3 * v + k], 4) == 0, "record vertex");
                            CHECK(x >= lo[k] && x <= hi[k], "leaf triangle outside its slot box");
                        }
                }
            } else {
                Item ch{cw, {lo[0], lo[1], lo[2]}, {hi[0], hi[1], hi[2]}};
                stack.push_back(ch);
            }
        }
    }
    for (size_t i = 0; i < seen.size(); ++i) CHECK(seen[i] == 1, "node %zu unreachable", i);
    for (uint32_t i = 0; i < n; ++i) CHECK(cover[i] == 1 && ids[i] == 1, "triangle %u not in exactly one leaf", i);
}

void bvh_cases(const Soup& s) {
    const uint32_t leafs[] = {1, 4, 8}, bins[] = {4, 16, 64};
    for (uint32_t L : leafs)
        for (uint32_t B : bins) {
            HostBVH h;
            build_bvh4(s.verts.data(), s.n, L, B, h);
            check_bvh4(h, s.verts.data(), s.n, L);
        }
}

std::vector<float> random_rays(uint32_t seed, float box, uint32_t n) {
    std::vector<float> r((size_t)n * 8);
    orc_random_rays(seed, box, n, r.data());
    return r;
}

void oracle_paths(const Soup& s) {
    orc_bvh* b = orc_bvh_build(s.verts.data(), s.n);
    CHECK(orc_bvh_validate(b) == 0, "oracle BVH invalid");
    uint32_t info[4];
    double sah = 0;
    orc_bvh_info(b, info, &sah);
    CHECK(info[3] <= 4 && sah > 0, "oracle BVH info");
    const uint32_t nr = 3000;
    std::vector<float> rays = random_rays(7, 12.0f, nr);
    std::vector<uint32_t> hb((size_t)nr * 4), hv((size_t)nr * 4);
    orc_closest_brute(s.verts.data(), s.n, rays.data(), nr, hb.data(), 0);
    uint64_t nb = 0, nt = 0;
    orc_trace_closest(b, rays.data(), nr, hv.data(), &nb, &nt, 0);
    int hits = 0;
    for (uint32_t i = 0; i < nr; ++i) {
        CHECK(hb[4 * i + 1] == hv[4 * i + 1], "BVH trace differs from brute force at ray %u", i);
        hits += hv[4 * i + 1] != 0xFFFFFFFFu;
    }
    CHECK(hits > 100 && nb > 0 && nt > 0, "too few hits");
    std::vector<uint8_t> occ(nr);
    orc_trace_any(b, rays.data(), nr, occ.data(), 2);
    for (uint32_t i = 0; i < nr; ++i) CHECK((occ[i] != 0) == (hv[4 * i + 1] != 0xFFFFFFFFu), "any-hit differs");
    float hit[4], tri[9];
    std::memcpy(tri, s.verts.data(), sizeof(tri));
    orc_intersect(rays.data(), tri, 0, hit);

    orc_camera cam{};
    cam.pos[0] = 0.0f; cam.pos[1] = -22.0f; cam.pos[2] = 3.5f;
    cam.yaw = 1.5707964f; cam.pitch = -0.157f; cam.vfov = 1.0471976f; cam.width = 40; cam.height = 24;
    float basis[14], basis2[14];
    orc_camera_basis(&cam, basis);
    mrt_camera mc{};
    std::memcpy(mc.position, cam.pos, 12);
    mc.yaw = cam.yaw; mc.pitch = cam.pitch; mc.vertical_fov = cam.vfov; mc.near_plane = 0.001f;
    mc.width = cam.width; mc.height = cam.height;
    CHECK(mrt_camera_basis(&mc, basis2) == MRT_OK && std::memcmp(basis, basis2, sizeof(basis)) == 0,
          "camera basis differs from the oracle's");
    const uint32_t np = cam.width * cam.height;
    std::vector<float> prim((size_t)np * 8), rgb((size_t)np * 3), rgb2((size_t)np * 3);
    orc_primary_rays(&cam, prim.data());
    std::vector<uint32_t> hits4((size_t)np * 4);
    const float ldir[3] = {-0.435286462f, 0.818654716f, 0.374606609f}, le[3] = {1, 1, 1};
    orc_render_lambert(b, s.mat.data(), s.alb.data(), &cam, ldir, le, hits4.data(), rgb.data(), 0);
    orc_path_params p{};
    p.frame_seed = 0x9E3779B9u; p.bounces = 4; p.rr_start = 3;
    for (int k = 0; k < 3; ++k) { p.sky[k] = 1.0f; p.sun_dir[k] = ldir[k]; p.sun_e[k] = 2.0f; }
    orc_render_path(b, s.mat.data(), s.alb.data(), &cam, &p, nullptr, rgb.data(), 0);
    for (float x : rgb) CHECK(std::isfinite(x) && x >= 0.0f, "path radiance not finite");
    std::vector<float> br((size_t)np * 8);
    std::vector<uint32_t> bp(np);
    const uint32_t nb1 = orc_bounce1_rays(b, &cam, &p, br.data(), bp.data(), 0);
    CHECK(nb1 > 0 && nb1 <= np, "bounce-1 rays");
    std::vector<uint32_t> pids;
    for (uint32_t i = 0; i < np; i += 7) pids.push_back(i);
    p.cache_vertex = 1;
    std::vector<float> smp(pids.size() * 16), tgt(pids.size() * 4);
    std::vector<uint8_t> valid(pids.size());
    orc_path_suffix(b, s.mat.data(), s.alb.data(), &cam, &p, pids.data(), (uint32_t)pids.size(), smp.data(), tgt.data(),
                    valid.data(), 0);
    std::vector<float> cache_rgb((size_t)np * 3, 0.25f);
    p.cache_mode = 1;
    orc_render_path(b, s.mat.data(), s.alb.data(), &cam, &p, cache_rgb.data(), rgb2.data(), 0);
    CHECK(orc_rel_mse(rgb2.data(), rgb.data(), rgb.size(), 1e-2) >= 0.0, "rel_mse");
    orc_bvh_free(b);
}

void oracle_cache() {
    orc_grid g{};
    g.levels = 6; g.log2_table = 12; g.base_res = 4.0f; g.max_res = 64.0f;
    for (int k = 0; k < 3; ++k) { g.lo[k] = -2.0f; g.hi[k] = 3.0f; }
    const size_t tf = (size_t)g.levels * (1u << g.log2_table) * 2;
    std::mt19937 rng(3);
    std::uniform_real_distribution<float> u(-0.5f, 0.5f);
    std::vector<float> table(tf), W(6 * 64 * 64);
    for (float& x : table) x = u(rng);
    for (float& x : W) x = 0.3f * u(rng);
    const uint32_t n = 257;
    std::vector<float> pos(n * 3), dir(n * 3), nrm(n * 3), alb(n * 3), x0((size_t)n * 64), y(n * 4), acts((size_t)5 * n * 64);
    for (uint32_t i = 0; i < n * 3; ++i) { pos[i] = 6.0f * u(rng) + 0.5f; dir[i] = u(rng); nrm[i] = u(rng); alb[i] = 0.5f + u(rng); }
    orc_cache_encode(&g, table.data(), pos.data(), dir.data(), nrm.data(), alb.data(), n, 1.0f, x0.data());
    orc_mlp_forward(W.data(), x0.data(), n, y.data(), acts.data());
    std::vector<float> target(n * 4, 0.3f), dW(W.size()), dx0((size_t)n * 64), gt(tf, 0.0f);
    const double loss = orc_mlp_backward(W.data(), x0.data(), n, target.data(), dW.data(), dx0.data());
    CHECK(std::isfinite(loss) && loss >= 0.0, "loss");
    orc_cache_encode_backward(&g, pos.data(), dx0.data(), n, gt.data());
    std::vector<float> m(tf, 0.0f), v(tf, 0.0f), mw(W.size(), 0.0f), vw(W.size(), 0.0f);
    orc_adam(table.data(), gt.data(), m.data(), v.data(), (uint32_t)tf, 1, 1e-2f, 0.9f, 0.99f, 1e-8f, 1);
    orc_adam(W.data(), dW.data(), mw.data(), vw.data(), (uint32_t)W.size(), 1, 1e-2f, 0.9f, 0.99f, 1e-8f, 0);
    for (float x : table) CHECK(std::isfinite(x), "table not finite after Adam");
    CHECK(orc_bf16_round(1.0f) == 1.0f, "bf16");
}

void oracle_sky_post() {
    orc_atmosphere a;
    orc_sky_earth(&a);
    const uint32_t tw = 32, th = 8, mw = 8, mh = 8, vw = 24, vh = 12;
    std::vector<float> trans(tw * th * 4), ms(mw * mh * 4), view(vw * vh * 4);
    orc_sky_transmittance(&a, tw, th, trans.data());
    orc_sky_multiscattering(&a, trans.data(), tw, th, mw, mh, ms.data());
    const float probe[3] = {0.0f, 0.0f, a.bottom_radius + 0.2f}, sun[3] = {-0.435286462f, 0.818654716f, 0.374606609f};
    const float illum[3] = {8, 8, 8};
    orc_sky_view(&a, trans.data(), tw, th, ms.data(), mw, mh, probe, sun, illum, vw, vh, view.data());
    const uint32_t nd = 64;
    std::vector<float> dirs(nd * 3), out(nd * 3);
    for (uint32_t i = 0; i < nd; ++i) {
        const float t = 0.1f * (float)i;
        dirs[3 * i] = std::cos(t) * 0.6f; dirs[3 * i + 1] = std::sin(t) * 0.6f; dirs[3 * i + 2] = 0.8f * (i % 2 ? 1.0f : -1.0f);
    }
    orc_sky_radiance(&a, trans.data(), tw, th, view.data(), vw, vh, probe, sun, illum, 1, dirs.data(), nd, out.data());
    for (float x : out) CHECK(std::isfinite(x) && x >= 0.0f, "sky radiance");
    float uv[2], pr[2];
    orc_sky_trans_uv(a.bottom_radius + 1.0f, 0.3f, a.bottom_radius, a.top_radius, uv);
    orc_sky_trans_params(uv[0], uv[1], a.bottom_radius, a.top_radius, pr);
    orc_sky_view_uv(0, 0.2f, 0.1f, (int)vw, (int)vh, probe[2], a.bottom_radius, uv);
    orc_sky_view_params(uv[0], uv[1], (int)vw, (int)vh, probe[2], a.bottom_radius, pr);
    CHECK(orc_sky_phase(1, 0.8f, 0.5f) > 0.0f, "phase");
    float xs[5] = {0.1f, 0.5f, 1.0f, 2.0f, 3.0f}, ys[5];
    for (int fn = 0; fn < 5; ++fn) orc_tx_eval(fn, xs, 5, ys);

    const uint32_t w = 32, h = 24, np = w * h;
    std::vector<float> rgb(np * 3), depth(np), normal(np * 4), den(np * 3), srgb(np * 3);
    std::vector<uint8_t> rgba(np * 4);
    for (uint32_t i = 0; i < np; ++i) {
        for (int k = 0; k < 3; ++k) rgb[3 * i + k] = 0.1f * (float)((i * 7 + k) % 13);
        depth[i] = (i % 9 == 0) ? 0.0f : 0.001f / (1.0f + 0.01f * (float)(i % w));
        normal[4 * i] = 0.0f; normal[4 * i + 1] = 0.0f; normal[4 * i + 2] = 1.0f; normal[4 * i + 3] = 0.0f;
    }
    orc_bilateral bp{5.0f, 2.0f, 0.12f, 0.001f, 3u};
    orc_denoise_bilateral(rgb.data(), depth.data(), normal.data(), w, h, &bp, den.data());
    for (uint32_t op = 0; op < 6; ++op) {
        orc_tonemap_params tp;
        orc_tonemap_defaults(&tp);
        tp.op = op;
        orc_tonemap(&tp, den.data(), np, srgb.data(), rgba.data());
        for (float x : srgb) CHECK(std::isfinite(x), "tonemap output not finite (op %u)", op);
    }
    float s1, c1, nn[3] = {0, 0, 1}, cd[3], off[3], pp[3] = {1, 2, 3};
    orc_sincos_turns(0.3f, &s1, &c1);
    orc_cosine_dir(nn, 0.25f, 0.75f, cd);
    orc_offset_ray(pp, nn, off);
    uint32_t st = 5;
    CHECK(orc_rand_float(&st) < 1.0f && orc_pcg(0) != 0u, "rng");
}

}  // namespace

int main() {
    const Soup s0 = scene(desc(0x5EED0000u, 64, 2, 1, 10.0f));
    const Soup s1 = scene(desc(0x5EED0001u, 40, 3, 0, 40.0f));
    const Soup s4 = scene(desc(0x5EED0004u, 3, 4, 0, 100.0f));
    std::printf("san_check: scenes %u %u %u triangles\n", s0.n, s1.n, s4.n);
    bvh_cases(s0);
    bvh_cases(s1);
    bvh_cases(s4);
    // degenerate soups: none, one, two, five, duplicates, coplanar, signed zeros, collinear
    std::mt19937 rng(7);
    std::uniform_real_distribution<float> u(-1.0f, 1.0f);
    const uint32_t sizes[] = {0, 1, 2, 5, 37, 500, 500, 64};
    for (int c = 0; c < 8; ++c) {
        Soup d;
        d.n = sizes[c];
        d.verts.resize((size_t)d.n * 9);
        for (float& x : d.verts) x = u(rng);
        if (c == 4) for (uint32_t t = 1; t < d.n; ++t) std::memcpy(&d.verts[9 * t], &d.verts[0], 36);
        if (c == 5 || c == 6) for (uint32_t t = 0; t < d.n; ++t) for (int v = 0; v < 3; ++v)
            d.verts[9 * t + 3 * v + 2] = (c == 6 && t % 2) ? -0.0f : 0.0f;
        if (c == 7) for (uint32_t t = 0; t < d.n; ++t) for (int v = 0; v < 3; ++v) {
            d.verts[9 * t + 3 * v] = (float)t; d.verts[9 * t + 3 * v + 1] = 0.0f; d.verts[9 * t + 3 * v + 2] = 0.0f;
        }
        bvh_cases(d);
    }
    std::printf("san_check: host BVH4 builds checked\n");
    oracle_paths(s0);
    oracle_cache();
    oracle_sky_post();
    std::printf("san_check: %s (%d failures)\n", g_fail ? "FAILED" : "ok", g_fail);
    return g_fail ? 1 : 0;
}
