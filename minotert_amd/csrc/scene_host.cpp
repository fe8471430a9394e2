// scene_host.cpp — synthetic icosphere-soup scenes and the camera basis (host side).
//
// Scene recipe (no reference counterpart; the reference scene is the compile-time sphere
// list of src/gpu/scene.glsl:4-11):
//   * unit icosahedron from the golden-ratio construction, subdivided `level` times into
//     4 triangles per triangle via edge midpoints pushed back onto the unit sphere, in double;
//   * per sphere, from one PCG32 stream (O'Neill 2014, XSH-RR 64/32, sequence 54, seeded with
//     the scene seed): centre x, y, z = -box + 2*box*U, radius = r_min + (r_max-r_min)*U,
//     albedo r, g, b = alb_min + (alb_max-alb_min)*U, U = (next >> 8) / 2^24;
//   * vertex = float(centre + radius * unit_vertex) in double, so shared edges of adjacent
//     triangles get bit-identical endpoints (watertight soups);
//   * optional 2-triangle ground plane z = ground_z spanning [-half, half]^2, last material.
// Camera basis: src/gfx/camera.ixx:26-44 and src/stx/math.ixx:822-860 in closed form.
#include "internal.h"
#include <cmath>
#include <cstring>

namespace {

struct Pcg32Stream {
    uint64_t s = 0, inc = 0;
    uint32_t step() {
        uint64_t prev = s;
        s = prev * 6364136223846793005ULL + inc;
        uint32_t x = (uint32_t)(((prev >> 18u) ^ prev) >> 27u);
        uint32_t r = (uint32_t)(prev >> 59u);
        return (x >> r) | (x << ((32u - r) & 31u));
    }
    Pcg32Stream(uint64_t seed, uint64_t seq) {
        inc = (seq << 1u) | 1u;
        step();
        s += seed;
        step();
    }
    double u01() { return (double)((float)(step() >> 8) * (1.0f / 16777216.0f)); }
};

struct P3 { double x, y, z; };
P3 unit(P3 p) {
    double n = std::sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
    return P3{p.x / n, p.y / n, p.z / n};
}
P3 midpoint_on_sphere(P3 a, P3 b) { return unit(P3{a.x + b.x, a.y + b.y, a.z + b.z}); }

void unit_icosphere(uint32_t level, std::vector<P3>& tri) {
    const double g = (1.0 + std::sqrt(5.0)) * 0.5;
    const P3 base[12] = {{-1, g, 0}, {1, g, 0}, {-1, -g, 0}, {1, -g, 0}, {0, -1, g}, {0, 1, g},
                         {0, -1, -g}, {0, 1, -g}, {g, 0, -1}, {g, 0, 1}, {-g, 0, -1}, {-g, 0, 1}};
    static const uint8_t faces[60] = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11,
                                      1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
                                      3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9,
                                      4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    tri.clear();
    for (int i = 0; i < 60; ++i) tri.push_back(unit(base[faces[i]]));
    std::vector<P3> next;
    for (uint32_t l = 0; l < level; ++l) {
        next.resize(tri.size() * 4);
        size_t w = 0;
        for (size_t i = 0; i < tri.size(); i += 3) {
            P3 a = tri[i], b = tri[i + 1], c = tri[i + 2];
            P3 ab = midpoint_on_sphere(a, b), bc = midpoint_on_sphere(b, c), ca = midpoint_on_sphere(c, a);
            // corner triangles a, b, c, then the centre triangle
            next[w++] = a;  next[w++] = ab; next[w++] = ca;
            next[w++] = b;  next[w++] = bc; next[w++] = ab;
            next[w++] = c;  next[w++] = ca; next[w++] = bc;
            next[w++] = ab; next[w++] = bc; next[w++] = ca;
        }
        tri.swap(next);
    }
}

}  // namespace

extern "C" uint32_t mrt_scene_desc_tri_count(const mrt_scene_desc* d) {
    if (!d) return 0;
    uint64_t per = 20;
    for (uint32_t l = 0; l < d->level; ++l) per *= 4;
    uint64_t n = per * d->n_spheres + (d->ground ? 2 : 0);
    return n > 0xFFFFFFFFull ? 0 : (uint32_t)n;
}

extern "C" mrt_status mrt_scene_generate(const mrt_scene_desc* d, float* verts, uint32_t* tri_material,
                                         float* material_albedo) {
    if (!d || !verts || !tri_material || !material_albedo) return MRT_ERR_INVALID;
    if (d->level > 8) return MRT_ERR_INVALID;
    std::vector<P3> unit_tris;
    unit_icosphere(d->level, unit_tris);
    const size_t per = unit_tris.size() / 3;
    Pcg32Stream rng(d->seed, 54u);
    size_t t = 0;
    for (uint32_t s = 0; s < d->n_spheres; ++s) {
        double box = d->box;
        float cx = (float)(-box + 2.0 * box * rng.u01());
        float cy = (float)(-box + 2.0 * box * rng.u01());
        float cz = (float)(-box + 2.0 * box * rng.u01());
        float r = (float)((double)d->r_min + ((double)d->r_max - (double)d->r_min) * rng.u01());
        for (int k = 0; k < 3; ++k)
            material_albedo[3 * s + k] =
                (float)((double)d->alb_min + ((double)d->alb_max - (double)d->alb_min) * rng.u01());
        for (size_t i = 0; i < per; ++i, ++t) {
            float* out = verts + t * 9;
            for (int v = 0; v < 3; ++v) {
                const P3& p = unit_tris[i * 3 + v];
                out[3 * v + 0] = (float)((double)cx + (double)r * p.x);
                out[3 * v + 1] = (float)((double)cy + (double)r * p.y);
                out[3 * v + 2] = (float)((double)cz + (double)r * p.z);
            }
            tri_material[t] = s;
        }
    }
    if (d->ground) {
        const float z = d->ground_z, h = d->ground_half;
        const float quad[18] = {-h, -h, z, h, -h, z, h, h, z, -h, -h, z, h, h, z, -h, h, z};
        std::memcpy(verts + t * 9, quad, sizeof(quad));
        tri_material[t] = tri_material[t + 1] = d->n_spheres;
        for (int k = 0; k < 3; ++k) material_albedo[3 * d->n_spheres + k] = d->ground_albedo;
    }
    return MRT_OK;
}

rt::CamBasis camera_basis(const mrt_camera& c) {
    // direction (camera.ixx:26-32); look (math.ixx:822-844): s = normalize(cross(up, dir)),
    // u = cross(dir, s); view x axis is -s. perspective (math.ixx:848-860): h = 1/tan(fov/2),
    // w = h * H / W; the inverse maps NDC (X, Y) to view (X/w, Y/h, 1).
    const double cp = std::cos((double)c.pitch), sp = std::sin((double)c.pitch);
    const double cy = std::cos((double)c.yaw), sy = std::sin((double)c.yaw);
    const double dir[3] = {cp * cy, cp * sy, sp};
    double s[3] = {-dir[1], dir[0], 0.0};
    const double sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (double& x : s) x /= sl;
    const double u[3] = {dir[1] * s[2] - dir[2] * s[1], dir[2] * s[0] - dir[0] * s[2],
                         dir[0] * s[1] - dir[1] * s[0]};
    const double h = 1.0 / std::tan(0.5 * (double)c.vertical_fov);
    const double w = h * ((double)c.height / (double)c.width);
    rt::CamBasis b;
    b.pos = rt::make_v3(c.position[0], c.position[1], c.position[2]);
    b.right = rt::make_v3((float)-s[0], (float)-s[1], (float)-s[2]);
    b.up = rt::make_v3((float)u[0], (float)u[1], (float)u[2]);
    b.fwd = rt::make_v3((float)dir[0], (float)dir[1], (float)dir[2]);
    b.sx = (float)(1.0 / w);
    b.sy = (float)(1.0 / h);
    b.pitch_x = 1.0f / (float)c.width;
    b.pitch_y = 1.0f / (float)c.height;
    b.width = c.width;
    b.height = c.height;
    return b;
}

extern "C" mrt_status mrt_camera_basis(const mrt_camera* cam, float out[14]) {
    if (!cam || !out || cam->width == 0 || cam->height == 0) return MRT_ERR_INVALID;
    rt::CamBasis b = camera_basis(*cam);
    const float v[14] = {b.pos.x, b.pos.y, b.pos.z, b.right.x, b.right.y, b.right.z, b.up.x, b.up.y,
                         b.up.z, b.fwd.x, b.fwd.y, b.fwd.z, b.sx, b.sy};
    std::memcpy(out, v, sizeof(v));
    return MRT_OK;
}
