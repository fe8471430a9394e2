// oracle_core.hpp — scalar CPU restatement of the MinoteRT hot-path semantics.
//
// TEST INFRASTRUCTURE ONLY. This file is part of the oracle: only the test suite, the smoke
// check and the bench's CPU-baseline leg may load it, and only as the checker. The product
// (minotert_amd/) never includes, links or calls anything under oracle/.
//
// Every function names the reference file:line whose behaviour it restates (paths relative
// to the MinoteRT repository root), or the published method it follows where the reference
// has no counterpart. All arithmetic is written with a fixed operation order and explicit
// fmaf where a fused operation is intended; the file is compiled with -ffp-contract=off so
// that the results are bit-identical to any other IEEE-754 implementation that performs the
// same sequence of correctly rounded operations.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>

namespace orc {

struct f3 { float x, y, z; };
static inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
static inline f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline f3 mul(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
static inline f3 cmul(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
static inline float get(f3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
// dot with a fixed order: fma(x, x', fma(y, y', z*z'))
static inline float dot(f3 a, f3 b) { return std::fmaf(a.x, b.x, std::fmaf(a.y, b.y, a.z * b.z)); }
// cross product with plain (unfused) products
static inline f3 cross(f3 a, f3 b) {
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// normalize: v * (1 / sqrt(dot(v, v))), correctly rounded sqrt and divide
static inline f3 normalize(f3 v) {
    float inv = 1.0f / std::sqrt(dot(v, v));
    return mul(v, inv);
}

// ---------------------------------------------------------------------------------------
// RNG. Restates random.glsl:22-27 (pcg) and random.glsl:29-31 (randomFloat).
// PCG-RXS-M-XS 32-bit hash (Jarzynski & Olano, JCGT 2020).
static inline uint32_t pcg(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
// randomFloat: advance the state in place and map the low 24 bits to [0, 1).
static inline float rand_float(uint32_t& state) {
    state = pcg(state);
    return (float)(state & 0xFFFFFFu) * (1.0f / 16777216.0f);
}

// ---------------------------------------------------------------------------------------
// sin/cos of 2*pi*x for x = k / 2^24 in [0, 1). Published method: exact Cody-Waite style
// reduction to the nearest quarter turn (exact in binary floating point because x has at
// most 24 significant bits), then Taylor polynomials of sin and cos on [-pi/4, pi/4]
// (coefficients 1/n!, truncation error < 2e-9), evaluated by Horner with fmaf.
// Replaces the driver-defined sin/cos of random.glsl:15-16.
static inline void sincos_turns(float x, float& s, float& c) {
    float q = std::nearbyint(x * 4.0f);             // quarter turn index, exact
    float r = x - q * 0.25f;                         // in [-1/8, 1/8], exact
    int qi = ((int)q) & 3;
    const float two_pi = 6.28318530717958647692f;
    float a = r * two_pi;
    float a2 = a * a;
    float sp = std::fmaf(a2, (float)(1.0 / 362880.0), (float)(-1.0 / 5040.0));
    sp = std::fmaf(a2, sp, (float)(1.0 / 120.0));
    sp = std::fmaf(a2, sp, (float)(-1.0 / 6.0));
    sp = a2 * sp;
    float sn = std::fmaf(a, sp, a);
    float cp = std::fmaf(a2, (float)(-1.0 / 3628800.0), (float)(1.0 / 40320.0));
    cp = std::fmaf(a2, cp, (float)(-1.0 / 720.0));
    cp = std::fmaf(a2, cp, (float)(1.0 / 24.0));
    cp = std::fmaf(a2, cp, -0.5f);
    float cs = std::fmaf(a2, cp, 1.0f);
    switch (qi) {
        case 0: s = sn; c = cs; break;
        case 1: s = cs; c = -sn; break;
        case 2: s = -sn; c = -cs; break;
        default: s = -cs; c = sn; break;
    }
}

// randomSpherePoint (random.glsl:10-19) fed with r = 2*xi - 1: phi = (r.x + 1) * pi = 2*pi*xi1
// (computed in turns), z = r.y = 2*xi2 - 1 (exact for 24-bit xi).
static inline f3 sphere_point(float xi1, float xi2) {
    float s, c;
    sincos_turns(xi1, s, c);
    float z = xi2 * 2.0f - 1.0f;
    float rr = std::sqrt(std::max(0.0f, std::fmaf(-z, z, 1.0f)));
    return f3{rr * c, rr * s, z};
}
// Cosine-weighted direction about n: normalize(n + randomSpherePoint(...))
// (secondaryRays.comp:72). A degenerate sum (|n + p| == 0) falls back to n.
static inline f3 cosine_dir(f3 n, float xi1, float xi2) {
    f3 v = add(n, sphere_point(xi1, xi2));
    float l2 = dot(v, v);
    if (!(l2 > 1e-20f)) return n;
    float inv = 1.0f / std::sqrt(l2);
    return mul(v, inv);
}

// ---------------------------------------------------------------------------------------
// Camera. Restates camera.ixx:26-44 (direction/view/projection), math.ixx:822-844 (look),
// math.ixx:848-860 (perspective, reverse-Z infinite) and the inversion done by
// pathtracer.ixx:100-101, in closed form: invProj*(X,Y,1,1).xyz = (X/w, Y/h, 1) with
// h = 1/tan(fov/2), w = h*H/W, and invView maps view x to -s, view y to u, view z to dir.
// The basis is derived in double on the host and rounded to float once.
struct Camera {
    float pos[3]; float yaw, pitch, vfov; uint32_t width, height;
};
struct CamBasis {
    f3 pos, right, up, fwd;   // right = -s, up = u, fwd = dir
    float sx, sy;             // 1/w, 1/h
    float pitch_x, pitch_y;   // 1/W, 1/H (OutPitch, primaryRay.comp:42)
};
static inline CamBasis camera_basis(const Camera& c) {
    double cp = std::cos((double)c.pitch), sp = std::sin((double)c.pitch);
    double cy = std::cos((double)c.yaw), sy = std::sin((double)c.yaw);
    double d[3] = {cp * cy, cp * sy, sp};
    // s = normalize(cross(up, dir)) with up = +Z: cross((0,0,1), d) = (-d.y, d.x, 0)
    double s[3] = {-d[1], d[0], 0.0};
    double sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    s[0] /= sl; s[1] /= sl; s[2] /= sl;
    // u = cross(dir, s)
    double u[3] = {d[1] * s[2] - d[2] * s[1], d[2] * s[0] - d[0] * s[2], d[0] * s[1] - d[1] * s[0]};
    double h = 1.0 / std::tan(0.5 * (double)c.vfov);
    double w = h * ((double)c.height / (double)c.width);
    CamBasis b;
    b.pos = f3{c.pos[0], c.pos[1], c.pos[2]};
    b.right = f3{(float)-s[0], (float)-s[1], (float)-s[2]};
    b.up = f3{(float)u[0], (float)u[1], (float)u[2]};
    b.fwd = f3{(float)d[0], (float)d[1], (float)d[2]};
    b.sx = (float)(1.0 / w);
    b.sy = (float)(1.0 / h);
    b.pitch_x = 1.0f / (float)c.width;
    b.pitch_y = 1.0f / (float)c.height;
    return b;
}
// Ray generation, primaryRay.comp:46-51: uv = (Gid + 0.5) * OutPitch, uv.y = 1 - uv.y,
// ndc = uv * 2 - 1, dir = normalize(invView * (invProj * (ndc, 1)).xyz).
static inline void primary_ray(const CamBasis& b, uint32_t x, uint32_t y, f3& o, f3& d) {
    float u = ((float)x + 0.5f) * b.pitch_x;
    float v = 1.0f - ((float)y + 0.5f) * b.pitch_y;
    float X = u * 2.0f - 1.0f;
    float Y = v * 2.0f - 1.0f;
    float a = X * b.sx, c = Y * b.sy;
    f3 dir{std::fmaf(b.right.x, a, std::fmaf(b.up.x, c, b.fwd.x)),
           std::fmaf(b.right.y, a, std::fmaf(b.up.y, c, b.fwd.y)),
           std::fmaf(b.right.z, a, std::fmaf(b.up.z, c, b.fwd.z))};
    d = normalize(dir);
    o = b.pos;
}

// ---------------------------------------------------------------------------------------
// Watertight ray/triangle intersection (Woop, Benthin & Wald, JCGT 2(1), 2013), two-sided,
// with the double-precision fallback for zero edge functions. Replaces raySphereIntersect
// (intersect.glsl:26-37) on the triangle scene; the closest-hit acceptance follows
// primaryRay.comp:28 (strict t < best) extended by the tie rule "smaller ID wins at equal t".
struct RayPre {
    f3 o; int kx, ky, kz; float Sx, Sy, Sz; float tmin, tmax; bool valid;
};
static inline RayPre ray_pre(f3 o, f3 d, float tmin, float tmax) {
    RayPre r;
    r.o = o; r.tmin = tmin; r.tmax = tmax;
    float ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    int kz = (ax >= ay) ? ((ax >= az) ? 0 : 2) : ((ay >= az) ? 1 : 2);
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    float dz = get(d, kz);
    if (dz < 0.0f) std::swap(kx, ky);
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.valid = (dz != 0.0f);
    r.Sx = get(d, kx) / dz;
    r.Sy = get(d, ky) / dz;
    r.Sz = 1.0f / dz;
    return r;
}
// Returns true and updates (best_t, best_id, u, v) if triangle `id` is a closer hit than the
// current best under the tie rule, inside [tmin, tmax].
static inline bool intersect_tri(const RayPre& r, f3 v0, f3 v1, f3 v2, uint32_t id,
                                 float& best_t, uint32_t& best_id, float& bu, float& bv) {
    if (!r.valid) return false;
    f3 A = sub(v0, r.o), B = sub(v1, r.o), C = sub(v2, r.o);
    float Akz = get(A, r.kz), Bkz = get(B, r.kz), Ckz = get(C, r.kz);
    float Ax = std::fmaf(-r.Sx, Akz, get(A, r.kx));
    float Ay = std::fmaf(-r.Sy, Akz, get(A, r.ky));
    float Bx = std::fmaf(-r.Sx, Bkz, get(B, r.kx));
    float By = std::fmaf(-r.Sy, Bkz, get(B, r.ky));
    float Cx = std::fmaf(-r.Sx, Ckz, get(C, r.kx));
    float Cy = std::fmaf(-r.Sy, Ckz, get(C, r.ky));
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
    float T = std::fmaf(U, Az, std::fmaf(V, Bz, W * Cz));
    float rcp = 1.0f / det;
    float t = T * rcp;
    if (!(t >= r.tmin && t <= r.tmax)) return false;
    if (t > best_t || (t == best_t && id >= best_id)) return false;
    best_t = t; best_id = id; bu = V * rcp; bv = W * rcp;
    return true;
}

// ---------------------------------------------------------------------------------------
// Ray origin offset for spawned rays. Published method: Waechter & Binder, "A Fast and
// Robust Method for Avoiding Self-Intersection", Ray Tracing Gems (2019), ch. 6.
// Integer ulp offset along the normal away from the origin, a float offset near zero.
static inline float off1(float p, float n) {
    const float origin = 1.0f / 32.0f, float_scale = 1.0f / 65536.0f, int_scale = 256.0f;
    int of = (int)(int_scale * n);
    int32_t pi; std::memcpy(&pi, &p, 4);
    pi += (p < 0.0f) ? -of : of;
    float pf; std::memcpy(&pf, &pi, 4);
    return std::fabs(p) < origin ? p + float_scale * n : pf;
}
static inline f3 offset_ray(f3 p, f3 n) { return f3{off1(p.x, n.x), off1(p.y, n.y), off1(p.z, n.z)}; }

// Surface point and two-sided geometric normal (facing against d) from a hit: p = v0 +
// u*(v1-v0) + v*(v2-v0) with fmaf, n = normalize(cross(v1-v0, v2-v0)) flipped to face -d.
static inline void hit_frame(f3 v0, f3 v1, f3 v2, float u, float v, f3 d, f3& p, f3& n) {
    f3 e1 = sub(v1, v0), e2 = sub(v2, v0);
    p = f3{std::fmaf(u, e1.x, std::fmaf(v, e2.x, v0.x)),
           std::fmaf(u, e1.y, std::fmaf(v, e2.y, v0.y)),
           std::fmaf(u, e1.z, std::fmaf(v, e2.z, v0.z))};
    n = normalize(cross(e1, e2));
    if (dot(n, d) > 0.0f) n = f3{-n.x, -n.y, -n.z};
}

}  // namespace orc
