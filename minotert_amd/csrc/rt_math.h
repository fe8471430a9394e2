// rt_math.h — scalar ray-tracing core shared by the HIP kernels and the host code of the
// minotert library (both sides compile it; -ffp-contract=off everywhere, so every fused
// multiply-add below is an explicit fmaf and everything else is a separately rounded op).
//
// Reference semantics (MinoteRT, paths relative to its repository root):
//   pcg / randomFloat ........ src/gpu/random.glsl:22-31
//   randomSpherePoint ........ src/gpu/random.glsl:10-19, used as normalize(n + p) at
//                              src/gpu/secondaryRays.comp:72 (cosine-weighted direction)
//   pinhole ray generation ... src/gpu/primaryRay.comp:46-51 with the matrices of
//                              src/gfx/camera.ixx:26-44 and src/stx/math.ixx:822-860
//   closest-hit acceptance ... src/gpu/primaryRay.comp:28 (strict t < best), extended by the
//                              rule that the smaller triangle ID wins at equal t
#pragma once
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define RT_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define RT_HD static inline
#endif

namespace rt {

struct v3 { float x, y, z; };
RT_HD v3 make_v3(float x, float y, float z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
RT_HD v3 operator-(v3 a, v3 b) { return make_v3(a.x - b.x, a.y - b.y, a.z - b.z); }
RT_HD v3 operator+(v3 a, v3 b) { return make_v3(a.x + b.x, a.y + b.y, a.z + b.z); }
RT_HD v3 operator*(v3 a, float s) { return make_v3(a.x * s, a.y * s, a.z * s); }
RT_HD v3 vmul(v3 a, v3 b) { return make_v3(a.x * b.x, a.y * b.y, a.z * b.z); }
RT_HD float vdot(v3 a, v3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }
RT_HD v3 vcross(v3 a, v3 b) {
    return make_v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
RT_HD v3 vnormalize(v3 v) { float inv = 1.0f / sqrtf(vdot(v, v)); return v * inv; }
RT_HD float comp(v3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

RT_HD uint32_t f2u(float f) { union { float f; uint32_t u; } c; c.f = f; return c.u; }
RT_HD float u2f(uint32_t u) { union { float f; uint32_t u; } c; c.u = u; return c.f; }

// ---- RNG: PCG-RXS-M-XS 32-bit hash (Jarzynski & Olano 2020), random.glsl:22-27 -------------
RT_HD uint32_t pcg_hash(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
// randomFloat (random.glsl:29-31): advance the state, low 24 bits / 2^24.
RT_HD float next_float(uint32_t& state) {
    state = pcg_hash(state);
    return (float)(state & 0xFFFFFFu) * (1.0f / 16777216.0f);
}
// Per-path seed: pcg(pixel ^ pcg(sample ^ pcg(frame_seed))). Counter based, so the random
// stream of a path does not depend on the order in which paths are processed.
RT_HD uint32_t path_seed(uint32_t pixel, uint32_t sample, uint32_t frame_seed) {
    return pcg_hash(pixel ^ pcg_hash(sample ^ pcg_hash(frame_seed)));
}

// ---- sin/cos of 2*pi*x, x = k/2^24 ---------------------------------------------------------
// Exact reduction to the nearest quarter turn, then degree-9/10 Taylor polynomials on
// [-pi/4, pi/4] (truncation error < 2e-9). Host and device evaluate the same sequence.
RT_HD void sincos_turns(float x, float& s, float& c) {
    float q = rintf(x * 4.0f);
    float r = x - q * 0.25f;
    int qi = ((int)q) & 3;
    const float two_pi = 6.28318530717958647692f;
    float a = r * two_pi;
    float a2 = a * a;
    float sp = fmaf(a2, (float)(1.0 / 362880.0), (float)(-1.0 / 5040.0));
    sp = fmaf(a2, sp, (float)(1.0 / 120.0));
    sp = fmaf(a2, sp, (float)(-1.0 / 6.0));
    sp = a2 * sp;
    float sn = fmaf(a, sp, a);
    float cp = fmaf(a2, (float)(-1.0 / 3628800.0), (float)(1.0 / 40320.0));
    cp = fmaf(a2, cp, (float)(-1.0 / 720.0));
    cp = fmaf(a2, cp, (float)(1.0 / 24.0));
    cp = fmaf(a2, cp, -0.5f);
    float cs = fmaf(a2, cp, 1.0f);
    float s0 = (qi & 1) ? cs : sn;
    float c0 = (qi & 1) ? sn : cs;
    s = (qi == 2 || qi == 3) ? -s0 : s0;
    c = (qi == 1 || qi == 2) ? -c0 : c0;
}

// Uniform point on the unit sphere from two uniforms (randomSpherePoint with r = 2*xi - 1).
RT_HD v3 sphere_point(float xi1, float xi2) {
    float s, c;
    sincos_turns(xi1, s, c);
    float z = xi2 * 2.0f - 1.0f;
    float rr = sqrtf(fmaxf(0.0f, fmaf(-z, z, 1.0f)));
    return make_v3(rr * c, rr * s, z);
}
// Cosine-weighted direction about the unit normal n: normalize(n + sphere point).
RT_HD v3 cosine_direction(v3 n, float xi1, float xi2) {
    v3 v = n + sphere_point(xi1, xi2);
    float l2 = vdot(v, v);
    if (!(l2 > 1e-20f)) return n;
    float inv = 1.0f / sqrtf(l2);
    return v * inv;
}

// ---- camera --------------------------------------------------------------------------------
// Float basis derived once on the host (see camera_basis in scene_host.cpp).
struct CamBasis {
    v3 pos, right, up, fwd;
    float sx, sy, pitch_x, pitch_y;
    uint32_t width, height;
};
RT_HD void camera_ray(const CamBasis& b, uint32_t x, uint32_t y, v3& o, v3& d) {
    float u = ((float)x + 0.5f) * b.pitch_x;
    float v = 1.0f - ((float)y + 0.5f) * b.pitch_y;
    float X = u * 2.0f - 1.0f;
    float Y = v * 2.0f - 1.0f;
    float a = X * b.sx, c = Y * b.sy;
    v3 dir = make_v3(fmaf(b.right.x, a, fmaf(b.up.x, c, b.fwd.x)),
                     fmaf(b.right.y, a, fmaf(b.up.y, c, b.fwd.y)),
                     fmaf(b.right.z, a, fmaf(b.up.z, c, b.fwd.z)));
    d = vnormalize(dir);
    o = b.pos;
}

// ---- watertight ray/triangle test (Woop, Benthin & Wald, JCGT 2013) ---------------------------
struct RayShear {
    v3 o;
    int kx, ky, kz;
    float Sx, Sy, Sz;
    float tmin;
    bool valid;
};
RT_HD RayShear make_shear(v3 o, v3 d, float tmin) {
    RayShear r;
    r.o = o; r.tmin = tmin;
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = (ax >= ay) ? ((ax >= az) ? 0 : 2) : ((ay >= az) ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    float dz = comp(d, kz);
    if (dz < 0.0f) { int t = kx; kx = ky; ky = t; }
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.valid = dz != 0.0f;
    r.Sx = comp(d, kx) / dz;
    r.Sy = comp(d, ky) / dz;
    r.Sz = 1.0f / dz;
    return r;
}
// Tests triangle `id` against the ray; on a closer hit (t in [tmin, best_t], smaller ID at
// equal t) it updates best_t / best_id / bu / bv and returns true.
RT_HD bool watertight_hit(const RayShear& r, v3 v0, v3 v1, v3 v2, uint32_t id,
                          float& best_t, uint32_t& best_id, float& bu, float& bv) {
    v3 A = v0 - r.o, B = v1 - r.o, C = v2 - r.o;
    float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
    float Ax = fmaf(-r.Sx, Akz, comp(A, r.kx));
    float Ay = fmaf(-r.Sy, Akz, comp(A, r.ky));
    float Bx = fmaf(-r.Sx, Bkz, comp(B, r.kx));
    float By = fmaf(-r.Sy, Bkz, comp(B, r.ky));
    float Cx = fmaf(-r.Sx, Ckz, comp(C, r.kx));
    float Cy = fmaf(-r.Sy, Ckz, comp(C, r.ky));
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
    float T = fmaf(U, Az, fmaf(V, Bz, W * Cz));
    float rcp = 1.0f / det;
    float t = T * rcp;
    if (!(t >= r.tmin && t <= best_t)) return false;
    if (t == best_t && id >= best_id) return false;
    best_t = t; best_id = id; bu = V * rcp; bv = W * rcp;
    return true;
}

// ---- spawned-ray origin (Waechter & Binder, Ray Tracing Gems 2019, ch. 6) ---------------------
RT_HD float offset1(float p, float n) {
    const float origin = 1.0f / 32.0f, float_scale = 1.0f / 65536.0f, int_scale = 256.0f;
    int of = (int)(int_scale * n);
    int32_t pi = (int32_t)f2u(p);
    pi += (p < 0.0f) ? -of : of;
    return fabsf(p) < origin ? p + float_scale * n : u2f((uint32_t)pi);
}
RT_HD v3 offset_origin(v3 p, v3 n) { return make_v3(offset1(p.x, n.x), offset1(p.y, n.y), offset1(p.z, n.z)); }

// Surface point from barycentrics and the two-sided face normal (facing against d).
RT_HD void surface_frame(v3 v0, v3 v1, v3 v2, float u, float v, v3 d, v3& p, v3& n) {
    v3 e1 = v1 - v0, e2 = v2 - v0;
    p = make_v3(fmaf(u, e1.x, fmaf(v, e2.x, v0.x)),
                fmaf(u, e1.y, fmaf(v, e2.y, v0.y)),
                fmaf(u, e1.z, fmaf(v, e2.z, v0.z)));
    n = vnormalize(vcross(e1, e2));
    if (vdot(n, d) > 0.0f) n = make_v3(-n.x, -n.y, -n.z);
}

// Conservative slab-test reciprocal: a zero component becomes +-1e30 so that (lo - o) * inv
// never forms 0 * inf.
RT_HD float safe_rcp(float d) { return d == 0.0f ? copysignf(1e30f, d) : 1.0f / d; }

}  // namespace rt
