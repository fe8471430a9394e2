// sky.h — device helpers of the physical sky shared by sky.hip (LUT generation) and the path
// tracer's miss shading (paths.hip). Functions restate src/gpu/sky/skyAccess.glsl (LUT
// parameterizations, :11-117), src/gpu/sky/sky.glsl:52-71 (ray/sphere) and :128-152 (sun disc)
// and skyColor (src/gpu/secondaryRays.comp:36-58). Plain separately rounded float operations;
// exp, sin, cos, acos and x^1.5 are the basic-operation versions of transc.h, so the CPU
// restatement reproduces every LUT texel bit for bit.
#pragma once
#include "internal.h"
#include "transc.h"

constexpr int kSkyTransW = 256, kSkyTransH = 64;   // Atmosphere::TransmittanceSize (sky.ixx)
constexpr int kSkyMsW = 32, kSkyMsH = 32;          // Atmosphere::MultiScatteringSize
constexpr int kSkyViewW = 192, kSkyViewH = 108;    // Sky::ViewSize

// Everything the miss shading needs, by value in kernel arguments.
struct SkyDev {
    mrt_atmosphere_params p;
    const float4* trans;
    const float4* view;
    float3 probe, sun, illum;
    int sun_disc;
};

namespace sky {

__device__ __forceinline__ float3 ld3(const float* v) { return make_float3(v[0], v[1], v[2]); }
__device__ __forceinline__ float dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float len(float3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ float3 normalize(float3 a) { return a * (1.0f / len(a)); }
__device__ __forceinline__ float3 cross(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float clampf(float x, float a, float b) { return fminf(fmaxf(x, a), b); }

__device__ __forceinline__ float sphere_nearest(float3 r0, float3 rd, float3 s0, float R) {
    const float a = dot(rd, rd);
    const float3 d = r0 - s0;
    const float b = 2.0f * dot(rd, d);
    const float c = dot(d, d) - R * R;
    const float delta = b * b - 4.0f * a * c;
    if (delta < 0.0f || a == 0.0f) return -1.0f;
    const float s0v = (-b - sqrtf(delta)) / (2.0f * a);
    const float s1v = (-b + sqrtf(delta)) / (2.0f * a);
    if (s0v < 0.0f && s1v < 0.0f) return -1.0f;
    if (s0v < 0.0f) return fmaxf(0.0f, s1v);
    if (s1v < 0.0f) return fmaxf(0.0f, s0v);
    return fmaxf(0.0f, fminf(s0v, s1v));
}

__device__ __forceinline__ float unit_to_sub(float u, float res) { return (u + 0.5f / res) * (res / (res + 1.0f)); }
__device__ __forceinline__ float sub_to_unit(float u, float res) { return (u - 0.5f / res) * (res / (res - 1.0f)); }

__device__ __forceinline__ void uv_to_trans_params(float u, float v, float bottom, float top, float& h, float& mu) {
    const float H = sqrtf(top * top - bottom * bottom);
    const float rho = H * v;
    h = sqrtf(rho * rho + bottom * bottom);
    const float dmin = top - h, dmax = rho + H;
    const float d = dmin + u * (dmax - dmin);
    mu = d == 0.0f ? 1.0f : (H * H - rho * rho - d * d) / (2.0f * h * d);
    mu = clampf(mu, -1.0f, 1.0f);
}

__device__ __forceinline__ void trans_params_to_uv(float h, float mu, float bottom, float top, float& u, float& v) {
    const float H = sqrtf(fmaxf(0.0f, top * top - bottom * bottom));
    const float rho = sqrtf(fmaxf(0.0f, h * h - bottom * bottom));
    const float disc = h * h * (mu * mu - 1.0f) + top * top;
    const float d = fmaxf(0.0f, (-h * mu + sqrtf(disc)));
    const float dmin = top - h, dmax = rho + H;
    u = (d - dmin) / (dmax - dmin);
    v = rho / H;
}

__device__ __forceinline__ void uv_to_view_params(float u, float v, int w, int hgt, float view_h, float bottom, float& vz,
                                                  float& lv) {
    u = sub_to_unit(u, (float)w);
    v = sub_to_unit(v, (float)hgt);
    const float vhor = sqrtf(view_h * view_h - bottom * bottom);
    const float beta = rt::tx_acos(vhor / view_h);
    const float zh = 3.14159265358979323846f - beta;
    if (v < 0.5f) {
        float c = 2.0f * v;
        c = 1.0f - c;
        c *= c;
        c = 1.0f - c;
        vz = rt::tx_cos(zh * c);
    } else {
        float c = v * 2.0f - 1.0f;
        c *= c;
        vz = rt::tx_cos(zh + beta * c);
    }
    float c = u;
    c *= c;
    lv = -(c * 2.0f - 1.0f);
}

// fmaxf(0, .) before the square roots: at the horizon the ground test and the angle can
// disagree by one rounding (the reference would take sqrt of a tiny negative number there).
__device__ __forceinline__ void view_params_to_uv(bool ground, float vz, float lv, int w, int hgt, float view_h,
                                                  float bottom, float& u, float& v) {
    const float vhor = sqrtf(view_h * view_h - bottom * bottom);
    const float beta = rt::tx_acos(vhor / view_h);
    const float zh = 3.14159265358979323846f - beta;
    if (!ground) {
        float c = rt::tx_acos(vz) / zh;
        c = 1.0f - c;
        c = sqrtf(fmaxf(c, 0.0f));
        c = 1.0f - c;
        v = c * 0.5f;
    } else {
        float c = (rt::tx_acos(vz) - zh) / beta;
        c = sqrtf(fmaxf(c, 0.0f));
        v = c * 0.5f + 0.5f;
    }
    float c = -lv * 0.5f + 0.5f;
    c = sqrtf(c);
    u = unit_to_sub(c, (float)w);
    v = unit_to_sub(v, (float)hgt);
}

// textureLod with a linear clamp-to-edge sampler over w x h float4 texels
__device__ __forceinline__ float3 lut(const float4* __restrict__ t, int w, int h, float u, float v) {
    const float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float ax = x - fx, ay = y - fy;
    int x0 = (int)fx, y0 = (int)fy, x1 = x0 + 1, y1 = y0 + 1;
    x0 = min(max(x0, 0), w - 1);
    x1 = min(max(x1, 0), w - 1);
    y0 = min(max(y0, 0), h - 1);
    y1 = min(max(y1, 0), h - 1);
    const float4 t00 = t[(size_t)y0 * w + x0], t10 = t[(size_t)y0 * w + x1];
    const float4 t01 = t[(size_t)y1 * w + x0], t11 = t[(size_t)y1 * w + x1];
    const float3 a = make_float3(t00.x, t00.y, t00.z) * (1.0f - ax) + make_float3(t10.x, t10.y, t10.z) * ax;
    const float3 b = make_float3(t01.x, t01.y, t01.z) * (1.0f - ax) + make_float3(t11.x, t11.y, t11.z) * ax;
    return a * (1.0f - ay) + b * ay;
}

// skyColor(dir) (secondaryRays.comp:36-58): sky-view LUT plus, with sun_disc, the limb-darkened
// sun disc (getSunLuminance, sky.glsl:128-152) scaled by 120000 / illuminance.
__device__ __forceinline__ float3 radiance(const SkyDev& s, float3 d) {
    const float R = s.p.bottom_radius;
    const float3 pos = s.probe + make_float3(0.0f, 0.0f, R);
    const float3 up = normalize(pos);
    const float vz = dot(d, up);
    const float vh = len(pos);
    const float3 side = normalize(cross(up, d));
    const float3 fwd = normalize(cross(side, up));
    float lx = dot(s.sun, fwd);
    const float ly = dot(s.sun, side);
    lx = lx / sqrtf(lx * lx + ly * ly);
    const bool ground = sphere_nearest(pos, d, make_float3(0.0f, 0.0f, 0.0f), R) >= 0.0f;
    float u, v;
    view_params_to_uv(ground, vz, lx, kSkyViewW, kSkyViewH, vh, R, u, v);
    float3 L = lut(s.view, kSkyViewW, kSkyViewH, u, v);
    if (s.sun_disc) {
        const float sun_r = 0.5f * 0.505f * 3.14159f / 180.0f;
        if (dot(d, s.sun) > rt::tx_cos(sun_r) && !ground) {
            float uu, vu, us, vs;
            trans_params_to_uv(R, 1.0f, R, s.p.top_radius, uu, vu);
            const float3 upv = pos * (1.0f / vh);
            trans_params_to_uv(vh, dot(s.sun, upv), R, s.p.top_radius, us, vs);
            const float ang = rt::tx_acos(clampf(dot(d, s.sun), -1.0f, 1.0f));
            const float rr = ang / sun_r;
            const float limb = sqrtf(clampf(1.0f - rr * rr, 0.0001f, 1.0f));
            const float3 space = s.illum / lut(s.trans, kSkyTransW, kSkyTransH, uu, vu);
            const float3 sun = (space * lut(s.trans, kSkyTransW, kSkyTransH, us, vs)) * limb;
            L = L + sun * (make_float3(120000.0f, 120000.0f, 120000.0f) / s.illum);
        }
    }
    return L;
}

}  // namespace sky

// The device view of a sky with a current view LUT (false: no sky or no view yet).
bool sky_device(const mrt_sky* s, SkyDev* out);
