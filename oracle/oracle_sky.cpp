// oracle_sky.cpp — scalar CPU restatement of the reference's physical sky (test infrastructure
// only): the atmosphere LUTs of Hillaire 2020 as the reference computes them
//   transmittance 256x64        src/gpu/sky/genTransmittance.comp:22-43
//   multiple scattering 32x32   src/gpu/sky/genMultiScattering.comp:25-145 (64 directions per
//                               texel, tree reduction 64 -> 1 in the shader's order)
//   sky view 192x108            src/gpu/sky/genView.comp:33-76
// with the shared functions of src/gpu/sky/sky.glsl (medium :94-123, phase functions :38-50,
// ray/sphere :52-71, integration :171-326, sun disc :128-152) and skyAccess.glsl (LUT
// parameterizations :11-117), and the miss radiance skyColor (src/gpu/secondaryRays.comp:36-58).
// LUTs are FP32 here (the reference stores RGBA16F / B10G11R11: unpinned), sampled bilinearly
// with clamp-to-edge like a LinearClamp sampler. exp, sin, cos, acos and x^1.5 are the
// basic-operation restatements of oracle_transc.hpp (the GLSL built-ins are driver-defined).
#include "oracle.h"
#include "oracle_transc.hpp"

#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kPlanetRadiusOffset = 0.01f;

struct V { float x, y, z; };
inline V v3(float x, float y, float z) { return V{x, y, z}; }
inline V add(V a, V b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V sub(V a, V b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V mul(V a, V b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V scl(V a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
inline V dvd(V a, V b) { return v3(a.x / b.x, a.y / b.y, a.z / b.z); }
inline float dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float len(V a) { return std::sqrt(dot(a, a)); }
inline V vexp(V a) { return v3(orc_tx::exp(a.x), orc_tx::exp(a.y), orc_tx::exp(a.z)); }
inline V load(const float* p) { return v3(p[0], p[1], p[2]); }
inline V vmax(V a, float m) { return v3(std::fmax(a.x, m), std::fmax(a.y, m), std::fmax(a.z, m)); }
inline float clampf(float x, float a, float b) { return std::fmin(std::fmax(x, a), b); }
inline V normalize(V a) { return scl(a, 1.0f / len(a)); }
inline V cross(V a, V b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

struct Lut {
    int w = 0, h = 0;
    const float* rgb = nullptr;  // w*h*3
    // textureLod with a linear clamp-to-edge sampler
    V sample(float u, float v) const {
        const float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
        const float fx = std::floor(x), fy = std::floor(y);
        const float ax = x - fx, ay = y - fy;
        int x0 = (int)fx, y0 = (int)fy, x1 = x0 + 1, y1 = y0 + 1;
        x0 = x0 < 0 ? 0 : (x0 >= w ? w - 1 : x0);
        x1 = x1 < 0 ? 0 : (x1 >= w ? w - 1 : x1);
        y0 = y0 < 0 ? 0 : (y0 >= h ? h - 1 : y0);
        y1 = y1 < 0 ? 0 : (y1 >= h ? h - 1 : y1);
        auto t = [&](int xx, int yy) { return load(rgb + ((size_t)yy * w + xx) * 3); };
        const V a = add(scl(t(x0, y0), 1.0f - ax), scl(t(x1, y0), ax));
        const V b = add(scl(t(x0, y1), 1.0f - ax), scl(t(x1, y1), ax));
        return add(scl(a, 1.0f - ay), scl(b, ay));
    }
};

struct Medium { V scattering, extinction, scat_mie, scat_ray; };

struct Atmo {
    const orc_atmosphere* p;
    Lut trans, ms;
    bool has_trans = false, has_ms = false;

    Medium medium(V pos) const {  // sky.glsl:94-123
        const float h = len(pos) - p->bottom_radius;
        const float dmie = orc_tx::exp(p->mie_density_exp_scale * h);
        const float dray = orc_tx::exp(p->rayleigh_density_exp_scale * h);
        const float dozo = clampf(h < p->absorption_density0_layer_width
                                      ? p->absorption_density0_linear * h + p->absorption_density0_constant
                                      : p->absorption_density1_linear * h + p->absorption_density1_constant,
                                  0.0f, 1.0f);
        Medium m;
        m.scat_mie = scl(load(p->mie_scattering), dmie);
        const V ext_mie = scl(load(p->mie_extinction), dmie);
        m.scat_ray = scl(load(p->rayleigh_scattering), dray);
        const V ext_ray = m.scat_ray;
        const V ext_ozo = scl(load(p->absorption_extinction), dozo);
        m.scattering = add(m.scat_mie, m.scat_ray);
        m.extinction = add(add(ext_mie, ext_ray), ext_ozo);
        return m;
    }
};

float mie_phase(float g, float cos_t) {  // Cornette-Shanks, sky.glsl:38-41 (called with -cos)
    const float k = 3.0f / (8.0f * kPi) * (1.0f - g * g) / (2.0f + g * g);
    return k * (1.0f + cos_t * cos_t) / orc_tx::pow15(1.0f + g * g - 2.0f * g * -cos_t);
}
float rayleigh_phase(float cos_t) { return 3.0f / (16.0f * kPi) * (1.0f + cos_t * cos_t); }

float sphere_nearest(V r0, V rd, V s0, float R) {  // sky.glsl:52-71
    const float a = dot(rd, rd);
    const V d = sub(r0, s0);
    const float b = 2.0f * dot(rd, d);
    const float c = dot(d, d) - R * R;
    const float delta = b * b - 4.0f * a * c;
    if (delta < 0.0f || a == 0.0f) return -1.0f;
    const float s0v = (-b - std::sqrt(delta)) / (2.0f * a);
    const float s1v = (-b + std::sqrt(delta)) / (2.0f * a);
    if (s0v < 0.0f && s1v < 0.0f) return -1.0f;
    if (s0v < 0.0f) return std::fmax(0.0f, s1v);
    if (s1v < 0.0f) return std::fmax(0.0f, s0v);
    return std::fmax(0.0f, std::fmin(s0v, s1v));
}

float unit_to_sub(float u, float res) { return (u + 0.5f / res) * (res / (res + 1.0f)); }
float sub_to_unit(float u, float res) { return (u - 0.5f / res) * (res / (res - 1.0f)); }

void uv_to_trans_params(float u, float v, float bottom, float top, float& h, float& mu) {  // skyAccess.glsl:19-35
    const float H = std::sqrt(top * top - bottom * bottom);
    const float rho = H * v;
    h = std::sqrt(rho * rho + bottom * bottom);
    const float dmin = top - h, dmax = rho + H;
    const float d = dmin + u * (dmax - dmin);
    mu = d == 0.0f ? 1.0f : (H * H - rho * rho - d * d) / (2.0f * h * d);
    mu = clampf(mu, -1.0f, 1.0f);
}

void trans_params_to_uv(float h, float mu, float bottom, float top, float& u, float& v) {  // skyAccess.glsl:37-55
    const float H = std::sqrt(std::fmax(0.0f, top * top - bottom * bottom));
    const float rho = std::sqrt(std::fmax(0.0f, h * h - bottom * bottom));
    const float disc = h * h * (mu * mu - 1.0f) + top * top;
    const float d = std::fmax(0.0f, (-h * mu + std::sqrt(disc)));
    const float dmin = top - h, dmax = rho + H;
    u = (d - dmin) / (dmax - dmin);
    v = rho / H;
}

void uv_to_view_params(float u, float v, int w, int hgt, float view_h, float bottom, float& vz, float& lv) {  // :57-87
    u = sub_to_unit(u, (float)w);
    v = sub_to_unit(v, (float)hgt);
    const float vhor = std::sqrt(view_h * view_h - bottom * bottom);
    const float cos_beta = vhor / view_h;
    const float beta = orc_tx::acos(cos_beta);
    const float zh = kPi - beta;
    if (v < 0.5f) {
        float c = 2.0f * v;
        c = 1.0f - c;
        c *= c;
        c = 1.0f - c;
        vz = orc_tx::cos(zh * c);
    } else {
        float c = v * 2.0f - 1.0f;
        c *= c;
        vz = orc_tx::cos(zh + beta * c);
    }
    float c = u;
    c *= c;
    lv = -(c * 2.0f - 1.0f);
}

void view_params_to_uv(bool ground, float vz, float lv, int w, int hgt, float view_h, float bottom, float& u,
                       float& v) {  // skyAccess.glsl:89-117
    const float vhor = std::sqrt(view_h * view_h - bottom * bottom);
    const float cos_beta = vhor / view_h;
    const float beta = orc_tx::acos(cos_beta);
    const float zh = kPi - beta;
    // fmax(0, .) before the square roots: at the horizon the ground test and the angle can
    // disagree by one rounding, which would give sqrt of a tiny negative number
    if (!ground) {
        float c = orc_tx::acos(vz) / zh;
        c = 1.0f - c;
        c = std::sqrt(std::fmax(c, 0.0f));
        c = 1.0f - c;
        v = c * 0.5f;
    } else {
        float c = (orc_tx::acos(vz) - zh) / beta;
        c = std::sqrt(std::fmax(c, 0.0f));
        v = c * 0.5f + 0.5f;
    }
    float c = -lv * 0.5f + 0.5f;
    c = std::sqrt(c);
    u = c;
    u = unit_to_sub(u, (float)w);
    v = unit_to_sub(v, (float)hgt);
}

V multiple_scattering(const Atmo& A, V pos, float vz) {  // sky.glsl:156-168
    float u = clampf(vz * 0.5f + 0.5f, 0.0f, 1.0f);
    float v = clampf((len(pos) - A.p->bottom_radius) / (A.p->top_radius - A.p->bottom_radius), 0.0f, 1.0f);
    u = unit_to_sub(u, (float)A.ms.w);
    v = unit_to_sub(v, (float)A.ms.h);
    return A.ms.sample(u, v);
}

struct Scatter { V L, optical_depth, ms_as1; };

// sky.glsl:171-326 (newMultiScatStep*Out are not used by any LUT and are left out)
Scatter integrate(const Atmo& A, V pos, V dir, V sun, bool ground, float spp_ini, bool variable, bool mie_ray_phase,
                  float tmax_max, V illum) {
    Scatter r{v3(0, 0, 0), v3(0, 0, 0), v3(0, 0, 0)};
    const V O = v3(0, 0, 0);
    const float t_bottom = sphere_nearest(pos, dir, O, A.p->bottom_radius);
    const float t_top = sphere_nearest(pos, dir, O, A.p->top_radius);
    float tmax = 0.0f;
    if (t_bottom < 0.0f) {
        if (t_top < 0.0f) return r;
        tmax = t_top;
    } else if (t_top > 0.0f) {
        tmax = std::fmin(t_top, t_bottom);
    }
    tmax = std::fmin(tmax, tmax_max);
    float spp = spp_ini, spp_floor = spp_ini, tmax_floor = tmax;
    if (variable) {
        const float a = clampf(tmax * 0.01f, 0.0f, 1.0f);
        spp = 4.0f * (1.0f - a) + 14.0f * a;  // mix(RAYMARCH_MIN_SPP, RAYMARCH_MAX_SPP, a)
        spp_floor = std::floor(spp);
        tmax_floor = tmax * spp_floor / spp;
    }
    float dt = tmax / spp;
    const float uniform_phase = 1.0f / (4.0f * kPi);
    const float cos_t = dot(sun, dir);
    const float mie_p = mie_phase(A.p->mie_phase_g, -cos_t);
    const float ray_p = rayleigh_phase(cos_t);
    V L = v3(0, 0, 0), thr = v3(1, 1, 1), od = v3(0, 0, 0);
    float t = 0.0f;
    const float seg = 0.3f;
    for (float s = 0.0f; s < spp; s += 1.0f) {
        if (variable) {
            float t0 = s / spp_floor, t1 = (s + 1.0f) / spp_floor;
            t0 = t0 * t0;
            t1 = t1 * t1;
            t0 = tmax_floor * t0;
            if (t1 > 1.0f) t1 = tmax;
            else t1 = tmax_floor * t1;
            t = t0 + (t1 - t0) * seg;
            dt = t1 - t0;
        } else {
            const float nt = tmax * (s + seg) / spp;
            dt = nt - t;
            t = nt;
        }
        const V P = add(pos, scl(dir, t));
        const Medium m = A.medium(P);
        const V sod = scl(m.extinction, dt);
        const V st = vexp(scl(sod, -1.0f));
        od = add(od, sod);
        const float ph = len(P);
        const V up = scl(P, 1.0f / ph);
        const float sun_z = dot(sun, up);
        float tu, tv;
        trans_params_to_uv(ph, sun_z, A.p->bottom_radius, A.p->top_radius, tu, tv);
        const V t_sun = A.has_trans ? A.trans.sample(tu, tv) : v3(0, 0, 0);
        const V pts = mie_ray_phase ? add(scl(m.scat_mie, mie_p), scl(m.scat_ray, ray_p)) : scl(m.scattering, uniform_phase);
        const float t_earth = sphere_nearest(P, sun, add(O, scl(up, kPlanetRadiusOffset)), A.p->bottom_radius);
        const float shadow = t_earth >= 0.0f ? 0.0f : 1.0f;
        const V msl = A.has_ms ? multiple_scattering(A, P, sun_z) : v3(0, 0, 0);
        const V S = mul(illum, add(scl(mul(t_sun, pts), shadow), mul(msl, m.scattering)));
        const V MS = m.scattering;
        const V MSint = dvd(sub(MS, mul(MS, st)), m.extinction);
        r.ms_as1 = add(r.ms_as1, mul(thr, MSint));
        const V Sint = dvd(sub(S, mul(S, st)), m.extinction);
        L = add(L, mul(thr, Sint));
        thr = mul(thr, st);
    }
    if (ground && tmax == t_bottom && t_bottom > 0.0f) {
        const V P = add(pos, scl(dir, t_bottom));
        const float ph = len(P);
        const V up = scl(P, 1.0f / ph);
        const float sun_z = dot(sun, up);
        float tu, tv;
        trans_params_to_uv(ph, sun_z, A.p->bottom_radius, A.p->top_radius, tu, tv);
        const V t_sun = A.has_trans ? A.trans.sample(tu, tv) : v3(0, 0, 0);
        const float ndl = clampf(dot(normalize(up), normalize(sun)), 0.0f, 1.0f);
        L = add(L, scl(mul(mul(mul(illum, t_sun), thr), load(A.p->ground_albedo)), ndl / kPi));
    }
    r.L = L;
    r.optical_depth = od;
    return r;
}

}  // namespace

extern "C" {

// Hook for the tests: fn 0 exp, 1 sin, 2 cos, 3 acos, 4 x^1.5, 5 log, 6 x^(1/2.4) of oracle_transc.hpp.
void orc_tx_eval(int fn, const float* x, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        const float v = x[i];
        out[i] = fn == 0 ? orc_tx::exp(v) : fn == 1 ? orc_tx::sin(v) : fn == 2 ? orc_tx::cos(v)
               : fn == 3 ? orc_tx::acos(v) : fn == 4 ? orc_tx::pow15(v) : fn == 5 ? orc_tx::log(v)
               : orc_tx::pow(v, 1.0f / 2.4f);
    }
}

// Hooks for the tests: the phase functions and LUT parameterizations used above.
float orc_sky_phase(int mie, float g, float cos_t) { return mie ? mie_phase(g, cos_t) : rayleigh_phase(cos_t); }
void orc_sky_trans_uv(float h, float mu, float bottom, float top, float out[2]) {
    trans_params_to_uv(h, mu, bottom, top, out[0], out[1]);
}
void orc_sky_trans_params(float u, float v, float bottom, float top, float out[2]) {
    uv_to_trans_params(u, v, bottom, top, out[0], out[1]);
}
void orc_sky_view_uv(int ground, float vz, float lv, int w, int h, float view_h, float bottom, float out[2]) {
    view_params_to_uv(ground != 0, vz, lv, w, h, view_h, bottom, out[0], out[1]);
}
void orc_sky_view_params(float u, float v, int w, int h, float view_h, float bottom, float out[2]) {
    uv_to_view_params(u, v, w, h, view_h, bottom, out[0], out[1]);
}

void orc_sky_earth(orc_atmosphere* p) {  // Atmosphere::Params::earth(), src/gfx/modules/sky.ixx
    std::memset(p, 0, sizeof(*p));
    p->bottom_radius = 6360.0f;
    p->top_radius = 6460.0f;
    p->rayleigh_density_exp_scale = -1.0f / 8.0f;
    const float ray[3] = {0.005802f, 0.013558f, 0.033100f};
    for (int k = 0; k < 3; ++k) {
        p->rayleigh_scattering[k] = ray[k];
        p->mie_scattering[k] = 0.003996f;
        p->mie_extinction[k] = 0.004440f;
        p->mie_absorption[k] = std::fmax(0.004440f - 0.003996f, 0.0f);
    }
    p->mie_density_exp_scale = -1.0f / 1.2f;
    p->mie_phase_g = 0.8f;
    p->absorption_density0_layer_width = 25.0f;
    p->absorption_density0_constant = -2.0f / 3.0f;
    p->absorption_density0_linear = 1.0f / 15.0f;
    p->absorption_density1_constant = 8.0f / 3.0f;
    p->absorption_density1_linear = -1.0f / 15.0f;
    const float ozo[3] = {0.000650f, 0.001881f, 0.000085f};
    for (int k = 0; k < 3; ++k) p->absorption_extinction[k] = ozo[k];
}

void orc_sky_transmittance(const orc_atmosphere* p, uint32_t w, uint32_t h, float* out) {  // genTransmittance.comp
    Atmo A{p, {}, {}};
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            const float u = ((float)x + 0.5f) / (float)w, v = ((float)y + 0.5f) / (float)h;
            float vh, mu;
            uv_to_trans_params(u, v, p->bottom_radius, p->top_radius, vh, mu);
            const V pos = v3(0.0f, 0.0f, vh);
            const V dir = v3(0.0f, std::sqrt(1.0f - mu * mu), mu);
            const Scatter s = integrate(A, pos, dir, v3(1, 1, 1), false, 40.0f, false, false, 9000000.0f, v3(1, 1, 1));
            const V tr = vexp(scl(s.optical_depth, -1.0f));
            float* o = out + ((size_t)y * w + x) * 3;
            o[0] = tr.x; o[1] = tr.y; o[2] = tr.z;
        }
}

void orc_sky_multiscattering(const orc_atmosphere* p, const float* trans, uint32_t tw, uint32_t th, uint32_t w,
                             uint32_t h, float* out) {  // genMultiScattering.comp
    Atmo A{p, {}, {}};
    A.trans = Lut{(int)tw, (int)th, trans};
    A.has_trans = true;
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            float u = ((float)x + 0.5f) / (float)w, v = ((float)y + 0.5f) / (float)h;
            u = sub_to_unit(u, (float)w);
            v = sub_to_unit(v, (float)h);
            const float cs = u * 2.0f - 1.0f;
            const V sun = v3(0.0f, std::sqrt(clampf(1.0f - cs * cs, 0.0f, 1.0f)), cs);
            const float vh = p->bottom_radius +
                             clampf(v + kPlanetRadiusOffset, 0.0f, 1.0f) * (p->top_radius - p->bottom_radius - kPlanetRadiusOffset);
            const V pos = v3(0.0f, 0.0f, vh);
            const float sphere = 4.0f * kPi, iso = 1.0f / sphere, sq = 8.0f;
            V as1[64], L[64];
            for (uint32_t z = 0; z < 64; ++z) {
                const float i = 0.5f + (float)(z / 8), j = 0.5f + (float)(z - (z / 8) * 8);
                const float theta = 2.0f * kPi * (i / sq), phi = kPi * (j / sq);
                const V dir = v3(orc_tx::cos(theta) * orc_tx::sin(phi), orc_tx::sin(theta) * orc_tx::sin(phi), orc_tx::cos(phi));
                const Scatter s = integrate(A, pos, dir, sun, true, 20.0f, false, false, 9000000.0f, v3(1, 1, 1));
                as1[z] = scl(s.ms_as1, sphere / (sq * sq));
                L[z] = scl(s.L, sphere / (sq * sq));
            }
            for (uint32_t half = 32; half >= 1; half >>= 1)
                for (uint32_t z = 0; z < half; ++z) {
                    as1[z] = add(as1[z], as1[z + half]);
                    L[z] = add(L[z], L[z + half]);
                }
            const V r = scl(as1[0], iso);
            const V in = scl(L[0], iso);
            // MULTI_SCATTERING_POWER_SERIE is never defined, so the shader's `#if ... == 0` branch
            // is compiled: the series 1 + r + r^2 + r^3 + r^4 (genMultiScattering.comp:133-136)
            const V r2 = mul(r, r);
            const V series = add(add(add(add(v3(1, 1, 1), r), r2), mul(r, r2)), mul(r2, r2));
            const V Lt = mul(in, series);
            float* o = out + ((size_t)y * w + x) * 3;
            o[0] = Lt.x; o[1] = Lt.y; o[2] = Lt.z;
        }
}

void orc_sky_view(const orc_atmosphere* p, const float* trans, uint32_t tw, uint32_t th, const float* ms, uint32_t mw,
                  uint32_t mh, const float probe[3], const float sun_dir[3], const float illum[3], uint32_t w, uint32_t h,
                  float* out) {  // genView.comp
    Atmo A{p, {}, {}};
    A.trans = Lut{(int)tw, (int)th, trans};
    A.ms = Lut{(int)mw, (int)mh, ms};
    A.has_trans = A.has_ms = true;
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            V pos = add(load(probe), v3(0.0f, 0.0f, p->bottom_radius));
            const float u = (float)x / (float)w, v = (float)y / (float)h;
            const float vh = len(pos);
            float vz, lv;
            uv_to_view_params(u, v, (int)w, (int)h, vh, p->bottom_radius, vz, lv);
            const V up = scl(pos, 1.0f / vh);
            const float sz = dot(up, load(sun_dir));
            const V sun = normalize(v3(std::sqrt(1.0f - sz * sz), 0.0f, sz));
            pos = v3(0.0f, 0.0f, vh);
            const float vs = std::sqrt(1.0f - vz * vz);
            const V dir = v3(vs * lv, vs * std::sqrt(1.0f - lv * lv), vz);
            float* o = out + ((size_t)y * w + x) * 3;
            // moveToTopAtmosphere (sky.glsl:73-92)
            const float hh = len(pos);
            if (hh > p->top_radius) {
                const float tt = sphere_nearest(pos, dir, v3(0, 0, 0), p->top_radius);
                if (tt < 0.0f) { o[0] = o[1] = o[2] = 0.0f; continue; }
                pos = add(add(pos, scl(dir, tt)), scl(scl(pos, 1.0f / hh), -kPlanetRadiusOffset));
            }
            const Scatter s = integrate(A, pos, dir, sun, false, 30.0f, true, true, 9000000.0f, load(illum));
            o[0] = s.L.x; o[1] = s.L.y; o[2] = s.L.z;
        }
}

// skyColor (secondaryRays.comp:36-58) with getSunLuminance (sky.glsl:128-152).
// sun_disc = 0 leaves the sun disc out (when the light is sampled explicitly).
void orc_sky_state_radiance(const orc_sky_state* st, const float dir[3], float out[3]) {
    const orc_atmosphere* p = &st->atmo;
    Atmo A{p, {}, {}};
    A.trans = Lut{(int)st->tw, (int)st->th, st->trans};
    A.has_trans = true;
    const Lut sky{(int)st->vw, (int)st->vh, st->view};
    const V sd = load(st->sun_dir), E = load(st->illum);
    const V d = load(dir);
    const V pos = add(load(st->probe), v3(0.0f, 0.0f, p->bottom_radius));
    const V up = normalize(pos);
    const float vz = dot(d, up);
    const float vh = len(pos);
    const V side = normalize(cross(up, d));
    const V fwd = normalize(cross(side, up));
    float lx = dot(sd, fwd);
    const float ly = dot(sd, side);
    lx = lx / std::sqrt(lx * lx + ly * ly);
    const bool ground = sphere_nearest(pos, d, v3(0, 0, 0), p->bottom_radius) >= 0.0f;
    float u, v;
    view_params_to_uv(ground, vz, lx, (int)st->vw, (int)st->vh, vh, p->bottom_radius, u, v);
    V L = sky.sample(u, v);
    if (st->sun_disc) {
        const float sun_r = 0.5f * 0.505f * 3.14159f / 180.0f;
        if (dot(d, sd) > orc_tx::cos(sun_r) && !ground) {
            float uu, vu, us, vs2;
            trans_params_to_uv(p->bottom_radius, 1.0f, p->bottom_radius, p->top_radius, uu, vu);
            const V upv = scl(pos, 1.0f / vh);
            trans_params_to_uv(vh, dot(sd, upv), p->bottom_radius, p->top_radius, us, vs2);
            const float ang = orc_tx::acos(clampf(dot(d, sd), -1.0f, 1.0f));
            const float rr = ang / sun_r;
            const float limb = std::sqrt(clampf(1.0f - rr * rr, 0.0001f, 1.0f));
            const V space = dvd(E, A.trans.sample(uu, vu));
            const V sun = scl(mul(space, A.trans.sample(us, vs2)), limb);
            L = add(L, mul(sun, dvd(v3(120000.0f, 120000.0f, 120000.0f), E)));
        }
    }
    out[0] = L.x; out[1] = L.y; out[2] = L.z;
}

void orc_sky_radiance(const orc_atmosphere* p, const float* trans, uint32_t tw, uint32_t th, const float* view,
                      uint32_t vw, uint32_t vh_, const float probe[3], const float sun_dir[3], const float illum[3],
                      int sun_disc, const float* dirs, uint32_t n, float* out) {
    orc_sky_state st;
    st.atmo = *p;
    st.trans = trans; st.tw = tw; st.th = th;
    st.view = view; st.vw = vw; st.vh = vh_;
    for (int k = 0; k < 3; ++k) { st.probe[k] = probe[k]; st.sun_dir[k] = sun_dir[k]; st.illum[k] = illum[k]; }
    st.sun_disc = sun_disc;
    for (uint32_t i = 0; i < n; ++i) orc_sky_state_radiance(&st, dirs + (size_t)i * 3, out + (size_t)i * 3);
}

}  // extern "C"
