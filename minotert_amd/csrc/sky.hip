// sky.hip — the reference's physical sky on gfx950: the atmosphere LUTs of Hillaire 2020 and the
// miss radiance of escaped paths.
//
//   transmittance 256x64       src/gpu/sky/genTransmittance.comp:22-43   (k_sky_transmittance)
//   multiple scattering 32x32  src/gpu/sky/genMultiScattering.comp:25-145 (k_sky_multiscattering:
//                              one workgroup per texel, 64 directions, LDS tree reduction in the
//                              shader's order)
//   sky view 192x108           src/gpu/sky/genView.comp:33-76            (k_sky_view)
//   miss radiance              src/gpu/secondaryRays.comp:36-58 skyColor + sky.glsl:128-152
//                              (sky_radiance, used by the path tracer on a miss)
// Shared functions: src/gpu/sky/sky.glsl (medium :94-123, phases :38-50, ray/sphere :52-71,
// integration :171-326) and skyAccess.glsl (LUT parameterizations :11-117). LUTs are FP32
// float4 texels here (the reference uses RGBA16F / B10G11R11), sampled bilinearly with
// clamp-to-edge as a LinearClamp sampler does. The power series of the multiple-scattering LUT
// is the shader's compiled branch: MULTI_SCATTERING_POWER_SERIE is never defined, so
// `#if ... == 0` selects 1 + r + r^2 + r^3 + r^4.
#include "internal.h"
#include "sky.h"

#include <cstring>
#include <new>

namespace sky {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kOffset = 0.01f;  // PLANET_RADIUS_OFFSET

struct Medium { float3 scat, ext, scat_mie, scat_ray; };

__device__ __forceinline__ Medium medium(const mrt_atmosphere_params& p, float3 pos) {
    const float h = len(pos) - p.bottom_radius;
    const float dmie = rt::tx_exp(p.mie_density_exp_scale * h);
    const float dray = rt::tx_exp(p.rayleigh_density_exp_scale * h);
    const float dozo = clampf(h < p.absorption_density0_layer_width
                                  ? p.absorption_density0_linear * h + p.absorption_density0_constant
                                  : p.absorption_density1_linear * h + p.absorption_density1_constant,
                              0.0f, 1.0f);
    Medium m;
    m.scat_mie = ld3(p.mie_scattering) * dmie;
    const float3 ext_mie = ld3(p.mie_extinction) * dmie;
    m.scat_ray = ld3(p.rayleigh_scattering) * dray;
    const float3 ext_ozo = ld3(p.absorption_extinction) * dozo;
    m.scat = m.scat_mie + m.scat_ray;
    m.ext = (ext_mie + m.scat_ray) + ext_ozo;
    return m;
}

__device__ __forceinline__ float mie_phase(float g, float c) {
    const float k = 3.0f / (8.0f * kPi) * (1.0f - g * g) / (2.0f + g * g);
    return k * (1.0f + c * c) / rt::tx_pow15(1.0f + g * g - 2.0f * g * -c);
}
__device__ __forceinline__ float rayleigh_phase(float c) { return 3.0f / (16.0f * kPi) * (1.0f + c * c); }

__device__ __forceinline__ float3 sample_ms(const mrt_atmosphere_params& p, const float4* ms, int w, int h, float3 pos,
                                            float vz) {
    float u = clampf(vz * 0.5f + 0.5f, 0.0f, 1.0f);
    float v = clampf((len(pos) - p.bottom_radius) / (p.top_radius - p.bottom_radius), 0.0f, 1.0f);
    return lut(ms, w, h, unit_to_sub(u, (float)w), unit_to_sub(v, (float)h));
}

struct Scatter { float3 L, od, as1; };

// integrateScatteredLuminance (sky.glsl:171-326) without the unused newMultiScatStep outputs
__device__ Scatter integrate(const mrt_atmosphere_params& p, const float4* trans, int tw, int th, const float4* ms,
                             int mw, int mh, float3 pos, float3 dir, float3 sun, bool ground, float spp_ini,
                             bool variable, bool mie_ray_phase, float tmax_max, float3 illum) {
    Scatter r;
    r.L = r.od = r.as1 = make_float3(0.0f, 0.0f, 0.0f);
    const float3 O = make_float3(0.0f, 0.0f, 0.0f);
    const float t_bottom = sphere_nearest(pos, dir, O, p.bottom_radius);
    const float t_top = sphere_nearest(pos, dir, O, p.top_radius);
    float tmax = 0.0f;
    if (t_bottom < 0.0f) {
        if (t_top < 0.0f) return r;
        tmax = t_top;
    } else if (t_top > 0.0f) {
        tmax = fminf(t_top, t_bottom);
    }
    tmax = fminf(tmax, tmax_max);
    float spp = spp_ini, spp_floor = spp_ini, tmax_floor = tmax;
    if (variable) {
        const float a = clampf(tmax * 0.01f, 0.0f, 1.0f);
        spp = 4.0f * (1.0f - a) + 14.0f * a;
        spp_floor = floorf(spp);
        tmax_floor = tmax * spp_floor / spp;
    }
    float dt = tmax / spp;
    const float uniform_phase = 1.0f / (4.0f * kPi);
    const float cos_t = dot(sun, dir);
    const float mie_p = mie_phase(p.mie_phase_g, -cos_t);
    const float ray_p = rayleigh_phase(cos_t);
    float3 L = O, thr = make_float3(1.0f, 1.0f, 1.0f), od = O;
    float t = 0.0f;
    for (float s = 0.0f; s < spp; s += 1.0f) {
        if (variable) {
            float t0 = s / spp_floor, t1 = (s + 1.0f) / spp_floor;
            t0 = t0 * t0;
            t1 = t1 * t1;
            t0 = tmax_floor * t0;
            t1 = t1 > 1.0f ? tmax : tmax_floor * t1;
            t = t0 + (t1 - t0) * 0.3f;
            dt = t1 - t0;
        } else {
            const float nt = tmax * (s + 0.3f) / spp;
            dt = nt - t;
            t = nt;
        }
        const float3 P = pos + dir * t;
        const Medium m = medium(p, P);
        const float3 sod = m.ext * dt;
        const float3 st = make_float3(rt::tx_exp(-sod.x), rt::tx_exp(-sod.y), rt::tx_exp(-sod.z));
        od = od + sod;
        const float ph = len(P);
        const float3 up = P * (1.0f / ph);
        const float sun_z = dot(sun, up);
        float tu, tv;
        trans_params_to_uv(ph, sun_z, p.bottom_radius, p.top_radius, tu, tv);
        const float3 t_sun = trans ? lut(trans, tw, th, tu, tv) : O;
        const float3 pts = mie_ray_phase ? m.scat_mie * mie_p + m.scat_ray * ray_p : m.scat * uniform_phase;
        const float t_earth = sphere_nearest(P, sun, O + up * kOffset, p.bottom_radius);
        const float shadow = t_earth >= 0.0f ? 0.0f : 1.0f;
        const float3 msl = ms ? sample_ms(p, ms, mw, mh, P, sun_z) : O;
        const float3 S = illum * ((t_sun * pts) * shadow + msl * m.scat);
        r.as1 = r.as1 + thr * ((m.scat - m.scat * st) / m.ext);
        L = L + thr * ((S - S * st) / m.ext);
        thr = thr * st;
    }
    if (ground && tmax == t_bottom && t_bottom > 0.0f) {
        const float3 P = pos + dir * t_bottom;
        const float ph = len(P);
        const float3 up = P * (1.0f / ph);
        float tu, tv;
        trans_params_to_uv(ph, dot(sun, up), p.bottom_radius, p.top_radius, tu, tv);
        const float3 t_sun = trans ? lut(trans, tw, th, tu, tv) : O;
        const float ndl = clampf(dot(normalize(up), normalize(sun)), 0.0f, 1.0f);
        L = L + ((illum * t_sun) * thr) * ld3(p.ground_albedo) * (ndl / kPi);
    }
    r.L = L;
    r.od = od;
    return r;
}

__global__ __launch_bounds__(64) void k_sky_transmittance(mrt_atmosphere_params p, int w, int h, float4* __restrict__ out) {
    const int x = blockIdx.x * 8 + (threadIdx.x & 7), y = blockIdx.y * 8 + (threadIdx.x >> 3);
    if (x >= w || y >= h) return;
    const float u = ((float)x + 0.5f) / (float)w, v = ((float)y + 0.5f) / (float)h;
    float vh, mu;
    uv_to_trans_params(u, v, p.bottom_radius, p.top_radius, vh, mu);
    const Scatter s = integrate(p, nullptr, 0, 0, nullptr, 0, 0, make_float3(0.0f, 0.0f, vh),
                                make_float3(0.0f, sqrtf(1.0f - mu * mu), mu), make_float3(1.0f, 1.0f, 1.0f), false, 40.0f,
                                false, false, 9000000.0f, make_float3(1.0f, 1.0f, 1.0f));
    out[(size_t)y * w + x] = make_float4(rt::tx_exp(-s.od.x), rt::tx_exp(-s.od.y), rt::tx_exp(-s.od.z), 1.0f);
}

// One workgroup of 64 per texel (the shader's local size 1x1x64).
__global__ __launch_bounds__(64) void k_sky_multiscattering(mrt_atmosphere_params p, const float4* __restrict__ trans,
                                                            int tw, int th, int w, int h, float4* __restrict__ out) {
    __shared__ float3 s_as1[64], s_L[64];
    const int x = blockIdx.x, y = blockIdx.y, z = threadIdx.x;
    float u = ((float)x + 0.5f) / (float)w, v = ((float)y + 0.5f) / (float)h;
    u = sub_to_unit(u, (float)w);
    v = sub_to_unit(v, (float)h);
    const float cs = u * 2.0f - 1.0f;
    const float3 sun = make_float3(0.0f, sqrtf(clampf(1.0f - cs * cs, 0.0f, 1.0f)), cs);
    const float vh = p.bottom_radius + clampf(v + kOffset, 0.0f, 1.0f) * (p.top_radius - p.bottom_radius - kOffset);
    const float sphere = 4.0f * kPi, sq = 8.0f;
    const float i = 0.5f + (float)(z / 8), j = 0.5f + (float)(z - (z / 8) * 8);
    const float theta = 2.0f * kPi * (i / sq), phi = kPi * (j / sq);
    const float3 dir = make_float3(rt::tx_cos(theta) * rt::tx_sin(phi), rt::tx_sin(theta) * rt::tx_sin(phi), rt::tx_cos(phi));
    const Scatter s = integrate(p, trans, tw, th, nullptr, 0, 0, make_float3(0.0f, 0.0f, vh), dir, sun, true, 20.0f, false,
                                false, 9000000.0f, make_float3(1.0f, 1.0f, 1.0f));
    s_as1[z] = s.as1 * (sphere / (sq * sq));
    s_L[z] = s.L * (sphere / (sq * sq));
    __syncthreads();
    for (int half = 32; half >= 1; half >>= 1) {
        if (z < half) {
            s_as1[z] = s_as1[z] + s_as1[z + half];
            s_L[z] = s_L[z] + s_L[z + half];
        }
        __syncthreads();
    }
    if (z > 0) return;
    const float iso = 1.0f / sphere;
    const float3 r = s_as1[0] * iso;
    const float3 in = s_L[0] * iso;
    const float3 r2 = r * r;
    const float3 series = (((make_float3(1.0f, 1.0f, 1.0f) + r) + r2) + r * r2) + r2 * r2;
    const float3 L = in * series;
    out[(size_t)y * w + x] = make_float4(L.x, L.y, L.z, 1.0f);
}

__global__ __launch_bounds__(64) void k_sky_view(mrt_atmosphere_params p, const float4* __restrict__ trans, int tw, int th,
                                                 const float4* __restrict__ ms, int mw, int mh, float3 probe, float3 sun_dir,
                                                 float3 illum, int w, int h, float4* __restrict__ out) {
    const int x = blockIdx.x * 8 + (threadIdx.x & 7), y = blockIdx.y * 8 + (threadIdx.x >> 3);
    if (x >= w || y >= h) return;
    float3 pos = probe + make_float3(0.0f, 0.0f, p.bottom_radius);
    const float u = (float)x / (float)w, v = (float)y / (float)h;
    const float vh = len(pos);
    float vz, lv;
    uv_to_view_params(u, v, w, h, vh, p.bottom_radius, vz, lv);
    const float3 up = pos * (1.0f / vh);
    const float sz = dot(up, sun_dir);
    const float3 sun = normalize(make_float3(sqrtf(1.0f - sz * sz), 0.0f, sz));
    pos = make_float3(0.0f, 0.0f, vh);
    const float vs = sqrtf(1.0f - vz * vz);
    const float3 dir = make_float3(vs * lv, vs * sqrtf(1.0f - lv * lv), vz);
    const float hh = len(pos);
    if (hh > p.top_radius) {  // moveToTopAtmosphere (sky.glsl:73-92)
        const float tt = sphere_nearest(pos, dir, make_float3(0.0f, 0.0f, 0.0f), p.top_radius);
        if (tt < 0.0f) { out[(size_t)y * w + x] = make_float4(0.0f, 0.0f, 0.0f, 1.0f); return; }
        pos = (pos + dir * tt) + (pos * (1.0f / hh)) * -kOffset;
    }
    const Scatter s = integrate(p, trans, tw, th, ms, mw, mh, pos, dir, sun, false, 30.0f, true, true, 9000000.0f, illum);
    out[(size_t)y * w + x] = make_float4(s.L.x, s.L.y, s.L.z, 1.0f);
}

__global__ __launch_bounds__(256) void k_sky_radiance(SkyDev sd, const float* __restrict__ dirs, uint32_t n,
                                                      float* __restrict__ rgb) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const float3 L = radiance(sd, make_float3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
    rgb[3 * i] = L.x;
    rgb[3 * i + 1] = L.y;
    rgb[3 * i + 2] = L.z;
}

}  // namespace sky

struct mrt_sky {
    mrt_atmosphere_params p{};
    float4* trans = nullptr;
    float4* ms = nullptr;
    float4* view = nullptr;
    SkyDev dev{};
    bool has_view = false;
};

bool sky_device(const mrt_sky* s, SkyDev* out) {
    if (!s || !s->has_view) return false;
    *out = s->dev;
    return true;
}

extern "C" {

mrt_status mrt_sky_earth(mrt_atmosphere_params* p) {
    if (!p) return api_fail(nullptr, MRT_ERR_INVALID, "mrt_sky_earth: null argument");
    std::memset(p, 0, sizeof(*p));
    p->bottom_radius = 6360.0f;
    p->top_radius = 6460.0f;
    p->rayleigh_density_exp_scale = -1.0f / 8.0f;
    const float ray[3] = {0.005802f, 0.013558f, 0.033100f};
    const float ozo[3] = {0.000650f, 0.001881f, 0.000085f};
    for (int k = 0; k < 3; ++k) {
        p->rayleigh_scattering[k] = ray[k];
        p->mie_scattering[k] = 0.003996f;
        p->mie_extinction[k] = 0.004440f;
        p->mie_absorption[k] = fmaxf(0.004440f - 0.003996f, 0.0f);
        p->absorption_extinction[k] = ozo[k];
    }
    p->mie_density_exp_scale = -1.0f / 1.2f;
    p->mie_phase_g = 0.8f;
    p->absorption_density0_layer_width = 25.0f;
    p->absorption_density0_constant = -2.0f / 3.0f;
    p->absorption_density0_linear = 1.0f / 15.0f;
    p->absorption_density1_constant = 8.0f / 3.0f;
    p->absorption_density1_linear = -1.0f / 15.0f;
    return MRT_OK;
}

mrt_status mrt_sky_create(mrt_ctx* ctx, const mrt_atmosphere_params* p, mrt_sky** out) {
    if (!ctx || !p || !out) return api_fail(ctx, MRT_ERR_INVALID, "mrt_sky_create: null argument");
    *out = nullptr;
    if (!(p->top_radius > p->bottom_radius && p->bottom_radius > 0.0f))
        return api_fail(ctx, MRT_ERR_INVALID, "mrt_sky_create: need 0 < bottom_radius < top_radius");
    auto* s = new (std::nothrow) mrt_sky();
    if (!s) return api_fail(ctx, MRT_ERR_OOM, "mrt_sky_create: host allocation failed");
    s->p = *p;
    hipError_t e = hipMalloc((void**)&s->trans, (size_t)kSkyTransW * kSkyTransH * 16);
    if (e == hipSuccess) e = hipMalloc((void**)&s->ms, (size_t)kSkyMsW * kSkyMsH * 16);
    if (e == hipSuccess) e = hipMalloc((void**)&s->view, (size_t)kSkyViewW * kSkyViewH * 16);
    if (e == hipSuccess) {
        ProfScope pr(ctx, "sky_luts");
        hipLaunchKernelGGL(sky::k_sky_transmittance, dim3((kSkyTransW + 7) / 8, (kSkyTransH + 7) / 8), dim3(64), 0,
                           ctx->stream, s->p, kSkyTransW, kSkyTransH, s->trans);
        hipLaunchKernelGGL(sky::k_sky_multiscattering, dim3(kSkyMsW, kSkyMsH), dim3(64), 0, ctx->stream, s->p, s->trans,
                           kSkyTransW, kSkyTransH, kSkyMsW, kSkyMsH, s->ms);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        hipFree(s->trans); hipFree(s->ms); hipFree(s->view);
        delete s;
        return api_fail(ctx, e == hipErrorOutOfMemory ? MRT_ERR_OOM : MRT_ERR_HIP, "mrt_sky_create: %s", hipGetErrorString(e));
    }
    *out = s;
    return MRT_OK;
}

mrt_status mrt_sky_update_view(mrt_ctx* ctx, mrt_sky* s, const float probe_km[3], const float sun_dir[3],
                               const float sun_illuminance[3], int sun_disc) {
    if (!ctx || !s || !probe_km || !sun_dir || !sun_illuminance)
        return api_fail(ctx, MRT_ERR_INVALID, "mrt_sky_update_view: null argument");
    const float3 probe = make_float3(probe_km[0], probe_km[1], probe_km[2]);
    const float3 sd = make_float3(sun_dir[0], sun_dir[1], sun_dir[2]);
    const float3 il = make_float3(sun_illuminance[0], sun_illuminance[1], sun_illuminance[2]);
    if (!(sun_illuminance[0] > 0.0f && sun_illuminance[1] > 0.0f && sun_illuminance[2] > 0.0f))
        return api_fail(ctx, MRT_ERR_INVALID, "mrt_sky_update_view: sun illuminance must be positive");
    {
        ProfScope pr(ctx, "sky_view");
        hipLaunchKernelGGL(sky::k_sky_view, dim3((kSkyViewW + 7) / 8, (kSkyViewH + 7) / 8), dim3(64), 0, ctx->stream, s->p,
                           s->trans, kSkyTransW, kSkyTransH, s->ms, kSkyMsW, kSkyMsH, probe, sd, il, kSkyViewW, kSkyViewH,
                           s->view);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return api_fail(ctx, MRT_ERR_HIP, "mrt_sky_update_view: %s", hipGetErrorString(e));
    SkyDev& d = s->dev;
    d.p = s->p;
    d.trans = s->trans;
    d.view = s->view;
    d.probe = probe;
    d.sun = sd;
    d.illum = il;
    d.sun_disc = sun_disc ? 1 : 0;
    s->has_view = true;
    return MRT_OK;
}

mrt_status mrt_sky_radiance(mrt_ctx* ctx, const mrt_sky* s, const float* dirs, uint32_t n, float* rgb) {
    if (!ctx || !s) return api_fail(ctx, MRT_ERR_INVALID, "mrt_sky_radiance: null argument");
    if (!s->has_view) return api_fail(ctx, MRT_ERR_STATE, "mrt_sky_radiance: no sky view yet (mrt_sky_update_view)");
    if (n == 0) return MRT_OK;
    if (!dirs || !rgb) return api_fail(ctx, MRT_ERR_INVALID, "mrt_sky_radiance: null buffer");
    hipLaunchKernelGGL(sky::k_sky_radiance, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, s->dev, dirs, n, rgb);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return api_fail(ctx, MRT_ERR_HIP, "mrt_sky_radiance: %s", hipGetErrorString(e));
    return MRT_OK;
}

// which: 0 transmittance (256x64), 1 multiple scattering (32x32), 2 sky view (192x108); 4 floats per texel
mrt_status mrt_sky_get(mrt_ctx* ctx, const mrt_sky* s, uint32_t which, float* host, size_t count) {
    if (!ctx || !s || !host) return api_fail(ctx, MRT_ERR_INVALID, "mrt_sky_get: null argument");
    const float4* src = which == 0 ? s->trans : which == 1 ? s->ms : which == 2 ? s->view : nullptr;
    const size_t texels = which == 0 ? (size_t)kSkyTransW * kSkyTransH : which == 1 ? (size_t)kSkyMsW * kSkyMsH
                                                                       : (size_t)kSkyViewW * kSkyViewH;
    if (!src || count != texels * 4) return api_fail(ctx, MRT_ERR_INVALID, "mrt_sky_get: which %u count %zu", which, count);
    hipError_t e = hipMemcpyAsync(host, src, count * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return api_fail(ctx, MRT_ERR_HIP, "mrt_sky_get: %s", hipGetErrorString(e));
    return MRT_OK;
}

mrt_status mrt_sky_destroy(mrt_ctx* ctx, mrt_sky* s) {
    if (!ctx) return api_fail(nullptr, MRT_ERR_INVALID, "mrt_sky_destroy: null context");
    if (!s) return MRT_OK;
    hipStreamSynchronize(ctx->stream);
    hipFree(s->trans); hipFree(s->ms); hipFree(s->view);
    delete s;
    return MRT_OK;
}

}  // extern "C"
