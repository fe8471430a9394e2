// oracle_post.cpp — scalar CPU restatement of the reference's post-processing (test
// infrastructure only):
//   edge-aware bilateral denoiser  src/gpu/denoise/bilateral.comp:23-75 (smartDeNoise + dither),
//                                  parameters Denoiser::BilateralParams (denoiser.ixx)
//   tonemappers + sRGB encode      src/gpu/tonemap/{linear,reinhard,hable,aces,uchimura,amd}.comp,
//                                  src/gpu/util.glsl:13-18, parameters tonemapper.ixx:19-55,
//                                  selection renderer.ixx:163-230
// Textures are sampled as a LinearClamp sampler does: bilinear between texel centres, clamped to
// the edge. exp and pow are the basic-operation versions of oracle_transc.hpp.
#include "oracle.h"
#include "oracle_core.hpp"
#include "oracle_transc.hpp"

#include <cmath>

namespace {

// Bilinear lookup at texel-space position (x, y) (texel centres at integers) of a W x H image
// with `ch` floats per texel, `nc` of them read.
void bilinear(const float* img, uint32_t w, uint32_t h, uint32_t ch, uint32_t nc, float x, float y, float* out) {
    const float fx = std::floor(x), fy = std::floor(y);
    const float ax = x - fx, ay = y - fy;
    auto cl = [](int v, int n) { return v < 0 ? 0 : (v >= n ? n - 1 : v); };
    const int x0 = cl((int)fx, (int)w), x1 = cl((int)fx + 1, (int)w);
    const int y0 = cl((int)fy, (int)h), y1 = cl((int)fy + 1, (int)h);
    for (uint32_t c = 0; c < nc; ++c) {
        const float t00 = img[((size_t)y0 * w + x0) * ch + c], t10 = img[((size_t)y0 * w + x1) * ch + c];
        const float t01 = img[((size_t)y1 * w + x0) * ch + c], t11 = img[((size_t)y1 * w + x1) * ch + c];
        const float a = t00 * (1.0f - ax) + t10 * ax;
        const float b = t01 * (1.0f - ax) + t11 * ax;
        out[c] = a * (1.0f - ay) + b * ay;
    }
}

float clampf(float x, float a, float b) { return std::fmin(std::fmax(x, a), b); }

float srgb(float c) { return c < 0.0031308f ? 12.92f * c : 1.055f * orc_tx::pow(c, 1.0f / 2.4f) - 0.055f; }

float hable_op(float x) {
    const float A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f;
    return ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F;
}

float rrt_odt(float v) {
    const float a = v * (v + 0.0245786f) - 0.000090537f;
    const float b = v * (0.983729f * v + 0.4329510f) + 0.238081f;
    return a / b;
}

float col_tone_b(float hdr, float con, float sh, float mi, float mo) {
    const float num = -orc_tx::pow(mi, con) +
                      (mo * (orc_tx::pow(hdr, con * sh) * orc_tx::pow(mi, con) -
                             orc_tx::pow(hdr, con) * orc_tx::pow(mi, con * sh) * mo)) /
                          (orc_tx::pow(hdr, con * sh) * mo - orc_tx::pow(mi, con * sh) * mo);
    return -(num / (orc_tx::pow(mi, con * sh) * mo));
}
float col_tone_c(float hdr, float con, float sh, float mi, float mo) {
    return (orc_tx::pow(hdr, con * sh) * orc_tx::pow(mi, con) - orc_tx::pow(hdr, con) * orc_tx::pow(mi, con * sh) * mo) /
           (orc_tx::pow(hdr, con * sh) * mo - orc_tx::pow(mi, con * sh) * mo);
}
float col_tone(float x, float con, float sh, float b, float c) {
    const float z = orc_tx::pow(x, con);
    return z / (orc_tx::pow(z, sh) * b + c);
}

void tonemap_one(const orc_tonemap_params* p, const float in[3], float out[3]) {
    float v[3] = {in[0] * p->exposure, in[1] * p->exposure, in[2] * p->exposure};
    float m[3];
    switch (p->op) {
        case 0:  // linear.comp
            for (int k = 0; k < 3; ++k) m[k] = v[k];
            break;
        case 1: {  // reinhard.comp
            const float mw = p->reinhard_max * p->reinhard_max;
            for (int k = 0; k < 3; ++k) m[k] = (v[k] * (1.0f + (v[k] / mw))) / (1.0f + v[k]);
            break;
        }
        case 2: {  // hable.comp
            const float wd = hable_op(11.2f);
            for (int k = 0; k < 3; ++k) m[k] = hable_op(2.0f * v[k]) / wd;
            break;
        }
        case 3: {  // aces.comp: v * ACESInputMat (row vector times the columns listed), fit, v * ACESOutputMat
            const float i0 = (0.59719f * v[0] + 0.35458f * v[1]) + 0.04823f * v[2];
            const float i1 = (0.07600f * v[0] + 0.90834f * v[1]) + 0.01566f * v[2];
            const float i2 = (0.02840f * v[0] + 0.13383f * v[1]) + 0.83777f * v[2];
            const float f0 = rrt_odt(i0), f1 = rrt_odt(i1), f2 = rrt_odt(i2);
            m[0] = clampf((1.60475f * f0 + -0.53108f * f1) + -0.07367f * f2, 0.0f, 1.0f);
            m[1] = clampf((-0.10208f * f0 + 1.10813f * f1) + -0.00605f * f2, 0.0f, 1.0f);
            m[2] = clampf((-0.00327f * f0 + -0.07276f * f1) + 1.07602f * f2, 0.0f, 1.0f);
            break;
        }
        case 4: {  // uchimura.comp
            const float P = p->uchimura[0], a = p->uchimura[1], mm = p->uchimura[2], l = p->uchimura[3];
            const float c = p->uchimura[4], b = p->uchimura[5];
            const float l0 = ((P - mm) * l) / a;
            const float S0 = mm + l0, S1 = mm + a * l0;
            const float C2 = (a * P) / (P - S1);
            const float CP = -C2 / P;
            for (int k = 0; k < 3; ++k) {
                const float x = v[k];
                const float t = clampf(x / mm, 0.0f, 1.0f);  // smoothstep(0, m, x)
                const float w0 = 1.0f - t * t * (3.0f - 2.0f * t);
                const float w2 = x < mm + l0 ? 0.0f : 1.0f;  // step(m + l0, x)
                const float w1 = 1.0f - w0 - w2;
                const float T = mm * orc_tx::pow(x / mm, c) + b;
                const float S = P - (P - S1) * orc_tx::exp(CP * (x - S0));
                const float L = mm + a * (x - mm);
                m[k] = (T * w0 + L * w1) + S * w2;
            }
            break;
        }
        default: {  // amd.comp (AMDTonemapper)
            const float hdr = p->amd[0], con = p->amd[1], sh = p->amd[2], mi = p->amd[3], mo = p->amd[4];
            const float b = col_tone_b(hdr, con, sh, mi, mo), c = col_tone_c(hdr, con, sh, mi, mo);
            float peak = std::fmax(v[0], std::fmax(v[1], v[2]));
            peak = std::fmax(1e-6f, peak);
            float ratio[3] = {v[0] / peak, v[1] / peak, v[2] / peak};
            peak = col_tone(peak, con, sh, b, c);
            const float sat = con, cross = con * 16.0f;
            const float tw = orc_tx::pow(peak, 4.0f);
            for (int k = 0; k < 3; ++k) {
                float r = orc_tx::pow(std::fabs(ratio[k]), sat / cross);
                r = r * (1.0f - tw) + 1.0f * tw;  // mix(ratio, white, peak^crosstalk)
                r = orc_tx::pow(std::fabs(r), cross);
                m[k] = peak * r;
            }
            break;
        }
    }
    for (int k = 0; k < 3; ++k) out[k] = srgb(m[k]);
}

}  // namespace

extern "C" {

void orc_tonemap_defaults(orc_tonemap_params* p) {  // renderer.ixx:183-187, tonemapper.ixx:27-55
    p->op = 5;
    p->exposure = 1.0f;
    p->reinhard_max = 8.0f;
    const float u[6] = {1.0f, 1.0f, 0.22f, 0.4f, 1.33f, 0.0f};
    const float a[5] = {16.0f, 2.0f, 1.0f, 0.18f, 0.18f};
    for (int k = 0; k < 6; ++k) p->uchimura[k] = u[k];
    for (int k = 0; k < 5; ++k) p->amd[k] = a[k];
}

void orc_tonemap(const orc_tonemap_params* p, const float* rgb, uint32_t n, float* out, uint8_t* rgba8) {
    for (uint32_t i = 0; i < n; ++i) {
        float o[3];
        tonemap_one(p, rgb + (size_t)i * 3, o);
        if (out) for (int k = 0; k < 3; ++k) out[(size_t)i * 3 + k] = o[k];
        if (rgba8) {
            for (int k = 0; k < 3; ++k) rgba8[(size_t)i * 4 + k] = (uint8_t)std::floor(clampf(o[k], 0.0f, 1.0f) * 255.0f + 0.5f);
            rgba8[(size_t)i * 4 + 3] = 255;
        }
    }
}

// smartDeNoise (bilateral.comp:23-61) + dither (:63-75). A pixel whose depth is <= 0 (no primary
// hit: reverse-Z depth 0) is returned unfiltered, extending the shader's `centrZ < 0` test to the
// depth its own G-buffer writes for the sky.
void orc_denoise_bilateral(const float* rgb, const float* depth, const float* normal, uint32_t w, uint32_t h,
                           const orc_bilateral* p, float* out) {
    const float radius = std::round(p->k_sigma * p->sigma);
    const float radq = radius * radius;
    const float inv_s2 = 0.5f / (p->sigma * p->sigma);
    const float inv_s2pi = 0.31830988618379067154f * inv_s2;
    const float inv_t2 = 0.5f / (p->threshold * p->threshold);
    const float inv_t_sqrt2pi = 0.39894228040143267794f / p->threshold;
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            const size_t i = (size_t)y * w + x;
            float* o = out + i * 3;
            const float cz = depth[i];
            float res[3];
            if (!(cz > 0.0f)) {
                for (int k = 0; k < 3; ++k) res[k] = rgb[i * 3 + k];
            } else {
                const float cn[3] = {normal[i * 4], normal[i * 4 + 1], normal[i * 4 + 2]};
                float zb = 0.0f, ab[3] = {0.0f, 0.0f, 0.0f};
                for (float dx = -radius; dx <= radius; dx += 1.0f) {
                    const float pt = std::sqrt(radq - dx * dx);
                    for (float dy = -pt; dy <= pt; dy += 1.0f) {
                        const float blur = orc_tx::exp(-(dx * dx + dy * dy) * inv_s2) * inv_s2pi;
                        const float sx = (float)x + dx, sy = (float)y + dy;
                        float px[3], wz, wn[3];
                        bilinear(rgb, w, h, 3, 3, sx, sy, px);
                        bilinear(depth, w, h, 1, 1, sx, sy, &wz);
                        bilinear(normal, w, h, 4, 3, sx, sy, wn);
                        float dz = p->near_plane / wz - p->near_plane / cz;
                        dz *= 100.0f;
                        const float dn = (wn[0] * cn[0] + wn[1] * cn[1]) + wn[2] * cn[2];
                        const float delta = orc_tx::exp(clampf(dn - dz * dz, 0.0f, 1.0f) * inv_t2) * inv_t_sqrt2pi * blur;
                        zb += delta;
                        for (int k = 0; k < 3; ++k) ab[k] += delta * px[k];
                    }
                }
                for (int k = 0; k < 3; ++k) res[k] = ab[k] / zb;
            }
            uint32_t seed = x * 709u + y * 1153u + p->frame * 1361u;
            const float noise = orc::rand_float(seed) * 0.005f;
            for (int k = 0; k < 3; ++k) o[k] = res[k] + noise;
        }
}

}  // extern "C"
