// post.hip — the reference's post-processing on gfx950:
//   edge-aware bilateral denoiser  src/gpu/denoise/bilateral.comp:23-75 (smartDeNoise + dither;
//                                  Denoiser::bilateral, denoiser.ixx:34-97)
//   tonemappers + sRGB encode      src/gpu/tonemap/{linear,reinhard,hable,aces,uchimura,amd}.comp
//                                  and src/gpu/util.glsl:13-18 (Tonemapper, tonemapper.ixx:57-373;
//                                  the AMD operator is the renderer's default, renderer.ixx:184)
// The denoiser samples colour, depth and normal bilinearly between texel centres, clamped to the
// edge (the reference's LinearClamp sampler) at (x + dx, y + dy): dx runs over integers and dy
// from -sqrt(r^2 - dx^2) in steps of one, so vertical taps fall between rows. A 16x16 block
// stages its neighbourhood of the three images (two float4 texels: colour + depth, normal) and the
// tap weights in LDS for radius r <= 12, else it reads global memory; both run the same per-pixel
// code. Output stays FP32 (the
// reference writes RGBA8 UNORM, which clamps HDR before tonemapping). exp and pow are transc.h's.
#include "internal.h"
#include "transc.h"

namespace post {

constexpr int kTile = 16;
constexpr int kMaxLdsRadius = 12;  // 2 float4 planes of 42^2 texels + 25^2 blur entries: 58 KB of LDS

struct DenoiseArgs {
    const float* rgb;     // W*H*3
    const float* depth;   // W*H
    const float* normal;  // W*H*4
    float* out;           // W*H*3
    int w, h;
    float radius, radq, inv_s2, inv_s2pi, inv_t2, inv_t_sqrt2pi, near_plane;
    uint32_t frame;
};

__device__ __forceinline__ float clampf(float x, float a, float b) { return fminf(fmaxf(x, a), b); }

// The taps. dx runs over the integers -R..R and pixel x is an integer, so every sample lies on
// a texel column: the horizontal weight is exactly 0 and the bilinear lookup reduces to the blend
// of two texels of one column (for finite texels t1, t0 * 1 + t1 * 0 == t0 up to the sign of a
// zero, which no later operation can observe). dy runs from -sqrt(R^2 - dx^2) in steps of one;
// the sequence is the same for every pixel, so the spatial weight
// blur = exp(-(dx^2 + dy^2) / (2 sigma^2)) / (2 pi sigma^2) of tap k of column c is tabled in LDS
// at c * (2R + 1) + k, filled by one thread per column with the same code.
__device__ __forceinline__ float blur_weight(const DenoiseArgs& a, float dx, float dy) {
    return rt::tx_exp(-(dx * dx + dy * dy) * a.inv_s2) * a.inv_s2pi;
}

__device__ void fill_blur_table(const DenoiseArgs& a, int r, float* s_blur) {
    const int c = threadIdx.x;
    if (c <= 2 * r) {
        const float dx = -a.radius + (float)c;
        const float pt = sqrtf(a.radq - dx * dx);
        int k = 0;
        for (float dy = -pt; dy <= pt; dy += 1.0f, ++k) s_blur[c * (2 * r + 1) + k] = blur_weight(a, dx, dy);
    }
}

// Texel sources: global memory with edge clamping, or an LDS tile of clamped texels.
struct GlobalSrc {
    const DenoiseArgs* a;
    int x;  // the tap column, clamped
    __device__ void texel(int y, float c[7]) const {
        y = min(max(y, 0), a->h - 1);
        const size_t i = (size_t)y * a->w + x;
        c[0] = a->rgb[i * 3]; c[1] = a->rgb[i * 3 + 1]; c[2] = a->rgb[i * 3 + 2];
        c[3] = a->depth[i];
        c[4] = a->normal[i * 4]; c[5] = a->normal[i * 4 + 1]; c[6] = a->normal[i * 4 + 2];
    }
};
struct LdsSrc {
    const float4* cd;  // t*t texels: r, g, b, depth
    const float4* nm;  // t*t texels: nx, ny, nz, 0
    int t, col;        // tile width, tile column of the tap
    int oy;
    __device__ void texel(int y, float c[7]) const {
        const int i = __mul24(y - oy, t) + col;
        const float4 p = cd[i], q = nm[i];
        c[0] = p.x; c[1] = p.y; c[2] = p.z; c[3] = p.w;
        c[4] = q.x; c[5] = q.y; c[6] = q.z;
    }
};

// smartDeNoise (bilateral.comp:23-61) for pixel (x, y) + the dither (:63-75). A pixel whose
// depth is <= 0 (no primary hit) is returned unfiltered (the shader's `centrZ < 0` test extended
// to the depth the G-buffer writes for the sky). MakeSrc(dx) gives the texel source of a column.
template <bool TABLE, class MakeSrc>
__device__ void denoise_pixel(const DenoiseArgs& a, const float* s_blur, int r, MakeSrc make_src, int x, int y) {
    const size_t i = (size_t)y * a.w + x;
    const float cz = a.depth[i];
    float res[3];
    if (!(cz > 0.0f)) {
        res[0] = a.rgb[i * 3]; res[1] = a.rgb[i * 3 + 1]; res[2] = a.rgb[i * 3 + 2];
    } else {
        const float cn0 = a.normal[i * 4], cn1 = a.normal[i * 4 + 1], cn2 = a.normal[i * 4 + 2];
        const float near_c = a.near_plane / cz;
        float zb = 0.0f, ab0 = 0.0f, ab1 = 0.0f, ab2 = 0.0f;
        int c = 0;
        for (float dx = -a.radius; dx <= a.radius; dx += 1.0f, ++c) {
            const float pt = sqrtf(a.radq - dx * dx);
            const auto src = make_src(x + (int)dx);
            int k = 0;
            for (float dy = -pt; dy <= pt; dy += 1.0f, ++k) {
                const float blur = TABLE ? s_blur[c * (2 * r + 1) + k] : blur_weight(a, dx, dy);
                const float sy = (float)y + dy;
                const float fy = floorf(sy);
                const float ay = sy - fy;
                const int y0 = (int)fy;
                float t0[7], t1[7], s[7];
                src.texel(y0, t0);
                src.texel(y0 + 1, t1);
#pragma unroll
                for (int q = 0; q < 7; ++q) s[q] = t0[q] * (1.0f - ay) + t1[q] * ay;
                float dz = a.near_plane / s[3] - near_c;
                dz *= 100.0f;
                const float dn = (s[4] * cn0 + s[5] * cn1) + s[6] * cn2;
                const float delta = rt::tx_exp(clampf(dn - dz * dz, 0.0f, 1.0f) * a.inv_t2) * a.inv_t_sqrt2pi * blur;
                zb += delta;
                ab0 += delta * s[0];
                ab1 += delta * s[1];
                ab2 += delta * s[2];
            }
        }
        res[0] = ab0 / zb; res[1] = ab1 / zb; res[2] = ab2 / zb;
    }
    uint32_t seed = (uint32_t)x * 709u + (uint32_t)y * 1153u + a.frame * 1361u;
    const float noise = rt::next_float(seed) * 0.005f;
    a.out[i * 3] = res[0] + noise;
    a.out[i * 3 + 1] = res[1] + noise;
    a.out[i * 3 + 2] = res[2] + noise;
}

// Hit pixels of the block's 16x16 tile, compacted to the first threads (block order kept), so
// that waves run the tap loop with full lanes; other pixels (sky, outside the image) are passed
// through by their own thread. Returns the number of hit pixels; s_list[i] = tile index.
__device__ __forceinline__ uint32_t compact_hits(const DenoiseArgs& a, uint32_t* s_list, uint32_t* s_wave) {
    const int x = blockIdx.x * kTile + (threadIdx.x & 15), y = blockIdx.y * kTile + (threadIdx.x >> 4);
    const bool inb = x < a.w && y < a.h;
    const bool hit = inb && a.depth[(size_t)y * a.w + x] > 0.0f;
    const uint64_t m = __ballot(hit);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) s_wave[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t off = 0, total = 0;
    for (int k = 0; k < 4; ++k) {
        off += k < wave ? s_wave[k] : 0u;
        total += s_wave[k];
    }
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    if (hit) s_list[off + (uint32_t)__popcll(m & below)] = threadIdx.x;
    return total;
}

// Radius > kMaxLdsRadius: texels from global memory (L2-resident for the tile's neighbourhood).
__global__ __launch_bounds__(256) void k_denoise_global(DenoiseArgs a) {
    __shared__ uint32_t s_list[256];
    __shared__ uint32_t s_wave[4];
    const uint32_t nhit = compact_hits(a, s_list, s_wave);
    __syncthreads();
    const DenoiseArgs* pa = &a;
    auto make_src = [pa](int col) { return GlobalSrc{pa, min(max(col, 0), pa->w - 1)}; };
    const int x = blockIdx.x * kTile + (threadIdx.x & 15), y = blockIdx.y * kTile + (threadIdx.x >> 4);
    if (x < a.w && y < a.h && !(a.depth[(size_t)y * a.w + x] > 0.0f)) denoise_pixel<false>(a, nullptr, 0, make_src, x, y);
    if (threadIdx.x < nhit) {
        const uint32_t t = s_list[threadIdx.x];
        denoise_pixel<false>(a, nullptr, 0, make_src, blockIdx.x * kTile + (t & 15), blockIdx.y * kTile + (t >> 4));
    }
}

// Radius <= kMaxLdsRadius: the block's (16 + 2r + 1) x (16 + 2r + 2) neighbourhood of clamped
// texels (tile width t = 16 + 2r + 2) as two float4 planes in LDS, plus the blur table.
__global__ __launch_bounds__(256) void k_denoise_lds(DenoiseArgs a, int r) {
    extern __shared__ float4 s_tile[];
    __shared__ uint32_t s_list[256];
    __shared__ uint32_t s_wave[4];
    const int t = kTile + 2 * r + 2;
    float4* s_cd = s_tile;
    float4* s_nm = s_tile + t * t;
    float* s_blur = reinterpret_cast<float*>(s_tile + 2 * t * t);
    const int ox = blockIdx.x * kTile - r, oy = blockIdx.y * kTile - r;
    const uint32_t nhit = compact_hits(a, s_list, s_wave);
    if (nhit > 0) {  // uniform: the whole block skips the staging when no pixel needs filtering
        for (int k = threadIdx.x; k < t * t; k += 256) {
            const int gx = min(max(ox + k % t, 0), a.w - 1), gy = min(max(oy + k / t, 0), a.h - 1);
            const size_t i = (size_t)gy * a.w + gx;
            s_cd[k] = make_float4(a.rgb[i * 3], a.rgb[i * 3 + 1], a.rgb[i * 3 + 2], a.depth[i]);
            s_nm[k] = make_float4(a.normal[i * 4], a.normal[i * 4 + 1], a.normal[i * 4 + 2], 0.0f);
        }
        fill_blur_table(a, r, s_blur);
    }
    __syncthreads();
    const float4* cd = s_cd;
    const float4* nm = s_nm;
    auto make_src = [cd, nm, t, ox, oy](int col) { return LdsSrc{cd, nm, t, col - ox, oy}; };
    const int x = blockIdx.x * kTile + (threadIdx.x & 15), y = blockIdx.y * kTile + (threadIdx.x >> 4);
    if (x < a.w && y < a.h && !(a.depth[(size_t)y * a.w + x] > 0.0f)) denoise_pixel<true>(a, s_blur, r, make_src, x, y);
    if (threadIdx.x < nhit) {
        const uint32_t q = s_list[threadIdx.x];
        denoise_pixel<true>(a, s_blur, r, make_src, blockIdx.x * kTile + (q & 15), blockIdx.y * kTile + (q >> 4));
    }
}

// ---- tonemappers ---------------------------------------------------------------------------
__device__ __forceinline__ float srgb(float c) {
    return c < 0.0031308f ? 12.92f * c : 1.055f * rt::tx_pow(c, 1.0f / 2.4f) - 0.055f;
}
__device__ __forceinline__ float hable_op(float x) {
    const float A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f;
    return ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F;
}
__device__ __forceinline__ float rrt_odt(float v) {
    const float a = v * (v + 0.0245786f) - 0.000090537f;
    const float b = v * (0.983729f * v + 0.4329510f) + 0.238081f;
    return a / b;
}
__device__ __forceinline__ float col_tone(float x, float con, float sh, float b, float c) {
    const float z = rt::tx_pow(x, con);
    return z / (rt::tx_pow(z, sh) * b + c);
}

// ColToneB / ColToneC of amd.comp for (hdrMax, contrast, shoulder, midIn, midOut); evaluated on
// the host with the same basic-operation pow as the device.
__host__ __device__ void amd_curve(const float amd[5], float* b, float* c) {
    const float hdr = amd[0], con = amd[1], sh = amd[2], mi = amd[3], mo = amd[4];
    const float pmc = rt::tx_pow(mi, con), phcs = rt::tx_pow(hdr, con * sh), phc = rt::tx_pow(hdr, con);
    const float pmcs = rt::tx_pow(mi, con * sh);
    const float num = -pmc + (mo * (phcs * pmc - phc * pmcs * mo)) / (phcs * mo - pmcs * mo);
    *b = -(num / (pmcs * mo));
    *c = (phcs * pmc - phc * pmcs * mo) / (phcs * mo - pmcs * mo);
}

// amd_b, amd_c: the AMD curve's ColToneB / ColToneC of p.amd (amd_curve), constant per launch.
__global__ __launch_bounds__(256) void k_tonemap(mrt_tonemap_params p, float amd_b, float amd_c, const float* __restrict__ rgb,
                                                 uint32_t n, float* __restrict__ out, uint8_t* __restrict__ rgba8) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    float v[3] = {rgb[3 * i] * p.exposure, rgb[3 * i + 1] * p.exposure, rgb[3 * i + 2] * p.exposure};
    float m[3];
    if (p.op == 0) {  // linear.comp
        for (int k = 0; k < 3; ++k) m[k] = v[k];
    } else if (p.op == 1) {  // reinhard.comp
        const float mw = p.reinhard_max * p.reinhard_max;
        for (int k = 0; k < 3; ++k) m[k] = (v[k] * (1.0f + (v[k] / mw))) / (1.0f + v[k]);
    } else if (p.op == 2) {  // hable.comp
        const float wd = hable_op(11.2f);
        for (int k = 0; k < 3; ++k) m[k] = hable_op(2.0f * v[k]) / wd;
    } else if (p.op == 3) {  // aces.comp
        const float i0 = (0.59719f * v[0] + 0.35458f * v[1]) + 0.04823f * v[2];
        const float i1 = (0.07600f * v[0] + 0.90834f * v[1]) + 0.01566f * v[2];
        const float i2 = (0.02840f * v[0] + 0.13383f * v[1]) + 0.83777f * v[2];
        const float f0 = rrt_odt(i0), f1 = rrt_odt(i1), f2 = rrt_odt(i2);
        m[0] = clampf((1.60475f * f0 + -0.53108f * f1) + -0.07367f * f2, 0.0f, 1.0f);
        m[1] = clampf((-0.10208f * f0 + 1.10813f * f1) + -0.00605f * f2, 0.0f, 1.0f);
        m[2] = clampf((-0.00327f * f0 + -0.07276f * f1) + 1.07602f * f2, 0.0f, 1.0f);
    } else if (p.op == 4) {  // uchimura.comp
        const float P = p.uchimura[0], a = p.uchimura[1], mm = p.uchimura[2], l = p.uchimura[3];
        const float c = p.uchimura[4], b = p.uchimura[5];
        const float l0 = ((P - mm) * l) / a;
        const float S0 = mm + l0, S1 = mm + a * l0;
        const float C2 = (a * P) / (P - S1);
        const float CP = -C2 / P;
        for (int k = 0; k < 3; ++k) {
            const float x = v[k];
            const float t = clampf(x / mm, 0.0f, 1.0f);
            const float w0 = 1.0f - t * t * (3.0f - 2.0f * t);
            const float w2 = x < mm + l0 ? 0.0f : 1.0f;
            const float w1 = 1.0f - w0 - w2;
            const float T = mm * rt::tx_pow(x / mm, c) + b;
            const float S = P - (P - S1) * rt::tx_exp(CP * (x - S0));
            const float L = mm + a * (x - mm);
            m[k] = (T * w0 + L * w1) + S * w2;
        }
    } else {  // amd.comp
        const float con = p.amd[1], sh = p.amd[2], b = amd_b, c = amd_c;
        float peak = fmaxf(v[0], fmaxf(v[1], v[2]));
        peak = fmaxf(1e-6f, peak);
        const float ratio[3] = {v[0] / peak, v[1] / peak, v[2] / peak};
        peak = col_tone(peak, con, sh, b, c);
        const float cross = con * 16.0f;
        const float tw = rt::tx_pow(peak, 4.0f);
        for (int k = 0; k < 3; ++k) {
            float r = rt::tx_pow(fabsf(ratio[k]), con / cross);
            r = r * (1.0f - tw) + 1.0f * tw;
            r = rt::tx_pow(fabsf(r), cross);
            m[k] = peak * r;
        }
    }
    float o[3];
    for (int k = 0; k < 3; ++k) o[k] = srgb(m[k]);
    if (out) { out[3 * i] = o[0]; out[3 * i + 1] = o[1]; out[3 * i + 2] = o[2]; }
    if (rgba8) {
        uchar4 q;
        q.x = (uint8_t)floorf(clampf(o[0], 0.0f, 1.0f) * 255.0f + 0.5f);
        q.y = (uint8_t)floorf(clampf(o[1], 0.0f, 1.0f) * 255.0f + 0.5f);
        q.z = (uint8_t)floorf(clampf(o[2], 0.0f, 1.0f) * 255.0f + 0.5f);
        q.w = 255;
        reinterpret_cast<uchar4*>(rgba8)[i] = q;
    }
}

}  // namespace post

extern "C" {

mrt_status mrt_denoise_bilateral(mrt_ctx* ctx, const float* rgb, const float* depth, const float* normal, uint32_t w,
                                 uint32_t h, const mrt_bilateral_params* p, float* out) {
    if (!ctx || !p) return api_fail(ctx, MRT_ERR_INVALID, "mrt_denoise_bilateral: null argument");
    if (w == 0 || h == 0) return MRT_OK;
    if (!rgb || !depth || !normal || !out) return api_fail(ctx, MRT_ERR_INVALID, "mrt_denoise_bilateral: null buffer");
    if (out == rgb) return api_fail(ctx, MRT_ERR_INVALID, "mrt_denoise_bilateral: in-place filtering is not supported");
    if (!(p->sigma > 0.0f) || !(p->k_sigma > 0.0f) || !(p->threshold > 0.0f) || !(p->near_plane > 0.0f))
        return api_fail(ctx, MRT_ERR_INVALID, "mrt_denoise_bilateral: sigma, k_sigma, threshold, near_plane must be > 0");
    post::DenoiseArgs a;
    a.rgb = rgb; a.depth = depth; a.normal = normal; a.out = out;
    a.w = (int)w; a.h = (int)h;
    a.radius = roundf(p->k_sigma * p->sigma);
    if (a.radius > 64.0f) return api_fail(ctx, MRT_ERR_INVALID, "mrt_denoise_bilateral: radius round(k_sigma*sigma) > 64");
    a.radq = a.radius * a.radius;
    a.inv_s2 = 0.5f / (p->sigma * p->sigma);
    a.inv_s2pi = 0.31830988618379067154f * a.inv_s2;
    a.inv_t2 = 0.5f / (p->threshold * p->threshold);
    a.inv_t_sqrt2pi = 0.39894228040143267794f / p->threshold;
    a.near_plane = p->near_plane;
    a.frame = p->frame;
    const dim3 grid((w + post::kTile - 1) / post::kTile, (h + post::kTile - 1) / post::kTile);
    const int r = (int)a.radius;
    {
        ProfScope pr(ctx, "denoise_bilateral");
        if (r <= post::kMaxLdsRadius) {
            const int t = post::kTile + 2 * r + 2;
            const size_t lds = (size_t)2 * t * t * sizeof(float4) + (size_t)(2 * r + 1) * (2 * r + 1) * sizeof(float);
            hipLaunchKernelGGL(post::k_denoise_lds, grid, dim3(256), lds, ctx->stream, a, r);
        } else {
            hipLaunchKernelGGL(post::k_denoise_global, grid, dim3(256), 0, ctx->stream, a);
        }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return api_fail(ctx, MRT_ERR_HIP, "mrt_denoise_bilateral: %s", hipGetErrorString(e));
    return MRT_OK;
}

mrt_status mrt_tonemap_defaults(mrt_tonemap_params* p) {
    if (!p) return api_fail(nullptr, MRT_ERR_INVALID, "mrt_tonemap_defaults: null argument");
    p->op = MRT_TONEMAP_AMD;
    p->exposure = 1.0f;
    p->reinhard_max = 8.0f;
    const float u[6] = {1.0f, 1.0f, 0.22f, 0.4f, 1.33f, 0.0f};
    const float a[5] = {16.0f, 2.0f, 1.0f, 0.18f, 0.18f};
    for (int k = 0; k < 6; ++k) p->uchimura[k] = u[k];
    for (int k = 0; k < 5; ++k) p->amd[k] = a[k];
    return MRT_OK;
}

mrt_status mrt_tonemap(mrt_ctx* ctx, const float* rgb, uint32_t n, const mrt_tonemap_params* p, float* out_rgb,
                       uint8_t* out_rgba8) {
    if (!ctx || !p) return api_fail(ctx, MRT_ERR_INVALID, "mrt_tonemap: null argument");
    if (p->op > MRT_TONEMAP_AMD) return api_fail(ctx, MRT_ERR_INVALID, "mrt_tonemap: unknown operator %u", p->op);
    if (n == 0) return MRT_OK;
    if (!rgb || (!out_rgb && !out_rgba8)) return api_fail(ctx, MRT_ERR_INVALID, "mrt_tonemap: null buffer");
    {
        ProfScope pr(ctx, "tonemap");
        float b = 0.0f, c = 0.0f;
        if (p->op == MRT_TONEMAP_AMD) post::amd_curve(p->amd, &b, &c);
        hipLaunchKernelGGL(post::k_tonemap, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, *p, b, c, rgb, n, out_rgb,
                           out_rgba8);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return api_fail(ctx, MRT_ERR_HIP, "mrt_tonemap: %s", hipGetErrorString(e));
    return MRT_OK;
}

}  // extern "C"
